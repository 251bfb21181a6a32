"""Pins the RNG oracle: Random123 known answers and the values printed in mocat.pdf p.125-126."""
import numpy as np

from oracle import rng


def test_threefry_random123_kat():
    # Random123 kat_vectors for threefry2x32_20 (also jax's own threefry test vectors)
    cases = [
        ((0, 0), (0, 0), (0x6B200159, 0x99BA4EFE)),
        ((0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFF, 0xFFFFFFFF), (0x1CB996FC, 0xBB002BE7)),
        ((0x13198A2E, 0x03707344), (0x243F6A88, 0x85A308D3), (0xC4923A9C, 0x483DF7A0)),
    ]
    for key, ctr, want in cases:
        a, b = rng.threefry2x32(key[0], key[1], np.array([ctr[0]]), np.array([ctr[1]]))
        assert (int(a[0]), int(b[0])) == want


def test_prngkey():
    assert rng.PRNGKey(0).tolist() == [0, 0]
    assert rng.PRNGKey(7).tolist() == [0, 7]


# mocat.pdf p.125: normal(PRNGKey(0)); split(PRNGKey(0), 3) -> normal(s1), normal(s2)
PDF_SCALARS = (-0.20584235, 0.5781487, 0.85355157)
# mocat.pdf p.126: x = -5 + 0.1 * normal(s1, (4, 2)); momenta = normal(s2, (4, 2))
PDF_X = np.array([[-5.1619368, -4.871614], [-5.1136875, -5.048856],
                  [-5.0195227, -4.8458843], [-5.037027, -5.017855]], np.float32)
PDF_P = np.array([[-0.38537452, -1.4707391], [0.5467919, 2.095505],
                  [1.1165614, 0.16117463], [-0.5371375, -0.89213836]], np.float32)


def test_pdf_golden_values():
    k = rng.PRNGKey(0)
    _, s1, s2 = rng.split(k, 3)
    got = [float(rng.normal(k)), float(rng.normal(s1)), float(rng.normal(s2))]
    np.testing.assert_allclose(got, PDF_SCALARS, rtol=5e-7)
    x = np.float32(-5) + np.float32(0.1) * rng.normal(s1, (4, 2))
    # printed to 8 significant digits at magnitude 5 -> exact to the f32 grid there
    np.testing.assert_array_equal(x, PDF_X)
    np.testing.assert_allclose(rng.normal(s2, (4, 2)), PDF_P, rtol=5e-7)
    np.testing.assert_allclose(rng.normal_exact(s2, (4, 2)), PDF_P, rtol=5e-7)


def test_split_layout_is_halves():
    # split(k, 2) uses counters (0, 2) and (1, 3): a plain interleaved layout would differ
    k = rng.PRNGKey(0)
    s = rng.split(k, 2)
    a0, a1 = rng.threefry2x32(0, 0, np.array([0, 1]), np.array([2, 3]))
    assert s.tolist() == [[int(a0[0]), int(a0[1])], [int(a1[0]), int(a1[1])]]


def test_odd_size_padding():
    k = rng.PRNGKey(3)
    b = rng.bits(k, 5)
    # halves of [0,1,2,3,4,0]: pairs (0,3), (1,4), (2,0)
    a, c = rng.threefry2x32(0, 3, np.array([0, 1, 2]), np.array([3, 4, 0]))
    assert b.tolist() == [int(v) for v in np.concatenate([a, c[:2]])]


def test_uniform_range_and_erfinv_against_scipy():
    from scipy.special import erfinv
    k = rng.PRNGKey(11)
    u = rng.uniform(k, (100000,))
    assert u.min() >= 0 and u.max() < 1
    x = rng.uniform(k, (100000,), rng.NORMAL_LO, 1.0)
    ref = erfinv(x.astype(np.float64))
    got = rng.erfinv_f32(x).astype(np.float64)
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-30)
    # Giles reports ~1e-6 relative in the central branch; the tail branch (|x| -> 1) is looser
    assert np.median(rel) < 2e-7 and rel.max() < 1e-5


def test_categorical_matches_bruteforce_argmax():
    k = rng.PRNGKey(5)
    logits = np.log(np.array([0.1, 0.2, 0.3, 0.4]))
    idx = rng.categorical(k, logits, 2000)
    freq = np.bincount(idx, minlength=4) / 2000
    np.testing.assert_allclose(freq, [0.1, 0.2, 0.3, 0.4], atol=0.04)


def test_batched_helpers_match_scalar_path():
    keys = rng.split(rng.PRNGKey(4), 17)
    for size in (1, 2, 5, 50):
        want = np.stack([rng.bits(k, size) for k in keys])
        np.testing.assert_array_equal(rng.bits_batch(keys, size), want)
    np.testing.assert_array_equal(rng.split_batch(keys, 3), np.stack([rng.split(k, 3) for k in keys]))
    np.testing.assert_array_equal(rng.normal_batch(keys, 7), np.stack([rng.normal(k, (7,)) for k in keys]))


def test_randint_special_cases():
    """randint's combination (see oracle/rng.py) reduces to the second key's bits modulo the span when
    the span is a power of two up to 2^16 (2^32 is then a multiple of the span), and to minval when
    the span is 1 or maxval <= minval; otherwise it stays in [minval, maxval)."""
    k = rng.PRNGKey(7)
    lo_bits = rng.bits(rng.split(k)[1], (1000,))
    for span in (1, 2, 64, 4096, 65536):
        np.testing.assert_array_equal(rng.randint(k, (1000,), 3, 3 + span), 3 + (lo_bits % span).astype(np.int64))
    np.testing.assert_array_equal(rng.randint(k, (5,), 9, 2), np.full(5, 9))
    for lo, hi in ((0, 3), (-7, 1000003), (5, 4096 + 5), (0, 2 ** 31 - 1)):
        r = rng.randint(k, (20000,), lo, hi).astype(np.int64)
        assert r.min() >= lo and r.max() < hi
        assert len(np.unique(r)) > min(hi - lo, 20000) // 4
    np.testing.assert_array_equal(rng.choice(k, 300, (4, 5)), rng.randint(k, (4, 5), 0, 300))


def test_randint_is_the_64_bit_draw_modulo_the_span():
    """For spans up to 2^16 the legacy combination (hi mod s) * ((2^16 mod s)^2 mod s) + (lo mod s),
    taken mod s, cannot wrap in uint32, and (2^16)^2 = 2^32, so it equals the 64-bit integer
    (hi * 2^32 + lo) mod s, with hi and lo the two split keys' bits (computed here with Python
    integers, none of the formula's steps)."""
    k = rng.PRNGKey(11)
    k1, k2 = rng.split(k)
    hi = rng.bits(k1, (2000,)).astype(object)
    lo = rng.bits(k2, (2000,)).astype(object)
    for span in (3, 7, 10, 1000, 4097, 65535, 65536):
        want = np.array([(int(h) * (1 << 32) + int(l)) % span for h, l in zip(hi, lo)], np.int64)
        got = rng.randint(k, (2000,), -5, -5 + span).astype(np.int64) + 5
        np.testing.assert_array_equal(got, want)


def test_argmax_margin_brute_force():
    rs = np.random.default_rng(3)
    s = rs.standard_normal((50, 7))
    s[3, 2] = s[3, 5] = 9.0  # an exact tie: margin 0, first maximum
    idx, mg = rng.argmax_margin(s)
    for r in range(50):
        order = np.sort(s[r])
        assert idx[r] == int(np.argmax(s[r])) and mg[r] == order[-1] - order[-2]
    assert mg[3] == 0.0 and idx[3] == 2
