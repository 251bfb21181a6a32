"""NumPy restatement of the jax.random pieces mocat calls (test infrastructure only).

jax/jaxlib are third-party and not part of the reference tree; mocat pins no version
(setup.py:15-19). Its era (jax 0.2.x-0.4.x, see mocat/src/core.py:17 `jax.util.unzip2` and
mocat/src/transport/svgd.py:11 `jax.example_libraries`) uses the non-partitionable threefry
layout restated here:

* threefry2x32-20 (Salmon et al. 2011). Pinned by the published Random123 known-answer vectors
  (tests/test_oracle_rng.py).
* PRNGKey(s) = [0, s]; split(k, n) = threefry_2x32(k, iota(2n)).reshape(n, 2); bits(k, shape)
  = threefry_2x32(k, iota(size)) with the counters cut into halves (x0 = first half, x1 = second
  half, a zero appended when the size is odd). Pinned by mocat.pdf p.125-126: normal(PRNGKey(0)),
  split(PRNGKey(0), 3) and the two 4x2 arrays printed there.
* uniform: bitcast(bits >> 9 | 0x3F800000) - 1, then max(lo, f * (hi - lo) + lo). Pinned through
  the same printed values (-5 + 0.1 * normal(...) to all printed digits).
* normal: sqrt(2) * erfinv(u), u ~ U[nextafter(-1, 0), 1). `normal_f32` mirrors Giles' single
  precision erfinv in f32 arithmetic; `normal_exact` uses scipy's FP64 erfinv. The printed
  values agree with both to <= 5e-7 relative; the residual is the f32 `log` of the XLA build
  that printed them (its implied log1p input error is ~1e-6 absolute), which no test in the
  reference pins: exact f32 bits of normals beyond that tolerance are "parity unpinned".
* randint (int32) / choice with replacement and no p (choice(key, n, shape) = randint(key, shape, 0,
  n), used by mocat/src/metrics.py:103 for KSD minibatches): k1, k2 = split(key); hi, lo =
  bits(k1), bits(k2); span = maxval - minval (uint32; 1 if maxval <= minval); m = (2^16 mod
  span)^2 mod span; result = minval + ((hi mod span) m + lo mod span) mod span, uint32 arithmetic
  wrapping at 32 bits (jax's legacy randint). No vector in the reference pins it. For spans up to
  2^16 the combination cannot wrap and equals the 64-bit draw (hi 2^32 + lo) mod span, which the
  tests check with Python integers; spans above 2^16 (where the product wraps) are "parity
  unpinned" beyond the range check.
* categorical (Gumbel-max, used by mocat/src/ssm/filtering.py:196-199 and
  mocat/src/transport/smc.py:61-71): argmax_j(logits_j - log(-log(U_ij))), U ~ uniform(key,
  (n, N), tiny, 1). Pinned only by brute-force agreement on tiny inputs; exact indices depend on
  f32 `log` rounding near ties, so bit-exactness is "parity unpinned".
"""
import numpy as np

U32 = np.uint32
F32 = np.float32

ROT = ((13, 15, 26, 6), (17, 29, 16, 24))


def _rotl(x, r):
    return ((x << U32(r)) | (x >> U32(32 - r))).astype(U32)


def threefry2x32(k0, k1, x0, x1):
    """Vectorised threefry2x32-20; keys and counters broadcast against each other."""
    k0 = np.asarray(k0, dtype=U32)
    k1 = np.asarray(k1, dtype=U32)
    shape = np.broadcast_shapes(k0.shape, k1.shape, np.shape(x0), np.shape(x1))
    x0 = np.array(np.broadcast_to(np.asarray(x0, U32), shape), dtype=U32, copy=True)
    x1 = np.array(np.broadcast_to(np.asarray(x1, U32), shape), dtype=U32, copy=True)
    ks = (k0, k1, (k0 ^ k1 ^ U32(0x1BD11BDA)).astype(U32))
    with np.errstate(over="ignore"):
        x0 += ks[0]
        x1 += ks[1]
        for i in range(5):
            for r in ROT[i % 2]:
                x0 += x1
                x1 = _rotl(x1, r)
                x1 ^= x0
            x0 += ks[(i + 1) % 3]
            x1 += ks[(i + 2) % 3] + U32(i + 1)
    return x0, x1


def threefry_2x32(key, count):
    """jax's threefry_2x32(keypair, count): halves layout over the flattened counter array."""
    key = np.asarray(key, dtype=U32)
    c = np.asarray(count, dtype=U32).ravel()
    odd = c.size % 2
    if odd:
        c = np.concatenate([c, np.zeros(1, U32)])
    h = c.size // 2
    a, b = threefry2x32(key[0], key[1], c[:h], c[h:])
    out = np.concatenate([a, b])
    if odd:
        out = out[:-1]
    return out.reshape(np.shape(count))


def PRNGKey(seed):
    seed = int(seed)
    assert -(2 ** 31) <= seed < 2 ** 31
    return np.array([0, seed & 0xFFFFFFFF], dtype=U32)


def split(key, num=2):
    return threefry_2x32(key, np.arange(2 * num, dtype=U32)).reshape(num, 2)


def bits(key, shape=()):
    shape = (shape,) if isinstance(shape, int) else tuple(shape)
    size = int(np.prod(shape)) if shape else 1
    if size >= 0xFFFFFFFF:
        raise ValueError("requesting more random bits than a single call provides")
    return threefry_2x32(key, np.arange(size, dtype=U32)).reshape(shape)


def uniform_from_bits(b, minval=0.0, maxval=1.0):
    lo = F32(minval)
    hi = F32(maxval)
    f = ((np.asarray(b, U32) >> U32(9)) | U32(0x3F800000)).view(F32) - F32(1)
    return np.maximum(lo, (f * (hi - lo)).astype(F32) + lo).astype(F32)


def uniform(key, shape=(), minval=0.0, maxval=1.0):
    return uniform_from_bits(bits(key, shape), minval, maxval)


def randint(key, shape=(), minval=0, maxval=1):
    """jax.random.randint (int32, legacy algorithm; see the module notes)."""
    k1, k2 = split(key)
    hi = bits(k1, shape).astype(np.uint64)
    lo = bits(k2, shape).astype(np.uint64)
    minval, maxval = int(minval), int(maxval)
    span = 1 if maxval <= minval else (maxval - minval) & 0xFFFFFFFF
    m = (65536 % span)
    m = ((m * m) & 0xFFFFFFFF) % span
    off = (((hi % span) * m) & 0xFFFFFFFF) + (lo % span)
    off = (off & 0xFFFFFFFF) % span
    return ((minval + off.astype(np.int64)) & 0xFFFFFFFF).astype(np.uint32).view(np.int32)


def choice(key, n, shape=()):
    """jax.random.choice(key, n, shape) with replacement and no p: randint(key, shape, 0, n)."""
    return randint(key, shape, 0, n)


NORMAL_LO = np.nextafter(F32(-1), F32(0))

# Giles (2011), single precision: central (w < 5) and tail coefficients, highest degree first.
_GILES_C = (2.81022636e-08, 3.43273939e-07, -3.5233877e-06, -4.39150654e-06, 0.00021858087,
            -0.00125372503, -0.00417768164, 0.246640727, 1.50140941)
_GILES_T = (-0.000200214257, 0.000100950558, 0.00134934322, -0.00367342844, 0.00573950773,
            -0.0076224613, 0.00943887047, 1.00167406, 2.83297682)


def erfinv_f32(x):
    x = np.asarray(x, F32)
    with np.errstate(divide="ignore", invalid="ignore"):
        w = (-np.log1p(-(x * x))).astype(F32)
        lt = w < F32(5)
        ww = np.where(lt, w - F32(2.5), np.sqrt(w).astype(F32) - F32(3)).astype(F32)
        p = np.where(lt, F32(_GILES_C[0]), F32(_GILES_T[0])).astype(F32)
        for c, t in zip(_GILES_C[1:], _GILES_T[1:]):
            p = (np.where(lt, F32(c), F32(t)) + p * ww).astype(F32)
        out = (p * x).astype(F32)
        return np.where(np.abs(x) == F32(1), x * F32(np.inf), out).astype(F32)


def normal_f32(key, shape=()):
    u = uniform(key, shape, NORMAL_LO, 1.0)
    return (F32(np.sqrt(2)) * erfinv_f32(u)).astype(F32)


def normal_exact(key, shape=()):
    from scipy.special import erfinv
    u = uniform(key, shape, NORMAL_LO, 1.0)
    return np.sqrt(2.0) * erfinv(u.astype(np.float64))


normal = normal_f32


def gumbel(key, shape):
    tiny = np.finfo(F32).tiny
    u = uniform(key, shape, tiny, 1.0)
    return -np.log(-np.log(u.astype(np.float64)))


def categorical(key, logits, n):
    """random.categorical(key, logits, shape=(n,)) by brute force (Gumbel-max)."""
    return categorical_margin(key, logits, n)[0]


def argmax_margin(scores):
    """Row-wise (argmax, top1 - top2) of an (m, n) FP64 score array: a row whose margin is within the
    f32 rounding of the scores can have a different first maximum in single precision."""
    scores = np.asarray(scores, np.float64)
    idx = np.argmax(scores, axis=-1)
    if scores.shape[-1] < 2:
        return idx.astype(np.int32), np.full(scores.shape[:-1], np.inf)
    top2 = np.partition(scores, scores.shape[-1] - 2, axis=-1)[..., -2:]
    return idx.astype(np.int32), top2[..., 1] - top2[..., 0]


def categorical_margin(key, logits, n):
    """categorical plus each draw's margin between the best and second-best Gumbel score (FP64)."""
    logits = np.asarray(logits, np.float64)
    g = gumbel(key, (n, logits.shape[-1]))
    return argmax_margin(g + logits[None, :])


def bits_batch(keys, size):
    """bits(keys[i], (size,)) for a batch of keys (C, 2) -> (C, size), vectorised over keys."""
    keys = np.asarray(keys, U32).reshape(-1, 2)
    S = int(size)
    h = (S + 1) // 2
    i = np.arange(h, dtype=U32)
    c1 = np.where(i + h < S, i + h, 0).astype(U32)
    a, b = threefry2x32(keys[:, 0:1], keys[:, 1:2], i[None, :], c1[None, :])
    out = np.concatenate([a, b], axis=1)[:, :S]
    return out


def split_batch(keys, num=2):
    return bits_batch(keys, 2 * num).reshape(-1, num, 2)


def uniform_batch(keys, size, minval=0.0, maxval=1.0):
    return uniform_from_bits(bits_batch(keys, size), minval, maxval)


def normal_batch(keys, size):
    u = uniform_batch(keys, size, NORMAL_LO, 1.0)
    return (F32(np.sqrt(2)) * erfinv_f32(u)).astype(F32)
