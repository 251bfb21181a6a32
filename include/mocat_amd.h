/* mocat_amd C ABI: the drop-in boundary between mocat-shaped host code and the gfx950 kernels.
 *
 * Conventions (every entry point):
 *   - pointers are device pointers unless named host_*; sizes are explicit int64_t;
 *   - RNG keys are uint32_t[2] pairs (jax.random key layout), batched as [n_keys][2];
 *   - particle/chain arrays are row-major [count][dim] f32 unless stated;
 *   - the last argument is the hipStream_t the work is queued on (work is asynchronous);
 *   - the return value is a status: 0 ok, -1 invalid argument, -2 HIP launch/runtime error,
 *     -3 size limit exceeded. No entry point allocates device memory unless stated.
 *
 * The reference (mocat 0.2.6) is pure Python/JAX and has no C interface; each declaration cites
 * the reference function (file:line) whose computation it replaces.
 */
#ifndef MOCAT_AMD_H
#define MOCAT_AMD_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Returns the ABI revision of the library (not a status). */
int mocat_abi_version(void);

/* ---------------------------------------------------------------------------------------- RNG
 * jax.random.split / _random_bits / uniform / normal as called throughout mocat, e.g.
 * mocat/src/ssm/filtering.py:181 (split(key, n)), mocat/src/mcmc/standard_mcmc.py:120-122
 * (split + normal), mocat/src/mcmc/metropolis.py:60-61 (split + uniform).
 * Batched over n_keys keys: out is [n_keys][...]. */

/* split(keys[i], num) -> out[i][num][2]. */
int mocat_random_split(const uint32_t* keys, int64_t n_keys, int64_t num, uint32_t* out, hipStream_t stream);
/* _random_bits(keys[i], 32, (size,)) -> out[i][size]. size < 2^32 - 1. */
int mocat_random_bits(const uint32_t* keys, int64_t n_keys, int64_t size, uint32_t* out, hipStream_t stream);
/* uniform(keys[i], (size,), minval, maxval) -> out[i][size] (f32). */
int mocat_random_uniform(const uint32_t* keys, int64_t n_keys, int64_t size, float minval, float maxval,
                         float* out, hipStream_t stream);
/* normal(keys[i], (size,)) -> out[i][size] (f32). */
int mocat_random_normal(const uint32_t* keys, int64_t n_keys, int64_t size, float* out, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* MOCAT_AMD_H */
