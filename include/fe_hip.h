/*
 * C ABI of libfe_hip.so, the gfx950 block-sparse SDF build path.
 *
 * Every entry point is `extern "C"`, noexcept, takes plain pointers and sizes
 * and returns a status: 0 success, 1 invalid input, 2 runtime failure (callers
 * latch 2 as "no retry"), as in the reference (FE/native_edt_v1.py:18-33).
 * Host pointers are caller-owned and never retained.  `stream` arguments are
 * hipStream_t values passed as void* (NULL = the null stream).  Symbols that
 * replace a reference export keep its name and signature; the comment above
 * each cites the reference interface (paths under src/field_engine/).
 */
#ifndef FE_HIP_H
#define FE_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- EDT */

/* Replaces edt_squared, edt_native_v1/edt.cu:18-20 (<=128 per axis, <=262144
 * voxels): squared distance of every nonzero voxel to the nearest zero voxel. */
int edt_squared(const uint8_t *mask, int nx, int ny, int nz, int32_t *output);

/* Replaces edt_squared_large, edt_native_large_v1/edt.cu:20 and
 * edt_native_linear_v1/edt.cu:33 (<=512 per axis, <=1048576 voxels). */
int edt_squared_large(const uint8_t *mask, int nx, int ny, int nz, int32_t *output);

/* Replaces edt_pair_create / edt_pair_query / edt_pair_destroy,
 * edt_persistent_pair_v1/edt.cu:44-78: persistent paired EDT handle; query
 * uploads the mask, runs both polarities and writes the signed field
 * (f32(sqrt(out)*h) - f32(sqrt(in)*h), half-pitch corrected) and optionally
 * the int32 pair (2,nx,ny,nz).  A status-2 failure latches the handle. */
void *edt_pair_create(int nx, int ny, int nz);
int edt_pair_query(void *handle, const uint8_t *mask, double pitch, float *output, int32_t *squared_or_null);
void edt_pair_destroy(void *handle);

/* Uncapped variants of the two calls above (axes <= 4096, no voxel cap);
 * same contracts as edt_native_linear_v1/edt.cu:33 and edt_persistent_pair_v1/edt.cu:48. */
int fe_edt_squared_host(const uint8_t *mask, int nx, int ny, int nz, int32_t *output);
void *fe_edt_pair_create(int nx, int ny, int nz);

/* Device-resident EDT (no reference counterpart; the reference's kernels
 * edt_native_linear_v1/edt.cu:13-29 and edt_persistent_pair_v1/edt.cu:12-35
 * run behind host copies).  mask: device uint8 (nx,ny,nz); work, out: device
 * int32 of planes*N; planes 1 = squared distance of nonzero voxels to the
 * nearest zero; planes 2 = plane 0 non-mask-to-mask, plane 1 mask-to-complement. */
int fe_edt_device(const uint8_t *mask, int nx, int ny, int nz, int planes, int32_t *work, int32_t *out,
                  void *stream);

/* Signed output of edt_persistent_pair_v1/edt.cu:36-42 from a device pair. */
int fe_edt_signed_device(const int32_t *pair, int64_t n, double pitch, float *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
