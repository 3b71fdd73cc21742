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

/* Text of the last runtime (status 2) failure on the calling thread, "" if
 * none: the failing call site and the HIP error string.  No reference
 * counterpart (the reference reports bare status codes). */
const char *fe_last_error(void);

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

/* fe_edt_device without the `work` argument, allocating nothing: its work
 * planes live in a static device array of 2^21 int32 words (8 MiB, both
 * planes of the reference EDT's 2^20-voxel cap) that each GPU keeps until the
 * process ends.  No counterpart in the reference; it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops).
 * Returns 1 when planes*nx*ny*nz exceeds 2^21.  Calls that overlap on one
 * GPU share the array and race. */
int fe_edt_device_static(const uint8_t *mask, int nx, int ny, int nz, int planes, int32_t *out, void *stream);

/* One axis pass of fe_edt_device (per-kernel timing and profiling), with the
 * intermediate layouts fe_edt_device uses: axis 2 reads the uint8 mask and
 * writes the z pass (z summaries when nz >= 16, else int32 planes), axis 1
 * reads that and runs the lower-envelope pass of
 * edt_native_linear_v1/edt.cu:13-29 into int32 planes, axis 0 runs it over
 * int32 planes.  The z-summary layout is internal to the library. */
int fe_edt_axis_pass(const void *src, int32_t *dst, int nx, int ny, int nz, int planes, int axis, void *stream);

/* Batched EDT of many uneven volumes (no reference counterpart; it replaces
 * one edt_pair_query / edt_squared_large call per volume,
 * edt_persistent_pair_v1/edt.cu:59-78, with one launch per axis pass over all
 * volumes).  dims: host int32 [nvol][3]; mask_off: byte offset of each
 * volume's mask in a device byte arena of mask_bytes; data_off/plane_stride:
 * element offset of each volume's plane 0 and the distance between its planes
 * in device int32 arenas of data_elems (out and work share the layout; plane
 * ranges must not overlap).  planes as in fe_edt_device.  create returns NULL
 * with *status 1 (invalid) or 2 (runtime); run launches three kernels on
 * `stream` and allocates nothing (capturable in a graph). */
void *fe_edt_batch_create(const int32_t *dims, const int64_t *mask_off, const int64_t *data_off,
                          const int64_t *plane_stride, int nvol, int planes, int64_t mask_bytes, int64_t data_elems,
                          int *status);
int fe_edt_batch_run(const void *handle, const uint8_t *masks, int32_t *work, int32_t *out, void *stream);
void fe_edt_batch_destroy(void *handle);

/* Signed output of edt_persistent_pair_v1/edt.cu:36-42 from a device pair. */
int fe_edt_signed_device(const int32_t *pair, int64_t n, double pitch, float *out, void *stream);

/* Both checksums of faltkarna_v1.py:238-255 (reduktion_checksum_single_process
 * :368-386) over n device float32 values: *out32 (device float, zeroed by the
 * caller) += every value < 0 by float atomics (order-dependent control),
 * *out64 (device int64, zeroed) += int64(f64(v) * scale) for every v < 0. */
int fe_checksum_pair(const float *vals, int64_t n, double scale, float *out32, int64_t *out64, void *stream);

/* ---------------------------------------------------------------- 2-bit adaptive block tree */
/* Replaces the Python builder and lookups of field_adaptive_blocks_v1.py:20-138
 * (build :97-138, sample :52-82).  dense: device float32 (nx,ny,nz); kinds[4]:
 * device int32 dense-source kinds at block sides 8, 4, 2, 1 (fe_dense_classify
 * with margin f32(sqrt(3)*side*pitch)).  build returns a handle or NULL with
 * *status 1 (invalid) / 2 (runtime).  info: sizes[4] = node count, value count,
 * packed bytes, checkpoints; counts[16] and first[16] per level (side 8,4,2,1)
 * and code.  copy: device destinations (any may be NULL).  sample: packed
 * codes zero-padded to 64-byte groups, device checkpoints, values,
 * root_shape[3] (host int32), in-domain int64 coordinates [n][3]; *flag (a
 * device int) is set to 1 if the tree is malformed. */
void *fe_octree_build(const float *dense, int nx, int ny, int nz, const int32_t *const *kinds, int *status,
                      void *stream);
int fe_octree_info(const void *handle, int64_t *sizes, int64_t *counts, int64_t *first);
int fe_octree_copy(const void *handle, uint8_t *packed, int32_t *branch_prefix, int32_t *value_prefix, float *values,
                   void *stream);
void fe_octree_free(void *handle);
int fe_octree_sample(const uint8_t *packed, const int32_t *branch_prefix, const int32_t *value_prefix,
                     const float *values, const int32_t *root_shape, int64_t root_count, double pitch,
                     const int64_t *ijk, int64_t n, float *out, int *flag, void *stream);

/* ---------------------------------------------------------------- field build */
/*
 * `descs` is a device table of R field descriptors (160 bytes each: grid
 * origin/pitch/margin, sizes, block edge B, superblock edge, primitive table
 * and output buffers; layout in the package's field.py FIELD_DESC).  Recipes
 * are arrays of 32-byte primitives {int32 kind, op; float cx,cy,cz,a,b,c}.
 */

/* Replaces the classify -> np.where -> eval launches of faltkarna_v1.py:258-299
 * and faltkarna_v1_g7_fix.py:56-112 for R recipes at once, with no host sync:
 * eight-corner classification (faltkarna_v1.py:128-164), ascending compaction
 * (:277), tile evaluation (:166-190).  sb_prefix/ch_prefix: device int32[R+1]
 * exclusive prefixes of superblock and 1024-block chunk counts; chunk_count:
 * device int32[total_chunks] scratch.  Every descriptor's block edge equals
 * B (4, 8 or 16).  tile_waves: waves of the persistent tile launch; the
 * active blocks of all R recipes, concatenated in order, are dealt to them
 * round-robin. */
int fe_field_build(const void *descs, int R, int B, const int32_t *sb_prefix, int total_sb, const int32_t *ch_prefix,
                   int total_chunks, int32_t *chunk_count, int tile_waves, int max_prims, void *stream);

/* fe_field_build that also counts the work of the culled evaluation (no
 * reference counterpart; for measurement): evals is a device uint64[6] added
 * to: [0] primitive evaluations at classified lattice points and [1] those
 * points, [2] primitive evaluations at tile voxels and [3] those voxels,
 * [4] / [5] interval evaluations of the culling in the classify / tile pass. */
int fe_field_build_counted(const void *descs, int R, int B, const int32_t *sb_prefix, int total_sb,
                           const int32_t *ch_prefix, int total_chunks, int32_t *chunk_count, int tile_waves,
                           int max_prims, unsigned long long *evals, void *stream);

/* k_classify_testfalt launch of faltkarna_v1.py:269-273 (kinds, corner_min_abs,
 * culled primitive lists) without compaction or tiles.  max_prims (both
 * calls): the longest recipe's primitive count (<= 4096), which sizes the
 * culling lists in LDS. */
int fe_field_classify(const void *descs, int R, int B, const int32_t *sb_prefix, int total_sb, int max_prims,
                      void *stream);

/* k_eval_tile_testfalt launch of faltkarna_v1.py:283-287 over host-compacted ids;
 * host_desc is a host copy of one descriptor whose tiles field names the output. */
int fe_field_tiles_ids(const void *host_desc, const int32_t *ids, int64_t n_active, void *stream);

/* k_fill_background + k_scatter_tiles of faltkarna_v1.py:192-236 (cropped),
 * plus the optional occupancy of faltkarna_v1_mesh_to_sdf.py:695-710. */
int fe_field_reconstruct(const void *descs, int r, int64_t n_voxels, float *dense, uint8_t *occupancy_or_null,
                         void *stream);

/* k_reduce_checksum_i64_safe of faltkarna_v1.py:249-255 (adds into *out). */
int fe_field_checksum(const void *descs, int r, double scale, int64_t *out, void *stream);

/* k_classify_dense / k_eval_tile_dense of faltkarna_v1_mesh_to_sdf.py:602-654. */
int fe_dense_classify(const float *dense, int nx, int ny, int nz, int B, float margin, int32_t *kind, void *stream);
int fe_dense_tiles(const float *dense, int nx, int ny, int nz, int B, const int32_t *ids, int64_t n_active,
                   float *tiles, void *stream);

/* Point SDF queries on a resident field: the recipe's fp32 fold at device
 * points (n,3), with the culled list of the containing superblock.  Replaces
 * the per-batch eval_batch launch of ikarus_v1/eval_warp.py:478-489 for a
 * built field.  From 2^18 points the points are first ordered by
 * superblock in stream-ordered scratch (hipMallocAsync / hipFreeAsync on
 * `stream`, 4 * (8194 + 3n) bytes); fewer points allocate nothing. */
int fe_query_points(const void *desc, const float *points, int64_t n, float *out, void *stream);
/* No counterpart in the reference (its builds are one per call): `nseg`
 * device-to-device copies in one launch, for a server moving every built
 * field of a batch out of the batch's pools.  segs: device table of 24-byte
 * records {const void *src; void *dst; int64 bytes}; a null src clears dst.
 * nseg <= 65535. */
int fe_copy_segments(const void *segs, int nseg, void *stream);

/* ---------------------------------------------------------------- per-block pitch */
/* The two-level field of faltkarna_v1_multires.py:84-106 (build :254-363,
 * band :211-252, queries :405-441), from a recipe instead of a mesh.  bricks:
 * device table of 32-byte records {int32 bi, bj, bk, level; int64 tile
 * (offset of its (B*2^level + 1)^3 fp32 samples), item0 (its first slab,
 * ascending from 0; a brick with m samples per axis has ceil(m / r) slabs of
 * r = max(1, 1024 / m^2) rows)}; grid: host 72-byte record {double lo[3], h,
 * margin; int32 B, nb[3], n[3], pad} (f64 origin and pitch, brick lattice,
 * coarse sample counts); slot / kind: device int32 per lattice brick (table
 * index or -1; -1 interior, 0 active, +1 exterior).  tiles: every active
 * brick's samples from one classified descriptor (its culled lists, refined
 * exactly per slab; nitems = all slabs), level 0
 * on the coarse lattice with indices clamped to the grid, level l at
 * x0 + f32(i) * (h / 2^l).  band: the face layers of the `fine` bricks (all
 * at one level) next to coarser active bricks, as
 * map_coordinates(order=1, mode="nearest") of those bricks.  query: signed
 * distance at device f64 points (n x 3) into device f64 out; brick_out
 * (optional, device int32 n x 4): the (clipped) brick and its level (-1
 * inactive) per point. */
int fe_multires_tiles(const void *desc, const void *bricks, int nbricks, int64_t nitems, float *tiles, void *stream);
int fe_multires_band(const void *grid, const void *bricks, const int32_t *slot, const int32_t *fine, int nfine,
                     float *tiles, void *stream);
int fe_multires_query(const void *grid, const void *bricks, const int32_t *slot, const int32_t *kind,
                      const float *tiles, const double *points, int64_t n, double *out, int32_t *brick_out,
                      void *stream);

/* ---------------------------------------------------------------- leaf-box packing */
/* Node records are 52 bytes {int32 lo[3], shape[3], stored[3], children[2];
 * int64 offset} in pre-order, as adaptive_sample_blocks_v1.py:14-16. */

/* Replace adaptive_pack / adaptive_nodes / adaptive_samples / adaptive_copy /
 * adaptive_free, adaptive_pack_native_v1/pack.cpp:57-79: pack a host fp32
 * (shape[0],shape[1],shape[2]) field (NULL on invalid or non-finite input);
 * the split decisions run on the device, the result is held on the host. */
void *adaptive_pack(const void *data, const int32_t *shape, int threshold);
int64_t adaptive_nodes(const void *packed);
int64_t adaptive_samples(const void *packed);
int adaptive_copy(const void *packed, void *nodes, int64_t count, void *payload, int64_t samples);
void adaptive_free(void *packed);

/* Device-resident packing of a device fp32 field (same split decisions as
 * adaptive_sample_blocks_v1.py:19-58); *status 1 on non-finite input, 2 on a
 * runtime failure.  fe_pack_copy copies into device or host buffers. */
void *fe_pack_device(const float *data, int nx, int ny, int nz, int leaf_samples, int *status, void *stream);
int64_t fe_pack_nodes(const void *packed);
int64_t fe_pack_samples(const void *packed);
int fe_pack_copy(const void *packed, void *nodes, int64_t count, void *payload, int64_t samples, void *stream);
void fe_pack_free(void *packed);

/* samples_at and query of adaptive_sample_blocks_v1.py:103-151 on device tables:
 * exact leaf lookups and the order-1 trilinear query (fp64, rounded to fp32). */
int fe_samples_at(const void *nodes, const uint32_t *payload, const int64_t *indices, int64_t n, uint32_t *out,
                  void *stream);
int fe_query_packed(const void *nodes, const uint32_t *payload, int nx, int ny, int nz, const double *origin,
                    double pitch, const double *points, int64_t n, float *out, void *stream);

/* Dense device reconstruction of a packed field (one workgroup per leaf). */
int fe_unpack(const void *nodes, int64_t count, const uint32_t *payload, int nx, int ny, int nz, float *dense,
              void *stream);

/* ---------------------------------------------------------------- tiled CSG */
/* Replaces csg_gather, csg_tile_native_v1/gather.cpp:11 (codes 0-5): exact
 * bits of every leaf intersecting [lower, upper) into a tile of <= 8192 samples. */
int csg_gather(const void *nodes, int64_t count, const void *payload, int64_t samples, const int32_t *lower,
               const int32_t *upper, void *output);

/* Replaces csg_execute, csg_batch_native_v1/execute.cpp:10 (codes 0-5):
 * ordered fp32 union/difference/offset program over packed operands; program
 * records are 24 bytes {int32 op, left, right, pad; double offset}. */
int csg_execute(const void *const *nodes, const int64_t *counts, const void *const *payloads, const int64_t *lengths,
                int32_t fields, const int32_t *shape, const void *program, int32_t instructions, float *output);

/* The same program over device-resident operands into a device dense output,
 * every used operand checked on each call; axes <= 4096 with no sample cap
 * (the reference caps stay on csg_execute).  Device buffers are kept between
 * calls (one set per device; concurrent calls are serialised). */
int fe_csg_device(const void *const *nodes, const int64_t *counts, const void *const *payloads,
                  const int64_t *lengths, int32_t fields, const int32_t *shape, const void *program,
                  int32_t instructions, float *output, void *stream);

/* Check-once form of fe_csg_device (no reference counterpart: the reference
 * re-gathers every operand tile by tile, csg_batch_native_v1/execute.cpp:25-30).
 * fe_csg_check validates one operand for a grid of `shape` with the
 * reference gather's rules (gather.cpp:28-47): *first_bad = the first
 * reference (16,32,16) tile, x-major, whose gather of this operand fails (0
 * for a malformed table), INT32_MAX when every sample is covered exactly once.
 * fe_csg_run then executes a program over checked operands (first_bad: their
 * host int32 results, per field) in a caller workspace of at least
 * fe_csg_workspace_bytes (-1: invalid program), allocating nothing; status
 * codes as csg_execute. */
int fe_csg_check(const void *nodes, int64_t count, const void *payload, int64_t samples, const int32_t *shape,
                 int32_t *first_bad, void *stream);
int64_t fe_csg_workspace_bytes(int32_t fields, const int32_t *shape, const void *program, int32_t instructions);
int fe_csg_run(const void *const *nodes, const int64_t *counts, const void *const *payloads, const int64_t *lengths,
               int32_t fields, const int32_t *shape, const void *program, int32_t instructions,
               const int32_t *first_bad, float *output, void *workspace, int64_t workspace_bytes, void *stream);

/* ---------------------------------------------------------------- mesh stage */
/* Replace the CUDA column scan of cuda_columns_v1/columns.cu:50-113 (status 0
 * ok, 1 invalid, 2 runtime, 3 foreign thread, 4 a query with more than `hits`
 * hits).  input: triangles x 9 doubles (corners a, b, c); hits per query in
 * face order, depths/signs [count][hits] (compact: [count][*width], *width =
 * largest count).  backend "cuda-scan-v1" or "hip-scan-v1". */
int columns_create(const char *backend, const double *input, uint32_t triangles, uint32_t capacity, uint32_t hits,
                   const double *anchor, void **handle);
int columns_query(void *handle, const double *xy, uint32_t count, int32_t *counts, double *depths, int32_t *signs);
int columns_query_compact(void *handle, const double *xy, uint32_t count, int32_t *counts, double *depths,
                          int32_t *signs, uint32_t *width);
int columns_destroy(void **handle);

/* Replaces column_mask_surface, column_mask_native_v1/mask.cu:17-37: mask and
 * 6-neighbour surface from sorted composite keys (CSR offsets per column). */
int column_mask_surface(const double *keys, const int32_t *signs, int hits, const int32_t *offsets,
                        const double *fractions, int nx, int ny, int nz, uint8_t *mask, uint8_t *surface);

/* Replace fused_create / fused_query / fused_destroy of
 * mesh_fused_native_v1/field.cu:39-94 and mesh_fused_graph_v1/field.cu:18-79:
 * column hits -> composite-key mask -> surface -> paired EDT -> signed
 * distance, captured once into a graph and replayed per query.  `path` (the
 * reference's ray-tracing program) is ignored: hits come from the column scan.
 * Query status 4 (hit overflow / non-finite depth) and 1 (mask lacks a class)
 * latch the handle. */
int fused_create(const char *path, const double *corners, uint32_t triangles, const double *anchor, const double *xy,
                 int nx, int ny, int nz, double pitch, const double *zs, void **handle);
int fused_query(void *handle, uint8_t *solid, uint8_t *surface, float *distance);
int fused_destroy(void **handle);

/* The same stage without the reference's size caps (axes <= 4096, < 2^31
 * voxels, 1..256 hits per column), and its open-mesh form: mask = |ordered
 * generalized winding number| > .5 at `points` (nx*ny*nz x 3), as
 * mesh_field_cuda_columns_v1.py:69-70 with generalized_winding_ordered_v1.py:11-32.
 * fe_mesh_query also returns 5 when a winding point lies on a mesh corner. */
int fe_mesh_create(const double *corners, uint32_t triangles, const double *anchor, const double *xy, int nx, int ny,
                   int nz, double pitch, const double *zs, int hit_capacity, void **handle);
int fe_gwn_create(const double *corners, uint32_t triangles, const double *points, int nx, int ny, int nz,
                  double pitch, void **handle);
int fe_mesh_query(void *handle, uint8_t *solid, uint8_t *surface, float *distance);
int fe_mesh_destroy(void **handle);

/* Device-resident use of a mesh handle: replay on a stream without copies,
 * the last run's status once the stream is done, and the device buffers. */
int fe_mesh_run(void *handle, void *stream);
int fe_mesh_status(const void *handle);
int fe_mesh_buffers(const void *handle, void **solid, void **surface, void **distance, void **pair, void **counts,
                    void **keys, void **signs, int *cap);

/* Winding numbers of device points against device corners (w, mask optional). */
int fe_gwn_device(const double *corners, int64_t triangles, const double *points, int64_t n, double *w, uint8_t *mask,
                  int *flags, void *stream);

/* ---------------------------------------------------------------- IKARUS expression trees */
/* Replaces the per-tree generated Warp kernels of ikarus_v1/eval_warp.py:160-485
 * (and eval_warp_uniform_v2.py:100-270 for rect_frustum) with one interpreter:
 * ins = n_ins 16-byte instructions {int32 op, arg, param offset, 0} compiled by
 * the package's ikarus.compile_tree; params fp32; grids = 48-byte grid-leaf
 * records {int64 offset; int32 nx, ny, nz, is_sdf; float ox, oy, oz, pitch,
 * iso, pad} over grid_data (null without grid leaves).  All pointers are
 * device pointers except origin_pitch_margin (5 floats) and blocks (3 int32),
 * which are host arrays.  eval: values at points (n x 3 fp32).  classify:
 * faltkarna_v1.py:128-164 rule per block (kind int32, corner_min_abs fp32);
 * tiles: faltkarna_v1.py:166-190 for the ascending active ids.
 * The entry points check pointers and n_ins only; the interpreter stops a
 * program that underflows or overflows its 17-value stack or holds an unknown
 * opcode (NaN values, no memory outside its own stack touched) and latches a
 * per-GPU fault word that fe_ikarus_fault reads.  Parameter offsets and grid
 * indices are trusted: run fe_ikarus_check on the host program first. */
int fe_ikarus_eval(const void *ins, int n_ins, const float *params, const void *grids, const float *grid_data,
                   const float *points, int64_t n, float *out, void *stream);
int fe_ikarus_classify(const void *ins, int n_ins, const float *params, const void *grids, const float *grid_data,
                       const float *origin_pitch_margin, const int32_t *blocks, int B, int32_t *kind, float *cma,
                       void *stream);
int fe_ikarus_tiles(const void *ins, int n_ins, const float *params, const void *grids, const float *grid_data,
                    const float *origin_pitch_margin, const int32_t *blocks, int B, const int32_t *ids,
                    int64_t n_active, float *tiles, void *stream);
/* No counterpart in the reference (its programs are generated Python): checks
 * a host copy of a program before upload -- known opcodes, parameters inside
 * [0, n_params), grid leaves below n_grids, axis and order arguments, and a
 * stack that never underflows, stays within 17 values and ends with one.
 * FE_INVALID names the first bad instruction in fe_last_error(). */
int fe_ikarus_check(const void *ins_host, int n_ins, int64_t n_params, int n_grids);
/* Waits for stream, then reads and clears this GPU's interpreter fault word:
 * FE_INVALID when a malformed program ran since the last call. */
int fe_ikarus_fault(void *stream);

/* The passes of the hierarchical volume queries of ikarus_v1/queries.py:138-277
 * on an implicit grid (lo: 3 host doubles, shape: 3 host int32, samples
 * lo + pitch*i per axis in f64, rounded to fp32).  grid: counts[0] += samples
 * with v < below, counts[1] += samples with |v| <= band (device uint64[2]);
 * the band samples' grid indices are appended to cells through *cursor
 * (device int32[n] and uint32, zeroed by the caller; both may be NULL).
 * refine: r^3 samples per listed cell at centre + ((a+0.5)/r*pitch - pitch/2);
 * counts[0] += v < 0, counts[1] += |v| <= band. */
int fe_ikarus_grid(const void *ins, int n_ins, const float *params, const void *grids, const float *grid_data,
                   const double *lo, double pitch, const int32_t *shape, float below, float band,
                   unsigned long long *counts, int32_t *cells, unsigned int *cursor, void *stream);
int fe_ikarus_refine(const void *ins, int n_ins, const float *params, const void *grids, const float *grid_data,
                     const double *lo, double pitch, const int32_t *shape, const int32_t *cells, int64_t ncells, int r,
                     float band, unsigned long long *counts, void *stream);

/* Grid-leaf pairs, replacing the fixed-shape multi-pair kernel of
 * ikarus_v1/eval_warp.py:491-560 (eval_batch_multi_pair): out[i] =
 * max(leaf leaf_a[pair_of_point[i]], leaf leaf_b[...]) at point i (selection
 * max), over nleaves grid records as above; out-of-range indices give NaN.
 * All pointers are device pointers. */
int fe_ikarus_pairs_eval(const void *grids, int nleaves, const float *grid_data, const float *points,
                         const int32_t *pair_of_point, const int32_t *leaf_a, const int32_t *leaf_b, int npairs,
                         int64_t n, float *out, void *stream);

/* Band tree of one grid leaf (grids[0]) with the plan as data, replacing
 * ikarus_v1/eval_warp.py:567-700 (compile_band_tree_kernel,
 * BandTreeEvaluator.eval): min over bands of max(leaf(p, p[axis] := zs),
 * max(p[axis] - z1, z0 - p[axis])), folded from 1e30; zs/z0/z1: nbands fp32
 * each (device). */
int fe_ikarus_band_eval(const void *grids, const float *grid_data, int axis, const float *zs, const float *z0,
                        const float *z1, int nbands, const float *points, int64_t n, float *out, void *stream);


#ifdef __cplusplus
}
#endif
#endif
