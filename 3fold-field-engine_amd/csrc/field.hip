// Block-sparse SDF build for gfx950: per-block eight-corner classification,
// ascending surface-block compaction, tile evaluation, reconstruction and the
// deterministic int64 checksum.
//
// Reference semantics (FE/faltkarna_v1.py): a block of B^3 voxels is
// classified from the recipe's fp32 values at its eight corners o + {0,B*h}
// (:128-164); mixed signs or min|v| <= margin (= fp32(sqrt(3)*B*h)) make it a
// surface block whose B^3 voxel values are stored in ascending block order
// (:166-190, :277); interior/exterior blocks store nothing.  Every value here
// is bit-identical to evaluating the full recipe at that point.
//
// MI355X design:
//  * One work item = one superblock (SB^3 blocks, 32-64 voxels on a side) of
//    one recipe; any number of recipes run in one launch through a table of
//    field descriptors (batched builds and the server use the same kernels).
//  * Exact Lipschitz culling per superblock: each primitive is bounded on the
//    superblock's box by value(centre) +- (half diagonal + fp32 slack); the
//    ordered min/max fold of those intervals is a composition of clamps, so a
//    workgroup-wide scan gives the fold interval at every step, and an
//    operand that is strictly worse everywhere than another operand of its
//    run of same-op steps is dropped.  The selection min/max returns the
//    extreme operand (the last of equal ones), which a dropped operand never
//    is, so results do not change.
//  * Corner and voxel loops run the primitive list in the outer loop with
//    scalar (wave-uniform) parameter loads and several points per lane.
//  * Compaction: wave ballots + popcount prefix; each chunk of 1024 blocks
//    finds its output offset by summing the counts of earlier chunks.
//  * Tiles: a persistent grid of waves walks every recipe's active blocks;
//    n_active is read on the device, so the whole build needs no host sync.
#include "recipe_eval.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>

__device__ __forceinline__ int find_item(const int32_t* prefix, int R, int item)
{
    int lo = 0, hi = R;  // largest r with prefix[r] <= item
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (prefix[mid] <= item) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Builds the ordered list of primitives that can change the fold somewhere in
// the box.  Called by all 256 threads; returns the list length (list in LDS).
__device__ int cull_box(const Prim* __restrict__ prims, int nprims, float cx, float cy, float cz, float rad,
                        float slack, uint16_t* list, Clamp2* wtot, int* wcnt)
{
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    float Flo = 0.0f, Fhi = 0.0f;  // fold interval after the processed chunks
    int n = 0;
    for (int base = 0; base < nprims; base += 256) {
        const int i = base + tid;
        Clamp2 cl = {-INFINITY, INFINITY, -INFINITY, INFINITY};
        float lo = 0.0f, hi = 0.0f;
        int op = 0;
        if (i < nprims) {
            const Prim p = prims[i];
            prim_interval(p, cx, cy, cz, rad, slack, lo, hi);
            op = p.op;
            if (i == 0) cl = {lo, lo, hi, hi};
            else if (op == 0) cl = {-INFINITY, lo, -INFINITY, hi};
            else if (op == 1) cl = {-hi, INFINITY, -lo, INFINITY};
            else cl = {lo, INFINITY, hi, INFINITY};
        }
        // inclusive wave scan
        Clamp2 inc = cl;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            Clamp2 o = cshfl_up(inc, d);
            if (lane >= d) inc = ccompose(o, inc);
        }
        if (lane == 63) wtot[wid] = inc;
        __syncthreads();
        Clamp2 pre = {-INFINITY, INFINITY, -INFINITY, INFINITY};
        for (int w = 0; w < wid; ++w) pre = ccompose(pre, wtot[w]);
        const Clamp2 run = ccompose(pre, run_end(inc, op, i == 0, i < nprims, i + 1 < nprims));
        const float rlo = capply(Flo, run.ll, run.lu), rhi = capply(Fhi, run.hl, run.hu);
        const bool keep = i < nprims && keep_step(op, i == 0, lo, hi, rlo, rhi);
        const uint64_t bal = __ballot(keep);
        if (lane == 0) wcnt[wid] = __popcll(bal);
        __syncthreads();
        int off = n;
        for (int w = 0; w < wid; ++w) off += wcnt[w];
        if (keep) list[off + __popcll(bal & ((1ull << lane) - 1))] = (uint16_t)i;
        Clamp2 all = {-INFINITY, INFINITY, -INFINITY, INFINITY};
        for (int w = 0; w < 4; ++w) all = ccompose(all, wtot[w]);
        n += wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
        const float nlo = capply(Flo, all.ll, all.lu), nhi = capply(Fhi, all.hl, all.hu);
        Flo = nlo;
        Fhi = nhi;
        __syncthreads();
    }
    return n;
}

// Box (centre, half diagonal) of blocks [bi, bi+nx) x [bj, bj+ny) x [bk, bk+nz),
// corners included.
__device__ __forceinline__ void block_box3(const FieldDesc& D, int bi, int bj, int bk, int nx, int ny, int nz,
                                           float& cx, float& cy, float& cz, float& rad)
{
    const int B = D.B;
    const float lx = D.x0 + (float)(bi * B) * D.pitch, ux = D.x0 + (float)((bi + nx) * B) * D.pitch;
    const float ly = D.y0 + (float)(bj * B) * D.pitch, uy = D.y0 + (float)((bj + ny) * B) * D.pitch;
    const float lz = D.z0 + (float)(bk * B) * D.pitch, uz = D.z0 + (float)((bk + nz) * B) * D.pitch;
    cx = 0.5f * (lx + ux);
    cy = 0.5f * (ly + uy);
    cz = 0.5f * (lz + uz);
    const float ex = ux - lx, ey = uy - ly, ez = uz - lz;
    rad = 0.5f * sqrtf(ex * ex + ey * ey + ez * ez) * 1.0001f;
}

// Box (centre, half diagonal) of blocks [b0, b0+nb) per axis, corners included.
__device__ __forceinline__ void block_box(const FieldDesc& D, int bi, int bj, int bk, int nb, float& cx, float& cy,
                                          float& cz, float& rad)
{
    const int B = D.B;
    const float lx = D.x0 + (float)(bi * B) * D.pitch, ux = D.x0 + (float)((bi + nb) * B) * D.pitch;
    const float ly = D.y0 + (float)(bj * B) * D.pitch, uy = D.y0 + (float)((bj + nb) * B) * D.pitch;
    const float lz = D.z0 + (float)(bk * B) * D.pitch, uz = D.z0 + (float)((bk + nb) * B) * D.pitch;
    cx = 0.5f * (lx + ux);
    cy = 0.5f * (ly + uy);
    cz = 0.5f * (lz + uz);
    const float ex = ux - lx, ey = uy - ly, ez = uz - lz;
    rad = 0.5f * sqrtf(ex * ex + ey * ey + ez * ez) * 1.0001f;
}

// ---------------------------------------------------------------- classify
// One workgroup per superblock: the workgroup culls the recipe for the
// superblock box (stored for the tile pass), then each wave takes groups of
// 4x2x2 blocks, refines the list for the group's box and evaluates the
// group's corner points (shared lattice points once when exact).
// The lists live in dynamic LDS sized to the launch's longest recipe: the
// superblock list holds lcap entries (>= every recipe's primitive count), each
// wave's group list min(lcap, FE_MAX_PRIMS / 2) (a longer superblock list is
// evaluated unrefined), so small recipes keep the block's LDS small.
// evals (optional, device uint64[6]): [0] primitive evaluations at classified
// lattice points, [1] those points, [4] interval evaluations of the culling
// (in this kernel); [2], [3], [5] the same for the tile pass (k_tiles).
__global__ __launch_bounds__(256) void k_classify(const FieldDesc* __restrict__ descs, const int32_t* __restrict__ sb_prefix,
                                                  int R, int total, uint32_t grid, int lcap,
                                                  unsigned long long* __restrict__ evals)
{
    extern __shared__ uint16_t lists_lds[];
    __shared__ Clamp2 wtot[4];
    __shared__ int wcnt[4];
    const int wcap = min(lcap, FE_MAX_PRIMS / 2);
    uint16_t* list = lists_lds;
    const uint32_t item = xcd_tile(blockIdx.x, grid);
    if (item >= (uint32_t)total) return;
    const int r = find_item(sb_prefix, R, item);
    const FieldDesc& D = descs[r];
    const int sb = item - sb_prefix[r];
    const int SB = D.SB, B = D.B;
    const int sbz = sb % D.nsbz, sby = (sb / D.nsbz) % D.nsby, sbx = sb / (D.nsbz * D.nsby);
    const int bx0 = sbx * SB, by0 = sby * SB, bz0 = sbz * SB;
    float cx, cy, cz, rad;
    block_box(D, bx0, by0, bz0, SB, cx, cy, cz, rad);
    // a recipe longer than the LDS lists (a launch given too small a max_prims)
    // is evaluated without culling: every list then means all primitives
    const bool fits = D.nprims <= lcap;
    const int n = fits ? cull_box(D.prims, D.nprims, cx, cy, cz, rad, D.slack, list, wtot, wcnt) : D.nprims;
    const int tid = threadIdx.x, lane = tid & 63, wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    if (tid == 0) D.list_len[sb] = (fits && n <= D.cap) ? n : -1;
    if (evals && tid == 0 && fits) atomicAdd(&evals[4], (unsigned long long)D.nprims);
    if (fits && n <= D.cap)
        for (int j = tid; j < n; j += 256) D.lists[(int64_t)sb * D.cap + j] = list[j];

    // Groups of 4 x 2 x 2 blocks, one per wave per pass.  The corner "+1" of a
    // block, origin + B*h, is usually the next block's origin bit for bit;
    // the wave checks that for its group and then evaluates the 5 x 3 x 3
    // lattice of block origins once (lane l = point l) and gathers each
    // block's eight corners by lane shuffles.  Otherwise (origin or pitch
    // rounding) every corner is evaluated as its own point, 8 blocks per pass.
    const int GX = SB / 4, GY = SB / 2, GZ = SB / 2, ngroups = GX * GY * GZ;
    const float side = (float)B * D.pitch;
    uint16_t* wl = lists_lds + lcap + wid * wcap;
    for (int g = wid; g < ngroups; g += 4) {
        const int gx = bx0 + 4 * (g / (GY * GZ)), gy = by0 + 2 * ((g / GZ) % GY), gz = bz0 + 2 * (g % GZ);
        block_box3(D, gx, gy, gz, 4, 2, 2, cx, cy, cz, rad);
        // a group list never exceeds the superblock list; fall back to it if it
        // would not fit the per-wave buffer
        const uint16_t* gl = fits ? list : nullptr;
        int gn = n;
        if (fits && n <= wcap) {
            // one pass: a second one (as the tiles run) costs more than the
            // few entries it would drop from a 45-point lattice's list
            gn = cull_wave(D.prims, list, n, cx, cy, cz, rad, D.slack, wl);
            gl = wl;
            if (evals && lane == 0) atomicAdd(&evals[4], (unsigned long long)n);
        }
        const int la = lane / 9, lb = (lane / 3) % 3, lc = lane % 3;
        const float lx = D.x0 + (float)((gx + la) * B) * D.pitch;
        const float ly = D.y0 + (float)((gy + lb) * B) * D.pitch;
        const float lz = D.z0 + (float)((gz + lc) * B) * D.pitch;
        auto same = [](float p, float q) { return __float_as_uint(p) == __float_as_uint(q); };
        const bool ok_x = la == 0 || same((D.x0 + (float)((gx + la - 1) * B) * D.pitch) + side, lx);
        const bool ok_y = lb == 0 || same((D.y0 + (float)((gy + lb - 1) * B) * D.pitch) + side, ly);
        const bool ok_z = lc == 0 || same((D.z0 + (float)((gz + lc - 1) * B) * D.pitch) + side, lz);
        const uint64_t lat = (1ull << 45) - 1;
        const bool lattice = (__ballot(ok_x && ok_y && ok_z) & lat) == lat;
        if (evals && lane == 0) {
            const unsigned long long pts = lattice ? 45ull : 128ull;
            atomicAdd(&evals[0], pts * (unsigned long long)gn);
            atomicAdd(&evals[1], pts);
        }
        float lv = 0.0f;
        if (lattice) {
            float x = lx, y = ly, z = lz;
            eval_points<1>(D.prims, gl, gn, &x, &y, &z, &lv);
        }
        const int c = lane & 7;
        for (int half = 0; half < 2; ++half) {
            const int bl = half * 8 + (lane >> 3);
            const int ba = bl >> 2, bb = (bl >> 1) & 1, bc = bl & 1;
            const int bi = gx + ba, bj = gy + bb, bk = gz + bc;
            float v;
            if (lattice) {
                const int idx = ((ba + (c & 1)) * 3 + bb + ((c >> 1) & 1)) * 3 + bc + ((c >> 2) & 1);
                v = __shfl(lv, idx);
            } else {
                const float ox = D.x0 + (float)(bi * B) * D.pitch;
                const float oy = D.y0 + (float)(bj * B) * D.pitch;
                const float oz = D.z0 + (float)(bk * B) * D.pitch;
                float x = ox + (float)(c & 1) * side, y = oy + (float)((c >> 1) & 1) * side,
                      z = oz + (float)((c >> 2) & 1) * side;
                eval_points<1>(D.prims, gl, gn, &x, &y, &z, &v);
            }
            float mn = fabsf(v);
            int pos = v >= 0.0f, neg = !pos;
#pragma unroll
            for (int s = 1; s < 8; s <<= 1) {
                mn = sel_min(mn, __shfl_xor(mn, s));
                pos |= __shfl_xor(pos, s);
                neg |= __shfl_xor(neg, s);
            }
            mn = sel_min(1.0e30f, mn);
            if (c == 0 && bi < D.nbx && bj < D.nby && bk < D.nbz) {
                const int64_t b = ((int64_t)bi * D.nby + bj) * D.nbz + bk;
                D.cma[b] = mn;
                D.kind[b] = (pos && neg) ? 0 : (mn <= D.margin ? 0 : (neg ? -1 : 1));
            }
        }
    }
}

// ---------------------------------------------------------------- compaction
__global__ __launch_bounds__(256) void k_compact_count(const FieldDesc* __restrict__ descs,
                                                       const int32_t* __restrict__ ch_prefix, int R, int total,
                                                       int32_t* __restrict__ chunk_count)
{
    __shared__ int wc[4];
    const int item = blockIdx.x;
    if (item >= total) return;
    const int r = find_item(ch_prefix, R, item);
    const FieldDesc& D = descs[r];
    const int64_t nb = (int64_t)D.nbx * D.nby * D.nbz;
    const int64_t b0 = (int64_t)(item - ch_prefix[r]) * FE_CHUNK;
    int c = 0;
    for (int k = 0; k < FE_CHUNK / 256; ++k) {
        const int64_t b = b0 + k * 256 + threadIdx.x;
        c += __popcll(__ballot(b < nb && D.kind[b] == 0));
    }
    if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) chunk_count[item] = wc[0] + wc[1] + wc[2] + wc[3];
}

__global__ __launch_bounds__(256) void k_compact_write(const FieldDesc* __restrict__ descs,
                                                       const int32_t* __restrict__ ch_prefix, int R, int total,
                                                       const int32_t* __restrict__ chunk_count)
{
    __shared__ int wsum[4];
    __shared__ int wc[4];
    const int item = blockIdx.x;
    if (item >= total) return;
    const int r = find_item(ch_prefix, R, item);
    const FieldDesc& D = descs[r];
    const int local = item - ch_prefix[r];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    int s = 0;
    for (int j = tid; j < local; j += 256) s += chunk_count[ch_prefix[r] + j];
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d);
    if (lane == 0) wsum[wid] = s;
    __syncthreads();
    int off = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    const int64_t nb = (int64_t)D.nbx * D.nby * D.nbz;
    const int64_t b0 = (int64_t)local * FE_CHUNK;
    for (int k = 0; k < FE_CHUNK / 256; ++k) {
        const int64_t b = b0 + k * 256 + tid;
        const bool act = b < nb && D.kind[b] == 0;
        const uint64_t bal = __ballot(act);
        __syncthreads();
        if (lane == 0) wc[wid] = __popcll(bal);
        __syncthreads();
        int o = off;
        for (int w = 0; w < wid; ++w) o += wc[w];
        o += __popcll(bal & ((1ull << lane) - 1));
        if (b < nb) {
            D.slot[b] = act ? o : -1;
            if (act) D.ids[o] = (int32_t)b;
        }
        off += wc[0] + wc[1] + wc[2] + wc[3];
    }
    if (local == ch_prefix[r + 1] - ch_prefix[r] - 1 && tid == 0) {
        D.counts[0] = off;
        D.counts[1] = off > D.tile_cap ? 1 : 0;
    }
}

// ---------------------------------------------------------------- tiles
// One wave evaluates one B^3 tile: it refines the superblock list for the
// block's own box (exact, as above) and evaluates KV voxels per lane per pass.
// `lin` is the block whose voxels are evaluated (the planted fault shifts it,
// and then evaluates the whole recipe), `bl` the listed block.  B (= D.B) is a
// compile-time edge: a lane's KV points of a pass then differ only in x (B = 8,
// 16) and the compiler shares their y and z terms in every primitive formula.
template <int B, int KV>
__device__ __forceinline__ void tile_eval(const FieldDesc& D, int64_t a, int bl, uint16_t* wl,
                                          unsigned long long* __restrict__ evals)
{
    const int lane = threadIdx.x & 63;
    constexpr int b3 = B * B * B;
    const int lin = bl + (D.corrupt ? 1 : 0);
    const int bk = lin % D.nbz, bj = (lin / D.nbz) % D.nby, bi = lin / (D.nbz * D.nby);
    const uint16_t* list = nullptr;
    int n = D.nprims;
    if (!D.corrupt) {
        const int sb = ((bl / (D.nbz * D.nby)) / D.SB * D.nsby + ((bl / D.nbz) % D.nby) / D.SB) * D.nsbz +
                       (bl % D.nbz) / D.SB;
        const int nl = D.list_len[sb];
        const uint16_t* src = nl >= 0 ? D.lists + (int64_t)sb * D.cap : nullptr;
        const int nsrc = nl >= 0 ? nl : D.nprims;
        if (nsrc <= FE_MAX_PRIMS / 2) {
            float cx, cy, cz, rad;
            block_box(D, bi, bj, bk, 1, cx, cy, cz, rad);
            n = cull_wave(D.prims, src, nsrc, cx, cy, cz, rad, D.slack, wl);
            if (evals && lane == 0) atomicAdd(&evals[5], (unsigned long long)(nsrc + (n > 8 ? n : 0)));
            // a second pass over the refined list (in place): its runs join
            // across dropped steps and are no longer cut every 64 entries
            if (n > 8) n = cull_wave(D.prims, wl, n, cx, cy, cz, rad, D.slack, wl);
            list = wl;
        } else {
            list = src;
            n = nsrc;
        }
    }
    if (evals && lane == 0) {
        atomicAdd(&evals[2], (unsigned long long)n * b3);
        atomicAdd(&evals[3], (unsigned long long)b3);
    }
    float* out = D.tiles + a * b3;
    for (int v0 = 0; v0 < b3; v0 += 64 * KV) {
        float x[KV], y[KV], z[KV], f[KV];
#pragma unroll
        for (int k = 0; k < KV; ++k) {
            const int l = min(v0 + k * 64 + lane, b3 - 1);
            const int li = l / (B * B), lj = (l / B) % B, lk = l % B;
            x[k] = D.x0 + (float)(bi * B + li) * D.pitch;
            y[k] = D.y0 + (float)(bj * B + lj) * D.pitch;
            z[k] = D.z0 + (float)(bk * B + lk) * D.pitch;
        }
        eval_points<KV>(D.prims, list, n, x, y, z, f);
#pragma unroll
        for (int k = 0; k < KV; ++k) {
            const int l = v0 + k * 64 + lane;
            if (l < b3) out[l] = f[k];
        }
    }
}

// Persistent: the active blocks of all recipes, concatenated in recipe order,
// are dealt round-robin over the waves (tile g of the concatenation goes to
// wave g mod nwaves), so a recipe with fewer tiles than waves does not land
// on the same low-numbered waves as every other recipe.
template <int B, int KV>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_tiles(const FieldDesc* __restrict__ descs, int R,
                                                                                          unsigned long long* __restrict__ evals)
{
    __shared__ uint16_t wlist[4][FE_MAX_PRIMS / 2];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    const int64_t wave = (int64_t)blockIdx.x * 4 + wv;
    const int64_t nwaves = (int64_t)gridDim.x * 4;
    int64_t base = 0;  // tiles of the recipes before r, mod nwaves
    for (int r = 0; r < R; ++r) {
        const FieldDesc& D = descs[r];
        const int64_t na = min((int64_t)D.counts[0], D.tile_cap);
        const int64_t a0 = wave >= base ? wave - base : wave - base + nwaves;
        for (int64_t a = a0; a < na; a += nwaves) tile_eval<B, KV>(D, a, D.ids[a], wlist[wv], evals);
        base = (base + na) % nwaves;
    }
}

// Tiles for an explicit host-provided id list (the plain per-call path); uses
// the culled lists that the classify launch stored.
template <int B, int KV>
__global__ __launch_bounds__(256) void k_tiles_ids(FieldDesc D, const int32_t* __restrict__ ids, int64_t na)
{
    __shared__ uint16_t wlist[4][FE_MAX_PRIMS / 2];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    const int64_t a = (int64_t)blockIdx.x * 4 + wv;
    if (a >= na) return;
    tile_eval<B, KV>(D, a, ids[a], wlist[wv], nullptr);
}

// ---------------------------------------------------------------- reconstruct / checksum
// Cropped dense field (FE/faltkarna_v1.py:302-325) and, when occ != nullptr,
// the occupancy dense < 0 (FE/faltkarna_v1_mesh_to_sdf.py:695-710).  One
// thread writes 4 consecutive z (one block, since B is a multiple of 4) with
// one 16-byte store (4-byte for the occupancy) when rows are 4-aligned.
__global__ __launch_bounds__(256) void k_reconstruct(const FieldDesc* __restrict__ descs, int r, float* __restrict__ dense,
                                                     uint8_t* __restrict__ occ)
{
    // One row (x, y) per wave at a time: the row's block row and in-tile
    // offset are wave-uniform (scalar), the lanes walk z four voxels apiece
    // with 32-bit arithmetic.
    const FieldDesc& D = descs[r];
    const int nz4 = (D.nz + 3) >> 2;
    const int64_t rows = (int64_t)D.nx * D.ny;
    const int B = D.B;
    const bool vec = (D.nz & 3) == 0;
    const int lane = threadIdx.x & 63;
    const int64_t first = __builtin_amdgcn_readfirstlane((int)(((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6));
    const int64_t stride = ((int64_t)gridDim.x * 256) >> 6;
    for (int64_t xy = first; xy < rows; xy += stride) {
        const int x = (int)(xy / D.ny), y = (int)(xy - (int64_t)x * D.ny);
        const int64_t brow = ((int64_t)(x / B) * D.nby + y / B) * D.nbz;
        const int toff = ((x % B) * B + (y % B)) * B;
        const int64_t o0 = xy * D.nz;
        for (int z4 = lane; z4 < nz4; z4 += 64) {
            const int z0 = z4 * 4, bz = z0 / B;
            const int64_t b = brow + bz;
            const int s = D.slot[b];
            float v[4];
            if (s >= 0 && s < D.tile_cap) {
                const float* t = D.tiles + (int64_t)s * B * B * B + toff + (z0 - bz * B);
                const float4 q = *reinterpret_cast<const float4*>(t);
                v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
            } else {
                const float bg = D.kind[b] < 0 ? -FE_BIG : FE_BIG;
                v[0] = v[1] = v[2] = v[3] = bg;
            }
            const int64_t o = o0 + z0;
            if (vec) {
                // the dense field streams past the caches; the occupancy, which
                // the EDT reads next, is stored normally and may stay cached
                typedef float f4 __attribute__((ext_vector_type(4)));
                const f4 w = {v[0], v[1], v[2], v[3]};
                __builtin_nontemporal_store(w, reinterpret_cast<f4*>(dense + o));
                if (occ) {
                    const uint32_t m = (v[0] < 0.0f ? 1u : 0u) | (v[1] < 0.0f ? 1u << 8 : 0u) |
                                       (v[2] < 0.0f ? 1u << 16 : 0u) | (v[3] < 0.0f ? 1u << 24 : 0u);
                    *reinterpret_cast<uint32_t*>(occ + o) = m;
                }
            } else {
                for (int k = 0; k < 4 && z0 + k < D.nz; ++k) {
                    dense[o + k] = v[k];
                    if (occ) occ[o + k] = v[k] < 0.0f ? 1 : 0;
                }
            }
        }
    }
}

// Sum over stored tile values v < 0 of int64(f64(v)*scale) (FE/faltkarna_v1.py:249-255).
__global__ __launch_bounds__(256) void k_checksum(const FieldDesc* __restrict__ descs, int r, double scale,
                                                  unsigned long long* __restrict__ out)
{
    const FieldDesc& D = descs[r];
    const int64_t n = min((int64_t)D.counts[0], D.tile_cap) * D.B * D.B * D.B;
    long long s = 0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const float v = D.tiles[i];
        if (v < 0.0f) s += (long long)((double)v * scale);
    }
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, (unsigned long long)s);
}

// ---------------------------------------------------------------- point queries
// Point SDF of a resident field: the recipe's fp32 fold at arbitrary points
// (the same arithmetic as the stored tiles), evaluated over the culled list
// of the superblock containing the point; points outside every superblock
// box, or in a superblock that kept no list, evaluate the whole recipe.

// The superblock whose list holds point (x, y, z) and that list (-1, nullptr:
// the whole recipe).
__device__ __forceinline__ int query_list(const FieldDesc& D, float x, float y, float z, const uint16_t*& list,
                                          int& nl)
{
    list = nullptr;
    nl = D.nprims;
    const float span = (float)(D.SB * D.B) * D.pitch;
    const int sx = (int)floorf((x - D.x0) / span), sy = (int)floorf((y - D.y0) / span), sz = (int)floorf((z - D.z0) / span);
    if (sx >= 0 && sy >= 0 && sz >= 0 && sx < D.nsbx && sy < D.nsby && sz < D.nsbz) {
        // the point must lie in the box the list was culled for
        const int bx0 = sx * D.SB, by0 = sy * D.SB, bz0 = sz * D.SB, B = D.B;
        const float lx = D.x0 + (float)(bx0 * B) * D.pitch, ux = D.x0 + (float)((bx0 + D.SB) * B) * D.pitch;
        const float ly = D.y0 + (float)(by0 * B) * D.pitch, uy = D.y0 + (float)((by0 + D.SB) * B) * D.pitch;
        const float lz = D.z0 + (float)(bz0 * B) * D.pitch, uz = D.z0 + (float)((bz0 + D.SB) * B) * D.pitch;
        const int sb = (sx * D.nsby + sy) * D.nsbz + sz;
        const int len = D.list_len[sb];
        if (len >= 0 && x >= lx && x <= ux && y >= ly && y <= uy && z >= lz && z <= uz) {
            list = D.lists + (int64_t)sb * D.cap;
            nl = len;
            return sb;
        }
    }
    return -1;
}

// One thread per point, each over its own list (small queries).
__global__ __launch_bounds__(256) void k_query_points(const FieldDesc* __restrict__ desc, const float* __restrict__ pts,
                                                      int64_t n, float* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const FieldDesc& D = *desc;
    const float x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
    const uint16_t* list;
    int nl;
    query_list(D, x, y, z, list, nl);
    float f = 0.0f;
    if (nl > 0) {
        // the next record and the index after it are in flight while a
        // primitive evaluates (a thread's list is one chain of loads)
        int inext = nl > 1 ? (list ? list[1] : 1) : 0;
        Prim pn = D.prims[list ? list[0] : 0];
        for (int j = 0; j < nl; ++j) {
            const Prim p = pn;
            if (j + 1 < nl) {
                pn = D.prims[inext];
                if (j + 2 < nl) inext = list ? list[j + 2] : j + 2;
            }
            const float v = prim_eval(p.kind, p.cx, p.cy, p.cz, p.a, p.b, p.c, x, y, z);
            f = j == 0 ? v : fold(p.op, f, v);
        }
    }
    out[i] = f;
}

// Large queries: the points are ordered by superblock first (FE_QBINS bins
// of superblock ids, one more for the whole recipe; a bin may mix
// superblocks, which costs only uniformity), so that a wave's lanes share one
// or two lists: each list then runs once for the wave with its primitive
// records read by scalar loads and one primitive formula at a time, where
// lanes of scattered points would step through every kind of formula and
// gather their own records.  Values are bit-identical to k_query_points.
#define FE_QBINS 8192
// Fewer points: one thread per point.  A wave then runs one list per lane at
// once, while sorted waves run each of their lists in turn, which pays only
// where many points share a superblock (a 512^3 part's 2^20 uniform points:
// 0.47 ms sorted against 1.06 ms; 10^5 points: 0.46 against 0.32 ms).
#define FE_QSORT_MIN (1 << 18)

__global__ __launch_bounds__(256) void k_qbin_zero(int32_t* __restrict__ cnt)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < FE_QBINS + 2) cnt[t] = 0;
}

__global__ __launch_bounds__(256) void k_qbin_count(const FieldDesc* __restrict__ desc, const float* __restrict__ pts,
                                                    int64_t n, int32_t* __restrict__ cnt, int32_t* __restrict__ bin,
                                                    int32_t* __restrict__ rank)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint16_t* list;
    int nl;
    const int sb = query_list(*desc, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], list, nl);
    const int b = sb < 0 ? FE_QBINS : sb % FE_QBINS;
    bin[i] = b;
    rank[i] = atomicAdd(cnt + b, 1);
}

// Exclusive scan of the FE_QBINS + 1 bin counts, in place (one workgroup).
__global__ __launch_bounds__(1024) void k_qbin_scan(int32_t* __restrict__ cnt)
{
    constexpr int NB = FE_QBINS + 1, PER = (NB + 1023) / 1024;
    __shared__ int part[1024];
    const int t = threadIdx.x;
    int loc[PER], s = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int idx = t * PER + k;
        loc[k] = idx < NB ? cnt[idx] : 0;
        s += loc[k];
    }
    part[t] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const int v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int base = part[t] - s;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int idx = t * PER + k;
        if (idx < NB) {
            cnt[idx] = base;
            base += loc[k];
        }
    }
}

__global__ __launch_bounds__(256) void k_qbin_place(const int32_t* __restrict__ off, const int32_t* __restrict__ bin,
                                                    const int32_t* __restrict__ rank, int64_t n,
                                                    int32_t* __restrict__ perm)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    perm[off[bin[i]] + rank[i]] = (int32_t)i;
}

__global__ __launch_bounds__(256) void k_query_sorted(const FieldDesc* __restrict__ desc, const float* __restrict__ pts,
                                                      const int32_t* __restrict__ perm, int64_t n,
                                                      float* __restrict__ out)
{
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool valid = k < n;
    const FieldDesc& D = *desc;
    const int64_t i = valid ? perm[k] : 0;
    float x = 0.0f, y = 0.0f, z = 0.0f;
    int key = -2;  // -2: no point; -1: the whole recipe; else the superblock
    if (valid) {
        x = pts[3 * i];
        y = pts[3 * i + 1];
        z = pts[3 * i + 2];
        const uint16_t* list;
        int nl;
        key = query_list(D, x, y, z, list, nl);
    }
    uint64_t todo = __ballot(valid);
    float f = 0.0f;
    while (todo) {  // the distinct lists of the wave, one at a time, all lanes running
        const int lead = __builtin_amdgcn_readlane(key, (int)__builtin_ctzll(todo));
        const bool mine = valid && key == lead;
        const uint16_t* list = lead >= 0 ? D.lists + (int64_t)lead * D.cap : nullptr;
        const int nl = lead >= 0 ? D.list_len[lead] : D.nprims;
        float v = 0.0f;
        eval_points<1>(D.prims, list, nl, &x, &y, &z, &v);
        if (mine) f = v;
        todo &= ~__ballot(mine);
    }
    if (valid) out[i] = f;
}

// ---------------------------------------------------------------- dense-source classify
// FE/faltkarna_v1_mesh_to_sdf.py:602-632 (corners clamped to n-1).
__global__ __launch_bounds__(256) void k_classify_dense(const float* __restrict__ dense, int nx, int ny, int nz, int B,
                                                        int nbx, int nby, int nbz, float margin, int32_t* __restrict__ kind)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (int64_t)nbx * nby * nbz) return;
    const int bk = (int)(t % nbz), bj = (int)((t / nbz) % nby), bi = (int)(t / ((int64_t)nbz * nby));
    float mn = 1.0e30f;
    int pos = 0, neg = 0;
    for (int c = 0; c < 8; ++c) {
        const int cx = min(bi * B + (c & 1) * B, nx - 1), cy = min(bj * B + ((c >> 1) & 1) * B, ny - 1),
                  cz = min(bk * B + ((c >> 2) & 1) * B, nz - 1);
        const float v = dense[((int64_t)cx * ny + cy) * nz + cz];
        mn = sel_min(mn, fabsf(v));
        if (v >= 0.0f) pos = 1;
        else neg = 1;
    }
    kind[t] = (pos && neg) ? 0 : (mn <= margin ? 0 : (neg ? -1 : 1));
}

// FE/faltkarna_v1_mesh_to_sdf.py:635-654 (voxels clamped to n-1).
__global__ __launch_bounds__(256) void k_tiles_dense(const float* __restrict__ dense, int nx, int ny, int nz, int B,
                                                     int nby, int nbz, const int32_t* __restrict__ ids, int64_t na,
                                                     float* __restrict__ tiles)
{
    const int64_t b3 = (int64_t)B * B * B;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= na * b3) return;
    const int64_t a = t / b3;
    const int l = (int)(t % b3);
    const int lin = ids[a];
    const int bk = lin % nbz, bj = (lin / nbz) % nby, bi = lin / (nbz * nby);
    const int x = min(bi * B + l / (B * B), nx - 1), y = min(bj * B + (l / B) % B, ny - 1), z = min(bk * B + l % B, nz - 1);
    tiles[t] = dense[((int64_t)x * ny + y) * nz + z];
}

// ---------------------------------------------------------------- C ABI
// Launches classify -> compaction -> tiles for R recipes described by the
// device descriptor table.  sb_prefix / ch_prefix: device int32[R+1]
// exclusive prefixes of superblock and 1024-block chunk counts; chunk_count:
// device int32[ch_prefix[R]].  Every recipe in one call shares one block edge.
// LDS list capacity (entries) of a classify launch whose longest recipe has
// max_prims primitives, and the dynamic LDS bytes it takes.
static int classify_lcap(int max_prims) { return (max_prims + 7) & ~7; }
static size_t classify_lds(int lcap) { return (size_t)2 * (lcap + 4 * std::min(lcap, FE_MAX_PRIMS / 2)); }

static int field_build(const void* descs, int R, int B, const int32_t* sb_prefix, int total_sb,
                       const int32_t* ch_prefix, int total_chunks, int32_t* chunk_count, int tile_waves, int max_prims,
                       unsigned long long* evals, hipStream_t st)
{
    if (!descs || R < 1 || (B != 4 && B != 8 && B != 16) || total_sb < 1 || total_chunks < 1 || !chunk_count ||
        max_prims < 1 || max_prims > FE_MAX_PRIMS)
        return FE_INVALID;
    const FieldDesc* d = static_cast<const FieldDesc*>(descs);
    const uint32_t g = round8(total_sb);
    const int lcap = classify_lcap(max_prims);
    hipLaunchKernelGGL(k_classify, dim3(g), dim3(256), classify_lds(lcap), st, d, sb_prefix, R, total_sb, g, lcap,
                       evals);
    FE_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_compact_count, dim3(total_chunks), dim3(256), 0, st, d, ch_prefix, R, total_chunks, chunk_count);
    FE_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_compact_write, dim3(total_chunks), dim3(256), 0, st, d, ch_prefix, R, total_chunks, chunk_count);
    FE_TRY(hipGetLastError());
    const int blocks = tile_waves > 0 ? (tile_waves + 3) / 4 : 2048;
    if (B == 4) hipLaunchKernelGGL((k_tiles<4, 1>), dim3(blocks), dim3(256), 0, st, d, R, evals);
    // B = 8: one pass of 8 voxels per lane; with the edge a compile-time
    // constant the 8 points share y and z and fit 64 VGPRs (8 waves/SIMD)
    // without spills
    else if (B == 8) hipLaunchKernelGGL((k_tiles<8, 8>), dim3(blocks), dim3(256), 0, st, d, R, evals);
    else hipLaunchKernelGGL((k_tiles<16, 4>), dim3(blocks), dim3(256), 0, st, d, R, evals);
    FE_TRY(hipGetLastError());
    return FE_OK;
}

extern "C" int fe_field_build(const void* descs, int R, int B, const int32_t* sb_prefix, int total_sb,
                              const int32_t* ch_prefix, int total_chunks, int32_t* chunk_count, int tile_waves,
                              int max_prims, void* stream) noexcept
{
    return field_build(descs, R, B, sb_prefix, total_sb, ch_prefix, total_chunks, chunk_count, tile_waves, max_prims,
                       nullptr, static_cast<hipStream_t>(stream));
}

// fe_field_build that also counts the primitive evaluations (evals: device
// uint64[6], added to; layout at k_classify).
extern "C" int fe_field_build_counted(const void* descs, int R, int B, const int32_t* sb_prefix, int total_sb,
                                      const int32_t* ch_prefix, int total_chunks, int32_t* chunk_count,
                                      int tile_waves, int max_prims, unsigned long long* evals, void* stream) noexcept
{
    if (!evals) return FE_INVALID;
    return field_build(descs, R, B, sb_prefix, total_sb, ch_prefix, total_chunks, chunk_count, tile_waves, max_prims,
                       evals, static_cast<hipStream_t>(stream));
}

// Classification only (the first launch of the plain per-call path).
extern "C" int fe_field_classify(const void* descs, int R, int B, const int32_t* sb_prefix, int total_sb,
                                 int max_prims, void* stream) noexcept
{
    if (!descs || R < 1 || (B != 4 && B != 8 && B != 16) || total_sb < 1 || max_prims < 1 ||
        max_prims > FE_MAX_PRIMS)
        return FE_INVALID;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const FieldDesc* d = static_cast<const FieldDesc*>(descs);
    const uint32_t g = round8(total_sb);
    const int lcap = classify_lcap(max_prims);
    hipLaunchKernelGGL(k_classify, dim3(g), dim3(256), classify_lds(lcap), st, d, sb_prefix, R, total_sb, g, lcap,
                       (unsigned long long*)nullptr);
    FE_TRY(hipGetLastError());
    return FE_OK;
}

// Tiles for host-compacted ids (plain per-call path): desc is a host copy of
// one descriptor; ids and the tile buffer are device pointers.
extern "C" int fe_field_tiles_ids(const void* host_desc, const int32_t* ids, int64_t n_active, void* stream) noexcept
{
    if (!host_desc || !ids || n_active < 0) return FE_INVALID;
    if (n_active == 0) return FE_OK;
    FieldDesc D;
    std::memcpy(&D, host_desc, sizeof D);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const unsigned blocks = (unsigned)((n_active + 3) / 4);
    if (D.B == 4) hipLaunchKernelGGL((k_tiles_ids<4, 1>), dim3(blocks), dim3(256), 0, st, D, ids, n_active);
    else if (D.B == 8) hipLaunchKernelGGL((k_tiles_ids<8, 4>), dim3(blocks), dim3(256), 0, st, D, ids, n_active);
    else if (D.B == 16) hipLaunchKernelGGL((k_tiles_ids<16, 4>), dim3(blocks), dim3(256), 0, st, D, ids, n_active);
    else return FE_INVALID;
    FE_TRY(hipGetLastError());
    return FE_OK;
}

// Cropped dense reconstruction of descriptor r (+ optional uint8 occupancy).
extern "C" int fe_field_reconstruct(const void* descs, int r, int64_t n_voxels, float* dense, uint8_t* occ,
                                    void* stream) noexcept
{
    if (!descs || r < 0 || !dense || n_voxels < 1) return FE_INVALID;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const unsigned blocks = (unsigned)std::min<int64_t>((n_voxels / 4 + 255) / 256 + 1, 65536);
    hipLaunchKernelGGL(k_reconstruct, dim3(blocks), dim3(256), 0, st, static_cast<const FieldDesc*>(descs), r, dense, occ);
    FE_TRY(hipGetLastError());
    return FE_OK;
}

// int64 fixed-point checksum of descriptor r's tiles, accumulated into *out.
extern "C" int fe_field_checksum(const void* descs, int r, double scale, int64_t* out, void* stream) noexcept
{
    if (!descs || r < 0 || !out) return FE_INVALID;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_checksum, dim3(1024), dim3(256), 0, st, static_cast<const FieldDesc*>(descs), r, scale,
                       reinterpret_cast<unsigned long long*>(out));
    FE_TRY(hipGetLastError());
    return FE_OK;
}

// The two checksums of FE/faltkarna_v1.py:238-255 over a flat array of values
// below zero: out32 accumulates every value with a float atomic add (order-
// dependent, the negative control), out64 the int64 of f64(v) * scale
// (order-independent).  Both must be zeroed by the caller.
__global__ __launch_bounds__(256) void k_checksum_pair(const float* __restrict__ vals, int64_t n, double scale,
                                                       float* __restrict__ out32, unsigned long long* __restrict__ out64)
{
    long long s = 0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const float v = vals[i];
        if (v < 0.0f) {
            atomicAdd(out32, v);
            s += (long long)((double)v * scale);
        }
    }
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d);
    if ((threadIdx.x & 63) == 0) atomicAdd(out64, (unsigned long long)s);
}

extern "C" int fe_checksum_pair(const float* vals, int64_t n, double scale, float* out32, int64_t* out64,
                                void* stream) noexcept
{
    if (!vals || !out32 || !out64 || n < 0) return FE_INVALID;
    if (n == 0) return FE_OK;
    hipLaunchKernelGGL(k_checksum_pair, dim3(1024), dim3(256), 0, static_cast<hipStream_t>(stream), vals, n, scale,
                       out32, reinterpret_cast<unsigned long long*>(out64));
    FE_TRY(hipGetLastError());
    return FE_OK;
}

// Dense-source classify and tile extraction (FE/faltkarna_v1_mesh_to_sdf.py:660-692 on the device).
extern "C" int fe_dense_classify(const float* dense, int nx, int ny, int nz, int B, float margin, int32_t* kind,
                                 void* stream) noexcept
{
    if (!dense || !kind || nx < 1 || ny < 1 || nz < 1 || B < 1) return FE_INVALID;
    const int nbx = (nx + B - 1) / B, nby = (ny + B - 1) / B, nbz = (nz + B - 1) / B;
    const int64_t nb = (int64_t)nbx * nby * nbz;
    hipLaunchKernelGGL(k_classify_dense, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), dense, nx, ny, nz, B, nbx, nby, nbz, margin, kind);
    FE_TRY(hipGetLastError());
    return FE_OK;
}

extern "C" int fe_dense_tiles(const float* dense, int nx, int ny, int nz, int B, const int32_t* ids, int64_t n_active,
                              float* tiles, void* stream) noexcept
{
    if (!dense || !tiles || nx < 1 || ny < 1 || nz < 1 || B < 1 || n_active < 0) return FE_INVALID;
    if (n_active == 0) return FE_OK;
    const int nby = (ny + B - 1) / B, nbz = (nz + B - 1) / B;
    const int64_t n = n_active * B * B * B;
    hipLaunchKernelGGL(k_tiles_dense, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       dense, nx, ny, nz, B, nby, nbz, ids, n_active, tiles);
    FE_TRY(hipGetLastError());
    return FE_OK;
}

// ---------------------------------------------------------------- segment copies
// Many device-to-device copies in one launch (a field server moving every
// built field of a batch out of the batch's pools): segment s copies bytes
// [src, src + n) to dst, or clears dst when src is null.  blockIdx.y picks
// the segment, blockIdx.x strides over it; 4-byte words when the segment
// allows (every array of a field is 4-byte aligned except the uint16 lists).
struct CopySeg {
    const char* src;
    char* dst;
    int64_t n;
};
static_assert(sizeof(CopySeg) == 24, "segment record");

__global__ __launch_bounds__(256) void k_copy_segments(const CopySeg* __restrict__ segs)
{
    const CopySeg g = segs[blockIdx.y];
    const int64_t stride = (int64_t)gridDim.x * 256, t0 = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (((reinterpret_cast<uintptr_t>(g.src) | reinterpret_cast<uintptr_t>(g.dst) | (uintptr_t)g.n) & 3) == 0) {
        const uint32_t* s = reinterpret_cast<const uint32_t*>(g.src);
        uint32_t* d = reinterpret_cast<uint32_t*>(g.dst);
        for (int64_t i = t0; i < g.n / 4; i += stride) d[i] = s ? s[i] : 0u;
    } else {
        for (int64_t i = t0; i < g.n; i += stride) g.dst[i] = g.src ? g.src[i] : (char)0;
    }
}

// nseg segments (device table of {src, dst, bytes} records) in one launch.
extern "C" int fe_copy_segments(const void* segs, int nseg, void* stream) noexcept
{
    if (nseg < 0 || nseg > 65535 || (nseg && !segs)) return FE_INVALID;
    if (nseg == 0) return FE_OK;
    hipLaunchKernelGGL(k_copy_segments, dim3(256, (unsigned)nseg), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const CopySeg*>(segs));
    FE_TRY(hipGetLastError());
    return FE_OK;
}

// Point SDF queries on one resident field (desc: device pointer to its
// descriptor; pts: device float32 (n,3); out: device float32 (n)).
extern "C" int fe_query_points(const void* desc, const float* pts, int64_t n, float* out, void* stream) noexcept
{
    if (!desc || !pts || !out || n < 0) return FE_INVALID;
    if (n == 0) return FE_OK;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const FieldDesc* d = static_cast<const FieldDesc*>(desc);
    const dim3 grid((unsigned)((n + 255) / 256));
    if (n < FE_QSORT_MIN || n > INT32_MAX) {
        hipLaunchKernelGGL(k_query_points, grid, dim3(256), 0, st, d, pts, n, out);
        FE_TRY(hipGetLastError());
        return FE_OK;
    }
    // stream-ordered scratch (pooled by the runtime): bin counts / offsets,
    // then per point its bin, its rank in the bin and the sorted order
    int32_t* buf = nullptr;
    FE_TRY(hipMallocAsync(reinterpret_cast<void**>(&buf), sizeof(int32_t) * ((size_t)FE_QBINS + 2 + 3 * (size_t)n), st));
    int32_t* cnt = buf;
    int32_t* bin = cnt + FE_QBINS + 2;
    int32_t* rank = bin + n;
    int32_t* perm = rank + n;
    hipLaunchKernelGGL(k_qbin_zero, dim3((FE_QBINS + 2 + 255) / 256), dim3(256), 0, st, cnt);
    hipLaunchKernelGGL(k_qbin_count, grid, dim3(256), 0, st, d, pts, n, cnt, bin, rank);
    hipLaunchKernelGGL(k_qbin_scan, dim3(1), dim3(1024), 0, st, cnt);
    hipLaunchKernelGGL(k_qbin_place, grid, dim3(256), 0, st, cnt, bin, rank, n, perm);
    hipLaunchKernelGGL(k_query_sorted, grid, dim3(256), 0, st, d, pts, perm, n, out);
    const hipError_t e = hipGetLastError();
    const hipError_t f = hipFreeAsync(buf, st);
    FE_TRY(e);
    FE_TRY(f);
    return FE_OK;
}
