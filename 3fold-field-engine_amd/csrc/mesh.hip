// Mesh stage on gfx950: closed triangle mesh -> per-column +z hits -> column
// mask and surface -> paired EDT -> signed distance, plus the ordered
// generalized winding number for open meshes.  The stage is the production
// caller of the EDT (FE/mesh_field_cuda_columns_v1.py:63-73,
// FE/mesh_fused_native_v1/field.cu, FE/mesh_fused_graph_v1/field.cu).
//
// Column scan.  Queries (one x,y per grid column) are grouped into tiles of
// 256; a tile is a 16x16 square of grid columns when the queries form the
// mesh stage's grid, else 256 consecutive queries.  Per tile the list of
// triangles whose xy bounding box meets the tile's query bounds is built once
// (face order kept, ballot + prefix compaction), so a query tests only the
// triangles near it instead of all of them.  The arithmetic of every
// (query, triangle) test and of the depth is the reference's, in f64 with no
// contraction (the library is built with -ffp-contract=off), so hits are
// bit-identical, in the same per-column order.
//
// Each wave further keeps, of every 256 staged triangles, only those meeting
// its own 16 x 4 columns' box (ballot compaction, order kept).
//
// Mask.  The depth range is reduced in the scan kernel through order-preserving
// int64 atomics (min/max are exact, so the order of arrival does not matter),
// keys and per-z fractions follow FE/mesh_fused_native_v1/field.cu:14-23; one
// wave per column sums the signs of the hits above each voxel (the column's
// hits are wave-uniform, so they are scalar loads) and derives the
// 6-neighbour surface in the same pass.  The paired EDT and the signed output are
// edt.hip's device passes.  The whole chain is captured once into a HIP graph
// (no memset nodes: a one-thread kernel clears the range cells and flags).
#include "fe_common.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

extern "C" int fe_edt_device(const uint8_t* mask, int nx, int ny, int nz, int planes, int32_t* work, int32_t* out,
                             void* stream) noexcept;
extern "C" int fe_edt_signed_device(const int32_t* pair, int64_t n, double pitch, float* out, void* stream) noexcept;

namespace {

constexpr int MT = 256;  // queries per tile, threads per block
constexpr int MQ = 16;   // edge of a square tile of grid columns

// Flag bits (FE/mesh_fused_native_v1/field.cu:10, 29-30, 70-71): 1 some voxel
// outside, 2 some voxel inside, 4 a hit count above the capacity, 8 a
// non-finite depth; 16 a winding query on a mesh corner.
enum { MF_OUT = 1, MF_IN = 2, MF_COUNT = 4, MF_NONFINITE = 8, MF_VERTEX = 16 };

// Extra status of the mesh handles beyond 0/1/2/3/4: a winding-number query
// point coincides with a mesh corner (the reference raises ValueError there,
// FE/generalized_winding_ordered_v1.py:25).
constexpr int FE_VERTEX = 5;

// Signed-int64 keys ordered like the doubles they encode.
__device__ __forceinline__ long long order_key(double d)
{
    long long b = __double_as_longlong(d);
    return b >= 0 ? b : (b ^ 0x7FFFFFFFFFFFFFFFLL);
}
__device__ __forceinline__ double order_val(long long k)
{
    return __longlong_as_double(k >= 0 ? k : (k ^ 0x7FFFFFFFFFFFFFFFLL));
}

// Which query a thread of a tile owns.  gny == 0: tiles of MT consecutive
// queries; else the queries are the columns (i, j) of an gnx x gny grid at
// index i*gny + j and a tile is a 16x16 square of them.
struct TileMap {
    int64_t nq;
    int gnx, gny, tiles_y;
};

__device__ __forceinline__ int64_t tile_query(const TileMap& m, uint32_t tile, int t)
{
    if (m.gny == 0) {
        int64_t q = (int64_t)tile * MT + t;
        return q < m.nq ? q : -1;
    }
    const int ti = (int)(tile / (uint32_t)m.tiles_y), tj = (int)(tile % (uint32_t)m.tiles_y);
    const int i = ti * MQ + t / MQ, j = tj * MQ + t % MQ;
    return (i < m.gnx && j < m.gny) ? (int64_t)i * m.gny + j : -1;
}

static uint32_t tile_count(const TileMap& m)
{
    if (m.gny == 0) return (uint32_t)((m.nq + MT - 1) / MT);
    return (uint32_t)(((m.gnx + MQ - 1) / MQ) * m.tiles_y);
}

__device__ __forceinline__ double wave_min(double v)
{
    for (int o = 32; o; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
    for (int o = 32; o; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// Per-triangle xy bounds; a triangle whose projected doubled area does not
// clear the floor (FE/cuda_columns_v1/columns.cu:36-37) gets empty bounds and
// never enters a tile list.
__global__ void k_tri_boxes(const double* __restrict__ tris, int64_t ntri, double floor_area, double* __restrict__ boxes)
{
    const int64_t t = (int64_t)blockIdx.x * MT + threadIdx.x;
    if (t >= ntri) return;
    const double* p = tris + 9 * t;
    const double ax = p[0], ay = p[1], bx = p[3], by = p[4], cx = p[6], cy = p[7];
    const double d = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
    double* b = boxes + 4 * t;
    if (!(fabs(d) > floor_area)) {
        b[0] = INFINITY;
        b[1] = -INFINITY;
        b[2] = INFINITY;
        b[3] = -INFINITY;
        return;
    }
    b[0] = fmin(ax, fmin(bx, cx));
    b[1] = fmax(ax, fmax(bx, cx));
    b[2] = fmin(ay, fmin(by, cy));
    b[3] = fmax(ay, fmax(by, cy));
}

// xy bounds of each tile's queries.
__global__ __launch_bounds__(MT) void k_tile_bounds(const double* __restrict__ xy, TileMap m, uint32_t ntiles,
                                                     uint32_t grid, double* __restrict__ bounds)
{
    const uint32_t tile = xcd_tile(blockIdx.x, grid);
    if (tile >= ntiles) return;
    __shared__ double red[4][MT / FE_WAVE];
    const int64_t q = tile_query(m, tile, threadIdx.x);
    double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
    if (q >= 0) {
        x0 = x1 = xy[2 * q];
        y0 = y1 = xy[2 * q + 1];
    }
    x0 = wave_min(x0);
    x1 = wave_max(x1);
    y0 = wave_min(y0);
    y1 = wave_max(y1);
    const int wave = threadIdx.x / FE_WAVE;
    if ((threadIdx.x & (FE_WAVE - 1)) == 0) {
        red[0][wave] = x0;
        red[1][wave] = x1;
        red[2][wave] = y0;
        red[3][wave] = y1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < MT / FE_WAVE; ++w) {
            x0 = fmin(x0, red[0][w]);
            x1 = fmax(x1, red[1][w]);
            y0 = fmin(y0, red[2][w]);
            y1 = fmax(y1, red[3][w]);
        }
        bounds[4 * tile] = x0;
        bounds[4 * tile + 1] = x1;
        bounds[4 * tile + 2] = y0;
        bounds[4 * tile + 3] = y1;
    }
}

// Tile lists: FILL == false counts the triangles whose bounds meet the tile's
// bounds; FILL == true writes their indices, in face order, from offset[tile].
template <bool FILL>
__global__ __launch_bounds__(MT) void k_tile_cull(const double* __restrict__ bounds, const double* __restrict__ boxes,
                                                   int64_t ntri, uint32_t ntiles, uint32_t grid,
                                                   int32_t* __restrict__ count, const int64_t* __restrict__ offset,
                                                   int32_t* __restrict__ list)
{
    const uint32_t tile = xcd_tile(blockIdx.x, grid);
    if (tile >= ntiles) return;
    __shared__ int wsum[MT / FE_WAVE];
    const double bx0 = bounds[4 * tile], bx1 = bounds[4 * tile + 1];
    const double by0 = bounds[4 * tile + 2], by1 = bounds[4 * tile + 3];
    const int lane = threadIdx.x & (FE_WAVE - 1), wave = threadIdx.x / FE_WAVE;
    const int64_t base = FILL ? offset[tile] : 0;
    int64_t total = 0;
    for (int64_t t0 = 0; t0 < ntri; t0 += MT) {
        const int64_t t = t0 + threadIdx.x;
        bool hit = false;
        if (t < ntri) {
            const double* b = boxes + 4 * t;
            hit = !(b[1] < bx0 || b[0] > bx1 || b[3] < by0 || b[2] > by1);
        }
        const uint64_t bal = __ballot(hit);
        if (lane == 0) wsum[wave] = __popcll(bal);
        __syncthreads();
        int before = 0, all = 0;
        for (int w = 0; w < MT / FE_WAVE; ++w) {
            const int c = wsum[w];
            before += w < wave ? c : 0;
            all += c;
        }
        if (FILL && hit) list[base + total + before + __popcll(bal & ((1ull << lane) - 1))] = (int32_t)t;
        total += all;
        __syncthreads();
    }
    if (!FILL && threadIdx.x == 0) count[tile] = (int32_t)total;
}

// Column scan over the tile lists (FE/cuda_columns_v1/columns.cu:28-48).  One
// thread per query; the tile's candidate triangles are staged through LDS 256
// at a time and tested in list (= face) order.  The first `cap` hits of a
// query land at [q*cap + k]; counts keep the full number.  When `cells` is
// given, the min/max stored depth goes to cells[0]/cells[1] and flag bits to
// the int at cells + 2.
__global__ __launch_bounds__(MT) void k_scan(const double* __restrict__ tris, const int32_t* __restrict__ list,
                                              const int64_t* __restrict__ offset, const double* __restrict__ xy,
                                              TileMap m, uint32_t ntiles, uint32_t grid, int cap,
                                              int32_t* __restrict__ counts, double* __restrict__ depths,
                                              int32_t* __restrict__ signs, long long* __restrict__ cells)
{
    const uint32_t tile = xcd_tile(blockIdx.x, grid);
    if (tile >= ntiles) return;
    __shared__ double tri[9][MT];
    __shared__ double red[2][MT / FE_WAVE];
    __shared__ uint8_t wlist[MT / FE_WAVE][MT];
    const int64_t q = tile_query(m, tile, threadIdx.x);
    double qx = 0.0, qy = 0.0;
    if (q >= 0) {
        qx = xy[2 * q];
        qy = xy[2 * q + 1];
    }
    // this wave's query box: each wave keeps only the staged triangles that meet it
    const int wave = threadIdx.x / FE_WAVE, lane = threadIdx.x & (FE_WAVE - 1);
    const double wx0 = wave_min(q >= 0 ? qx : INFINITY), wx1 = wave_max(q >= 0 ? qx : -INFINITY);
    const double wy0 = wave_min(q >= 0 ? qy : INFINITY), wy1 = wave_max(q >= 0 ? qy : -INFINITY);
    const int64_t beg = offset[tile], end = offset[tile + 1];
    int found = 0;
    bool bad = false;
    double lo = INFINITY, hi = -INFINITY;
    for (int64_t k0 = beg; k0 < end; k0 += MT) {
        const int n = (int)(end - k0 < MT ? end - k0 : MT);
        __syncthreads();
        if ((int)threadIdx.x < n) {
            const double* p = tris + 9 * (int64_t)list[k0 + threadIdx.x];
#pragma unroll
            for (int c = 0; c < 9; ++c) tri[c][threadIdx.x] = p[c];
        }
        __syncthreads();
        int cnt = 0;  // the wave's candidates among the staged triangles, in list (= face) order
        for (int r = 0; r < MT / FE_WAVE; ++r) {
            const int k = r * FE_WAVE + lane;
            bool hit = false;
            if (k < n) {
                const double ax = tri[0][k], ay = tri[1][k], bx = tri[3][k], by = tri[4][k], cx = tri[6][k],
                             cy = tri[7][k];
                hit = !(fmax(ax, fmax(bx, cx)) < wx0 || fmin(ax, fmin(bx, cx)) > wx1 ||
                        fmax(ay, fmax(by, cy)) < wy0 || fmin(ay, fmin(by, cy)) > wy1);
            }
            const uint64_t m = __ballot(hit);
            if (hit) wlist[wave][cnt + __popcll(m & ((1ull << lane) - 1))] = (uint8_t)k;
            cnt += __popcll(m);
        }
        __syncthreads();
        if (q < 0) continue;
        for (int u = 0; u < cnt; ++u) {
            const int k = wlist[wave][u];
            const double ax = tri[0][k], ay = tri[1][k], bx = tri[3][k], by = tri[4][k], cx = tri[6][k], cy = tri[7][k];
            if (qx < fmin(ax, fmin(bx, cx)) || qx > fmax(ax, fmax(bx, cx)) || qy < fmin(ay, fmin(by, cy)) ||
                qy > fmax(ay, fmax(by, cy)))
                continue;
            const double d = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
            const double l0 = (bx - qx) * (cy - qy) - (by - qy) * (cx - qx);
            const double l1 = (cx - qx) * (ay - qy) - (cy - qy) * (ax - qx);
            const double l2 = (ax - qx) * (by - qy) - (ay - qy) * (bx - qx);
            const int s = d > 0 ? 1 : -1;
            if (l0 * s < 0 || l1 * s < 0 || l2 * s < 0) continue;
            const double z = (l0 * tri[2][k] + l1 * tri[5][k] + l2 * tri[8][k]) / d;
            if (found < cap) {
                depths[q * cap + found] = z;
                signs[q * cap + found] = s;
                if (isfinite(z)) {
                    lo = fmin(lo, z);
                    hi = fmax(hi, z);
                } else {
                    bad = true;
                }
            }
            ++found;
        }
    }
    if (q >= 0) counts[q] = found;
    if (!cells) return;
    const int f = (found > cap ? MF_COUNT : 0) | (bad ? MF_NONFINITE : 0);
    if (f) atomicOr(reinterpret_cast<int*>(cells + 2), f);
    lo = wave_min(lo);
    hi = wave_max(hi);
    if (lane == 0) {
        red[0][wave] = lo;
        red[1][wave] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < MT / FE_WAVE; ++w) {
            lo = fmin(lo, red[0][w]);
            hi = fmax(hi, red[1][w]);
        }
        if (lo <= hi) {
            atomicMin(cells, order_key(lo));
            atomicMax(cells + 1, order_key(hi));
        }
    }
}

// Clears the range cells and the flags (a kernel, so the graph holds no memset node).
__global__ void k_mesh_init(long long* cells)
{
    cells[0] = order_key(INFINITY);
    cells[1] = order_key(-INFINITY);
    cells[2] = 0;
}

// Keys in place of the depths and the per-z fractions
// (FE/mesh_fused_native_v1/field.cu:14-23).
__global__ void k_keys(const int32_t* __restrict__ counts, double* __restrict__ depths, int64_t ncol, int cap,
                       const long long* __restrict__ cells, const double* __restrict__ zs, int nz,
                       double* __restrict__ fractions)
{
    const int64_t i = (int64_t)blockIdx.x * MT + threadIdx.x;
    double lo = order_val(cells[0]), span = 1.0;
    const double hi = order_val(cells[1]);
    if (!isfinite(lo)) lo = 0.0;
    else span = fmax(hi - lo, 1e-12);
    if (i < nz) {
        const double f = .25 + (.5 * (zs[i] - lo)) / span;
        fractions[i] = fmax(0.0, fmin(.999, f));
    }
    if (i < ncol) {
        const int n = min(max(counts[i], 0), cap);
        const double base = (double)i + .25;
        double* k = depths + i * cap;
        for (int j = 0; j < n; ++j) k[j] = base + (.5 * (k[j] - lo)) / span;
    }
}

// ORs each thread's flag bits into *flags with at most one global atomic per
// block, and none once the bits are already there: one atomic per wave on a
// single word serialises at L2 (millions of waves on a large grid).  Every
// thread of the block must call it.
__device__ __forceinline__ void block_flags(int local, int* flags)
{
    __shared__ int acc;
    if (threadIdx.x == 0) acc = 0;
    __syncthreads();
    int w = 0;
    for (int bit = 1; bit <= MF_VERTEX; bit <<= 1)
        if (__ballot(local & bit)) w |= bit;
    if ((threadIdx.x & (FE_WAVE - 1)) == 0 && w) atomicOr(&acc, w);
    __syncthreads();
    if (threadIdx.x == 0 && acc) {
        const int seen = __hip_atomic_load(flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((seen & acc) != acc) atomicOr(flags, acc);
    }
}


// Sum of the signs of a column's stored hits whose key exceeds q
// (FE/mesh_fused_native_v1/field.cu:24-30).  col must be wave-uniform: the
// hit loads are then scalar loads, shared by the wave.
__device__ __forceinline__ int column_sum(const int32_t* __restrict__ counts, const int32_t* __restrict__ signs,
                                          const double* __restrict__ keys, int cap, int64_t col, double q)
{
    const int n = min(max(counts[col], 0), cap);
    const double* k = keys + col * cap;
    const int32_t* s = signs + col * cap;
    int sum = 0;
    for (int j = 0; j < n; ++j)
        if (k[j] > q) sum += s[j];
    return sum;
}

// A column's hits held one per lane (columns of at most 64 stored hits): the
// sum over them reads each key and sign with readlane, so the loop has no
// memory latency.
struct LaneHits {
    double key;
    int sign, n;
};

__device__ __forceinline__ LaneHits lane_hits(const int32_t* __restrict__ counts, const int32_t* __restrict__ signs,
                                              const double* __restrict__ keys, int cap, int64_t col, int lane)
{
    LaneHits h;
    h.n = min(max(counts[col], 0), cap);
    h.key = 0.0;
    h.sign = 0;
    if (lane < h.n && lane < FE_WAVE) {
        h.key = keys[col * cap + lane];
        h.sign = signs[col * cap + lane];
    }
    return h;
}

__device__ __forceinline__ int lane_sum(const LaneHits& h, double q)
{
    int sum = 0;
    for (int j = 0; j < h.n; ++j) {
        const double k = __longlong_as_double((long long)readlane64((uint64_t)__double_as_longlong(h.key), j));
        const int sg = __builtin_amdgcn_readlane(h.sign, j);
        sum += k > q ? sg : 0;
    }
    return sum;
}

// Mask and surface of the closed-mesh stage in one pass: one wave per column
// (x, y).  The column's mask (one voxel per lane, 64 z at a time) is written
// and kept in LDS; a solid voxel off every border is surface unless its z
// neighbours (LDS) and its four xy-neighbour columns (evaluated from their
// hits at this z) are all solid -- the 6-neighbour rule of
// FE/mesh_fused_native_v1/field.cu:31-36.  Dynamic LDS: 4 * nz bytes.
__global__ __launch_bounds__(MT) void k_mask_surface(const int32_t* __restrict__ counts,
                                                      const int32_t* __restrict__ signs,
                                                      const double* __restrict__ keys,
                                                      const double* __restrict__ fractions, int cap, int nx, int ny,
                                                      int nz, uint8_t* __restrict__ mask,
                                                      uint8_t* __restrict__ surface, int* flags)
{
    extern __shared__ uint8_t colmask[];
    const int wave = threadIdx.x / FE_WAVE, lane = threadIdx.x & (FE_WAVE - 1);
    const int64_t ncol = (int64_t)nx * ny;
    const int64_t col = (int64_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * (MT / FE_WAVE) + wave));
    const bool valid = col < ncol;
    uint8_t* own = colmask + wave * nz;
    int local = 0;
    if (valid) {
        const LaneHits h = lane_hits(counts, signs, keys, cap, col, lane);
        for (int z0 = 0; z0 < nz; z0 += FE_WAVE) {
            const int z = z0 + lane;
            const double q = (double)col + (z < nz ? fractions[z] : 0.0);
            const int sum = h.n <= FE_WAVE ? lane_sum(h, q) : column_sum(counts, signs, keys, cap, col, q);
            if (z < nz) {
                const bool b = sum != 0;
                own[z] = b;
                mask[col * nz + z] = b;
                local |= b ? MF_IN : MF_OUT;
            }
        }
    }
    __syncthreads();
    if (valid) {
        const int x = (int)(col / ny), y = (int)(col % ny);
        const bool inner = x > 0 && x + 1 < nx && y > 0 && y + 1 < ny;
        LaneHits nb[4];
        if (inner)
            for (int d = 0; d < 4; ++d)
                nb[d] = lane_hits(counts, signs, keys, cap, d == 0 ? col - ny : d == 1 ? col + ny : d == 2 ? col - 1
                                                                                                           : col + 1,
                                  lane);
        for (int z0 = 0; z0 < nz; z0 += FE_WAVE) {
            const int z = z0 + lane;
            bool sfc = false;
            bool need = false;
            double f = 0.0;
            if (z < nz) {
                sfc = own[z];
                need = sfc && inner && z > 0 && z + 1 < nz && own[z - 1] && own[z + 1];
                f = fractions[z];
            }
            bool all = need;
            if (__ballot(need))
                for (int d = 0; d < 4; ++d) {  // uniform loop: the neighbour column stays wave-uniform
                    const int64_t c = d == 0 ? col - ny : d == 1 ? col + ny : d == 2 ? col - 1 : col + 1;
                    const int sum = nb[d].n <= FE_WAVE ? lane_sum(nb[d], (double)c + f)
                                                       : column_sum(counts, signs, keys, cap, c, (double)c + f);
                    all = all && sum != 0;
                }
            if (need) sfc = !all;
            if (z < nz) surface[col * nz + z] = sfc;
        }
    }
    block_flags(local, flags);
}

// The same classification from sorted keys in CSR form
// (FE/column_mask_native_v1/mask.cu:5-10).
__global__ __launch_bounds__(MT) void k_classify_csr(const double* __restrict__ keys, const int32_t* __restrict__ signs,
                                                      const int32_t* __restrict__ offsets,
                                                      const double* __restrict__ fractions, int nz, uint32_t size,
                                                      uint8_t* __restrict__ mask)
{
    const uint32_t i = blockIdx.x * MT + threadIdx.x;
    if (i >= size) return;
    const uint32_t col = i / (uint32_t)nz;
    const double q = (double)col + fractions[i - col * (uint32_t)nz];
    int sum = 0;
    for (int j = offsets[col]; j < offsets[col + 1]; ++j)
        if (keys[j] > q) sum += signs[j];
    mask[i] = sum != 0;
}

// Solid voxels on the grid border or with a 6-neighbour outside
// (FE/mesh_fused_native_v1/field.cu:31-36; equal to solid & ~binary_erosion).
__global__ __launch_bounds__(MT) void k_surface(const uint8_t* __restrict__ mask, uint8_t* __restrict__ surface, int nx,
                                                 int ny, int nz)
{
    const uint32_t i = blockIdx.x * MT + threadIdx.x;
    const uint32_t size = (uint32_t)nx * ny * nz;
    if (i >= size) return;
    const uint32_t plane = (uint32_t)ny * nz;
    const uint32_t x = i / plane, r = i - x * plane, y = r / (uint32_t)nz, z = r - y * (uint32_t)nz;
    bool s = mask[i];
    if (s && x > 0 && x + 1 < (uint32_t)nx && y > 0 && y + 1 < (uint32_t)ny && z > 0 && z + 1 < (uint32_t)nz)
        s = !(mask[i - plane] && mask[i + plane] && mask[i - nz] && mask[i + nz] && mask[i - 1] && mask[i + 1]);
    surface[i] = s;
}

// Ordered generalized winding number (FE/generalized_winding_ordered_v1.py:11-32):
// per point, atan2 of the solid-angle numerator and denominator summed over
// triangles in face order, divided by 2*pi.  Triangles are staged through LDS
// 256 at a time.  mask = |w| > .5 (FE/mesh_field_cuda_columns_v1.py:69).
__global__ __launch_bounds__(MT) void k_gwn(const double* __restrict__ tris, int64_t ntri, const double* __restrict__ pts,
                                             uint32_t n, double* __restrict__ w_out, uint8_t* __restrict__ mask,
                                             int* flags)
{
    __shared__ double tri[9][MT];
    const uint32_t i = blockIdx.x * MT + threadIdx.x;
    const bool valid = i < n;
    double px = 0.0, py = 0.0, pz = 0.0;
    if (valid) {
        px = pts[3 * (int64_t)i];
        py = pts[3 * (int64_t)i + 1];
        pz = pts[3 * (int64_t)i + 2];
    }
    double total = 0.0;
    bool bad = false;
    for (int64_t t0 = 0; t0 < ntri; t0 += MT) {
        const int cnt = (int)(ntri - t0 < MT ? ntri - t0 : MT);
        __syncthreads();
        if ((int)threadIdx.x < cnt) {
            const double* p = tris + 9 * (t0 + threadIdx.x);
#pragma unroll
            for (int c = 0; c < 9; ++c) tri[c][threadIdx.x] = p[c];
        }
        __syncthreads();
        if (!valid) continue;
        for (int k = 0; k < cnt; ++k) {
            const double a0 = tri[0][k] - px, a1 = tri[1][k] - py, a2 = tri[2][k] - pz;
            const double b0 = tri[3][k] - px, b1 = tri[4][k] - py, b2 = tri[5][k] - pz;
            const double c0 = tri[6][k] - px, c1 = tri[7][k] - py, c2 = tri[8][k] - pz;
            const double la = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
            const double lb = sqrt(b0 * b0 + b1 * b1 + b2 * b2);
            const double lc = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
            bad |= la == 0 || lb == 0 || lc == 0;
            const double x0 = b1 * c2 - b2 * c1, x1 = b2 * c0 - b0 * c2, x2 = b0 * c1 - b1 * c0;
            const double num = a0 * x0 + a1 * x1 + a2 * x2;
            const double den = la * lb * lc + (a0 * b0 + a1 * b1 + a2 * b2) * lc + (b0 * c0 + b1 * c1 + b2 * c2) * la +
                               (c0 * a0 + c1 * a1 + c2 * a2) * lb;
            total += atan2(num, den);
        }
    }
    int local = bad ? MF_VERTEX : 0;
    if (valid) {
        const double w = total / (2.0 * 3.141592653589793);
        if (w_out) w_out[i] = w;
        const bool inside = fabs(w) > .5;
        if (mask) mask[i] = inside;
        local |= inside ? MF_IN : MF_OUT;
    }
    block_flags(local, flags);
}

// ------------------------------------------------------------------ host side

struct Geometry {
    double* tris = nullptr;   // ntri x 9
    double* boxes = nullptr;  // ntri x 4
    int64_t ntri = 0;
};

struct TileLists {
    double* bounds = nullptr;   // ntiles x 4
    int64_t* offset = nullptr;  // ntiles + 1
    int32_t* count = nullptr;   // ntiles
    int32_t* list = nullptr;
    int64_t list_cap = 0;
    uint32_t tiles_cap = 0;
};

template <class T>
void release(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

void release(Geometry& g)
{
    release(g.tris);
    release(g.boxes);
}

void release(TileLists& l)
{
    release(l.bounds);
    release(l.offset);
    release(l.count);
    release(l.list);
    l.list_cap = 0;
    l.tiles_cap = 0;
}

// Host validation of FE/cuda_columns_v1/columns.cu:52-54 and the area floor
// scale of :61-65.
bool valid_corners(const double* c, uint32_t ntri, double* floor_area)
{
    if (!c || !ntri || ntri > 1000000) return false;
    double scale = 1.0;
    for (size_t i = 0; i < (size_t)ntri * 9; ++i)
        if (!std::isfinite(c[i]) || std::abs(c[i]) > 1e12) return false;
    for (size_t t = 0; t < ntri; ++t)
        for (int j = 0; j < 3; ++j)
            scale = std::max({scale, std::abs(c[9 * t + 3 * j]), std::abs(c[9 * t + 3 * j + 1])});
    *floor_area = 1e-12 * scale * scale;
    return true;
}

hipError_t upload_geometry(Geometry& g, const double* corners, uint32_t ntri, double floor_area, hipStream_t st)
{
    hipError_t e;
    g.ntri = ntri;
    if ((e = hipMalloc(&g.tris, (size_t)ntri * 9 * sizeof(double))) != hipSuccess) return e;
    if ((e = hipMalloc(&g.boxes, (size_t)ntri * 4 * sizeof(double))) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(g.tris, corners, (size_t)ntri * 9 * sizeof(double), hipMemcpyHostToDevice, st)) !=
        hipSuccess)
        return e;
    hipLaunchKernelGGL(k_tri_boxes, dim3((unsigned)((ntri + MT - 1) / MT)), dim3(MT), 0, st, g.tris, (int64_t)ntri,
                       floor_area, g.boxes);
    return hipGetLastError();
}

// Builds the tile lists for the queries at device `xy` (synchronises `st`).
hipError_t build_lists(const Geometry& g, const double* xy, const TileMap& m, TileLists& l, hipStream_t st)
{
    hipError_t e;
    const uint32_t ntiles = tile_count(m), grid = round8(ntiles);
    if (ntiles > l.tiles_cap) {
        release(l.bounds);
        release(l.offset);
        release(l.count);
        if ((e = hipMalloc(&l.bounds, (size_t)ntiles * 4 * sizeof(double))) != hipSuccess) return e;
        if ((e = hipMalloc(&l.offset, ((size_t)ntiles + 1) * sizeof(int64_t))) != hipSuccess) return e;
        if ((e = hipMalloc(&l.count, (size_t)ntiles * sizeof(int32_t))) != hipSuccess) return e;
        l.tiles_cap = ntiles;
    }
    hipLaunchKernelGGL(k_tile_bounds, dim3(grid), dim3(MT), 0, st, xy, m, ntiles, grid, l.bounds);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_tile_cull<false>, dim3(grid), dim3(MT), 0, st, l.bounds, g.boxes, g.ntri, ntiles, grid, l.count,
                       nullptr, nullptr);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    std::vector<int32_t> count(ntiles);
    std::vector<int64_t> offset((size_t)ntiles + 1, 0);
    if ((e = hipMemcpyAsync(count.data(), l.count, (size_t)ntiles * 4, hipMemcpyDeviceToHost, st)) != hipSuccess)
        return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    for (uint32_t t = 0; t < ntiles; ++t) offset[t + 1] = offset[t] + count[t];
    const int64_t total = std::max<int64_t>(offset[ntiles], 1);
    if (total > l.list_cap) {
        release(l.list);
        if ((e = hipMalloc(&l.list, (size_t)total * sizeof(int32_t))) != hipSuccess) return e;
        l.list_cap = total;
    }
    if ((e = hipMemcpyAsync(l.offset, offset.data(), offset.size() * sizeof(int64_t), hipMemcpyHostToDevice, st)) !=
        hipSuccess)
        return e;
    hipLaunchKernelGGL(k_tile_cull<true>, dim3(grid), dim3(MT), 0, st, l.bounds, g.boxes, g.ntri, ntiles, grid,
                       nullptr, l.offset, l.list);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return hipStreamSynchronize(st);
}

hipError_t launch_scan(const Geometry& g, const TileLists& l, const double* xy, const TileMap& m, int cap,
                       int32_t* counts, double* depths, int32_t* signs, long long* cells, hipStream_t st)
{
    const uint32_t ntiles = tile_count(m), grid = round8(ntiles);
    hipLaunchKernelGGL(k_scan, dim3(grid), dim3(MT), 0, st, g.tris, l.list, l.offset, xy, m, ntiles, grid, cap, counts,
                       depths, signs, cells);
    return hipGetLastError();
}

// ----------------------------------------------------- FE/cuda_columns_v1 handle
struct Columns {
    Geometry geo;
    TileLists lists;
    uint32_t capacity = 0, hits = 0;
    double* xy = nullptr;
    int32_t* counts = nullptr;
    double* depths = nullptr;
    int32_t* signs = nullptr;
    hipStream_t stream = nullptr;
    std::thread::id owner = std::this_thread::get_id();
};

int columns_release(Columns* s)
{
    int status = FE_OK;
    if (s->stream && hipStreamSynchronize(s->stream) != hipSuccess) status = FE_RUNTIME;
    release(s->geo);
    release(s->lists);
    release(s->xy);
    release(s->counts);
    release(s->depths);
    release(s->signs);
    if (s->stream && hipStreamDestroy(s->stream) != hipSuccess) status = FE_RUNTIME;
    delete s;
    return status;
}

int columns_run(Columns& s, const double* xy, uint32_t count, int32_t* counts, double* depths, int32_t* signs,
                uint32_t* width, bool compact)
{
    if (s.owner != std::this_thread::get_id()) return 3;
    if (!xy || !counts || !depths || !signs || !count || count > s.capacity) return FE_INVALID;
    for (size_t i = 0; i < (size_t)count * 2; ++i)
        if (!std::isfinite(xy[i]) || std::abs(xy[i]) > 1e12) return FE_INVALID;
    hipStream_t st = s.stream;
    FE_TRY(hipMemcpyAsync(s.xy, xy, (size_t)count * 2 * sizeof(double), hipMemcpyHostToDevice, st));
    TileMap m{(int64_t)count, 0, 0, 1};
    FE_TRY(build_lists(s.geo, s.xy, m, s.lists, st));
    FE_TRY(launch_scan(s.geo, s.lists, s.xy, m, (int)s.hits, s.counts, s.depths, s.signs, nullptr, st));
    FE_TRY(hipMemcpyAsync(counts, s.counts, (size_t)count * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    FE_TRY(hipStreamSynchronize(st));
    uint32_t maximum = 0;
    for (uint32_t i = 0; i < count; ++i) {
        if (counts[i] > (int32_t)s.hits) return 4;
        maximum = std::max(maximum, (uint32_t)counts[i]);
    }
    const uint32_t columns = compact ? maximum : s.hits;
    if (columns) {
        FE_TRY(hipMemcpy2DAsync(depths, columns * sizeof(double), s.depths, s.hits * sizeof(double),
                                columns * sizeof(double), count, hipMemcpyDeviceToHost, st));
        FE_TRY(hipMemcpy2DAsync(signs, columns * sizeof(int32_t), s.signs, s.hits * sizeof(int32_t),
                                columns * sizeof(int32_t), count, hipMemcpyDeviceToHost, st));
        FE_TRY(hipStreamSynchronize(st));
    }
    if (width) *width = maximum;
    return FE_OK;
}

// ------------------------------------------------------------ mesh field handle
struct MeshField {
    bool gwn = false;
    int nx = 0, ny = 0, nz = 0, cap = 0;
    int64_t columns = 0, size = 0;
    double pitch = 0.0;
    std::thread::id owner = std::this_thread::get_id();
    hipStream_t stream = nullptr;
    Geometry geo;
    TileLists lists;
    TileMap map{};
    double* xy = nullptr;
    double* zs = nullptr;
    double* fractions = nullptr;
    int32_t* counts = nullptr;
    double* depths = nullptr;  // keys after the keys kernel
    int32_t* signs = nullptr;
    long long* cells = nullptr;  // [0] min depth key, [1] max depth key, [2] flags
    double* points = nullptr;
    uint8_t* mask = nullptr;
    uint8_t* surface = nullptr;
    int32_t* work = nullptr;
    int32_t* pair = nullptr;
    float* distance = nullptr;
    int* flags_host = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool failed = false;
};

int mesh_release(MeshField* f)
{
    int status = FE_OK;
    if (f->stream && hipStreamSynchronize(f->stream) != hipSuccess) status = FE_RUNTIME;
    if (f->exec && hipGraphExecDestroy(f->exec) != hipSuccess) status = FE_RUNTIME;
    if (f->graph && hipGraphDestroy(f->graph) != hipSuccess) status = FE_RUNTIME;
    release(f->geo);
    release(f->lists);
    release(f->xy);
    release(f->zs);
    release(f->fractions);
    release(f->counts);
    release(f->depths);
    release(f->signs);
    release(f->cells);
    release(f->points);
    release(f->mask);
    release(f->surface);
    release(f->work);
    release(f->pair);
    release(f->distance);
    if (f->flags_host && hipHostFree(f->flags_host) != hipSuccess) status = FE_RUNTIME;
    if (f->stream && hipStreamDestroy(f->stream) != hipSuccess) status = FE_RUNTIME;
    delete f;
    return status;
}

// The stage's work on stream st: clear, sign (scan + keys + classify, or
// winding numbers), surface, paired EDT, signed output, flags to the host.
int mesh_enqueue(MeshField& f, hipStream_t st)
{
    hipLaunchKernelGGL(k_mesh_init, dim3(1), dim3(1), 0, st, f.cells);
    FE_TRY(hipGetLastError());
    int* flags = reinterpret_cast<int*>(f.cells + 2);
    const uint32_t vox_blocks = (uint32_t)((f.size + MT - 1) / MT);
    if (!f.gwn) {
        FE_TRY(launch_scan(f.geo, f.lists, f.xy, f.map, f.cap, f.counts, f.depths, f.signs, f.cells, st));
        const int64_t n = std::max<int64_t>(f.columns, f.nz);
        hipLaunchKernelGGL(k_keys, dim3((unsigned)((n + MT - 1) / MT)), dim3(MT), 0, st, f.counts, f.depths, f.columns,
                           f.cap, f.cells, f.zs, f.nz, f.fractions);
        FE_TRY(hipGetLastError());
        const unsigned col_blocks = (unsigned)((f.columns + MT / FE_WAVE - 1) / (MT / FE_WAVE));
        hipLaunchKernelGGL(k_mask_surface, dim3(col_blocks), dim3(MT), (size_t)(MT / FE_WAVE) * f.nz, st, f.counts,
                           f.signs, f.depths, f.fractions, f.cap, f.nx, f.ny, f.nz, f.mask, f.surface, flags);
        FE_TRY(hipGetLastError());
    } else {
        hipLaunchKernelGGL(k_gwn, dim3(vox_blocks), dim3(MT), 0, st, f.geo.tris, f.geo.ntri, f.points,
                           (uint32_t)f.size, nullptr, f.mask, flags);
        FE_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_surface, dim3(vox_blocks), dim3(MT), 0, st, f.mask, f.surface, f.nx, f.ny, f.nz);
        FE_TRY(hipGetLastError());
    }
    if (fe_edt_device(f.mask, f.nx, f.ny, f.nz, 2, f.work, f.pair, st) != FE_OK) return FE_RUNTIME;
    if (fe_edt_signed_device(f.pair, f.size, f.pitch, f.distance, st) != FE_OK) return FE_RUNTIME;
    FE_TRY(hipMemcpyAsync(f.flags_host, flags, sizeof(int), hipMemcpyDeviceToHost, st));
    return FE_OK;
}

// Allocates the voxel buffers and captures the stage into a graph.
int mesh_finish(MeshField& f)
{
    const size_t n = (size_t)f.size;
    FE_TRY(hipMalloc(&f.cells, 4 * sizeof(long long)));
    FE_TRY(hipMalloc(&f.mask, n));
    FE_TRY(hipMalloc(&f.surface, n));
    FE_TRY(hipMalloc(&f.work, 2 * n * sizeof(int32_t)));
    FE_TRY(hipMalloc(&f.pair, 2 * n * sizeof(int32_t)));
    FE_TRY(hipMalloc(&f.distance, n * sizeof(float)));
    FE_TRY(hipHostMalloc(&f.flags_host, sizeof(int), hipHostMallocDefault));
    *f.flags_host = 0;
    FE_TRY(hipStreamSynchronize(f.stream));
    FE_TRY(hipStreamBeginCapture(f.stream, hipStreamCaptureModeThreadLocal));
    const int status = mesh_enqueue(f, f.stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(f.stream, &g);
    f.graph = g;
    if (status != FE_OK) return status;
    FE_TRY(e);
    FE_TRY(hipGraphInstantiate(&f.exec, f.graph, nullptr, nullptr, 0));
    return FE_OK;
}

// Validation of FE/mesh_fused_native_v1/field.cu:54-56 with caller-chosen caps.
bool valid_grid(const double* xy, const double* zs, int nx, int ny, int nz, double pitch, int axis_cap,
                int64_t size_cap, int64_t column_cap)
{
    if (!xy || !zs || nx < 1 || ny < 1 || nz < 1 || nx > axis_cap || ny > axis_cap || nz > axis_cap) return false;
    if ((int64_t)nx * ny * nz > size_cap || (int64_t)nx * ny > column_cap) return false;
    if (!std::isfinite(pitch) || pitch < 1e-12 || pitch > 1e12) return false;
    for (int64_t i = 0; i < 2 * (int64_t)nx * ny; ++i)
        if (!std::isfinite(xy[i]) || std::abs(xy[i]) > 1e12) return false;
    for (int i = 0; i < nz; ++i)
        if (!std::isfinite(zs[i])) return false;
    return true;
}

bool valid_anchor(const double* a)
{
    if (!a) return false;
    for (int i = 0; i < 3; ++i)
        if (!std::isfinite(a[i]) || std::abs(a[i]) > 1e12) return false;
    return true;
}

int mesh_create(const double* corners, uint32_t ntri, const double* anchor, const double* xy, int nx, int ny, int nz,
                double pitch, const double* zs, int cap, int axis_cap, int64_t size_cap, int64_t column_cap,
                void** handle)
{
    if (!handle || *handle) return FE_INVALID;
    if (!valid_grid(xy, zs, nx, ny, nz, pitch, axis_cap, size_cap, column_cap)) return FE_INVALID;
    double floor_area = 0.0;
    if (!valid_corners(corners, ntri, &floor_area) || !valid_anchor(anchor) || cap < 1 || cap > 256) return FE_INVALID;
    MeshField* f = new (std::nothrow) MeshField();
    if (!f) return FE_RUNTIME;
    f->nx = nx;
    f->ny = ny;
    f->nz = nz;
    f->cap = cap;
    f->pitch = pitch;
    f->columns = (int64_t)nx * ny;
    f->size = f->columns * nz;
    f->map = TileMap{f->columns, nx, ny, (ny + MQ - 1) / MQ};
    int status = FE_RUNTIME;
    do {
        if (hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking) != hipSuccess) break;
        if (upload_geometry(f->geo, corners, ntri, floor_area, f->stream) != hipSuccess) break;
        if (hipMalloc(&f->xy, (size_t)f->columns * 2 * sizeof(double)) != hipSuccess) break;
        if (hipMalloc(&f->zs, (size_t)nz * sizeof(double)) != hipSuccess) break;
        if (hipMalloc(&f->fractions, (size_t)nz * sizeof(double)) != hipSuccess) break;
        if (hipMalloc(&f->counts, (size_t)f->columns * sizeof(int32_t)) != hipSuccess) break;
        if (hipMalloc(&f->depths, (size_t)f->columns * cap * sizeof(double)) != hipSuccess) break;
        if (hipMalloc(&f->signs, (size_t)f->columns * cap * sizeof(int32_t)) != hipSuccess) break;
        if (hipMemcpyAsync(f->xy, xy, (size_t)f->columns * 2 * sizeof(double), hipMemcpyHostToDevice, f->stream) !=
            hipSuccess)
            break;
        if (hipMemcpyAsync(f->zs, zs, (size_t)nz * sizeof(double), hipMemcpyHostToDevice, f->stream) != hipSuccess)
            break;
        if (build_lists(f->geo, f->xy, f->map, f->lists, f->stream) != hipSuccess) break;
        status = mesh_finish(*f);
    } while (0);
    if (status != FE_OK) {
        fe_note_error(hipGetLastError(), "mesh field create");
        mesh_release(f);
        return status;
    }
    *handle = f;
    return FE_OK;
}

int mesh_query(void* handle, uint8_t* solid, uint8_t* surface, float* distance)
{
    if (!handle || !solid || !surface || !distance) return FE_INVALID;
    MeshField& f = *static_cast<MeshField*>(handle);
    if (f.owner != std::this_thread::get_id()) return 3;
    if (f.failed) return FE_RUNTIME;
    const size_t n = (size_t)f.size;
    hipStream_t st = f.stream;
    hipError_t e = hipGraphLaunch(f.exec, st);
    if (e == hipSuccess) e = hipMemcpyAsync(solid, f.mask, n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(surface, f.surface, n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(distance, f.distance, n * sizeof(float), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        fe_note_error(e, "mesh field query");
        (void)hipStreamSynchronize(st);
        f.failed = true;
        return FE_RUNTIME;
    }
    const int flags = *f.flags_host;
    if (flags & (MF_COUNT | MF_NONFINITE)) {
        f.failed = true;
        return 4;
    }
    if (flags & MF_VERTEX) {
        f.failed = true;
        return FE_VERTEX;
    }
    if ((flags & (MF_IN | MF_OUT)) != (MF_IN | MF_OUT)) {
        f.failed = true;
        return FE_INVALID;
    }
    return FE_OK;
}

int mesh_destroy(void** handle)
{
    if (!handle) return FE_INVALID;
    if (!*handle) return FE_OK;
    MeshField* f = static_cast<MeshField*>(*handle);
    if (f->owner != std::this_thread::get_id()) return 3;
    const int status = mesh_release(f);
    *handle = nullptr;
    return status;
}

}  // namespace

// ------------------------------------------------------------------ C ABI

// FE/cuda_columns_v1/columns.cu:50-79.  backend: "cuda-scan-v1" (the name the
// reference's callers pass) or "hip-scan-v1".  Status 0 ok, 1 invalid, 2 runtime.
extern "C" int columns_create(const char* backend, const double* input, uint32_t triangles, uint32_t capacity,
                              uint32_t hits, const double* anchor, void** handle) noexcept
{
    if (!handle || *handle || !backend || (std::strcmp(backend, "cuda-scan-v1") && std::strcmp(backend, "hip-scan-v1")))
        return FE_INVALID;
    double floor_area = 0.0;
    if (!input || !valid_anchor(anchor) || !capacity || capacity > 262144 || !hits || hits > 256) return FE_INVALID;
    if (!valid_corners(input, triangles, &floor_area)) return FE_INVALID;
    Columns* s = new (std::nothrow) Columns();
    if (!s) return FE_RUNTIME;
    s->capacity = capacity;
    s->hits = hits;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = upload_geometry(s->geo, input, triangles, floor_area, s->stream);
    if (e == hipSuccess) e = hipMalloc(&s->xy, (size_t)capacity * 2 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&s->counts, (size_t)capacity * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&s->depths, (size_t)capacity * hits * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&s->signs, (size_t)capacity * hits * sizeof(int32_t));
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        fe_note_error(e, "columns_create");
        columns_release(s);
        return FE_RUNTIME;
    }
    *handle = s;
    return FE_OK;
}

// FE/cuda_columns_v1/columns.cu:81-105: hits per query in face order, depths
// and signs [count][hits]; status 3 from a foreign thread, 4 when a query has
// more than `hits` hits (counts are still written).
extern "C" int columns_query(void* handle, const double* xy, uint32_t count, int32_t* counts, double* depths,
                             int32_t* signs) noexcept
{
    if (!handle) return FE_INVALID;
    try {
        return columns_run(*static_cast<Columns*>(handle), xy, count, counts, depths, signs, nullptr, false);
    } catch (...) {
        return FE_RUNTIME;
    }
}

// FE/cuda_columns_v1/columns.cu:106-108: as columns_query with rows of
// *width = the largest count.
extern "C" int columns_query_compact(void* handle, const double* xy, uint32_t count, int32_t* counts, double* depths,
                                     int32_t* signs, uint32_t* width) noexcept
{
    if (!handle || !width) return FE_INVALID;
    try {
        return columns_run(*static_cast<Columns*>(handle), xy, count, counts, depths, signs, width, true);
    } catch (...) {
        return FE_RUNTIME;
    }
}

// FE/cuda_columns_v1/columns.cu:109-113.
extern "C" int columns_destroy(void** handle) noexcept
{
    if (!handle) return FE_INVALID;
    if (!*handle) return FE_OK;
    Columns* s = static_cast<Columns*>(*handle);
    if (s->owner != std::this_thread::get_id()) return 3;
    const int status = columns_release(s);
    *handle = nullptr;
    return status;
}

// FE/column_mask_native_v1/mask.cu:17-37: mask and surface from sorted
// composite keys in CSR form, with the reference's validation and caps
// (<= 512 per axis, <= 1048576 voxels, <= 1000000 hits).
extern "C" int column_mask_surface(const double* keys, const int32_t* signs, int hits, const int32_t* offsets,
                                   const double* fractions, int nx, int ny, int nz, uint8_t* mask,
                                   uint8_t* surface) noexcept
{
    if (!keys || !signs || !offsets || !fractions || !mask || !surface || hits < 0 || hits > 1000000 || nx < 1 ||
        ny < 1 || nz < 1 || nx > 512 || ny > 512 || nz > 512 || (int64_t)nx * ny * nz > 1048576)
        return FE_INVALID;
    const int columns = nx * ny, size = columns * nz;
    if (offsets[0] != 0 || offsets[columns] != hits) return FE_INVALID;
    for (int i = 0; i < columns; ++i)
        if (offsets[i] < 0 || offsets[i] > offsets[i + 1] || offsets[i + 1] > hits) return FE_INVALID;
    for (int i = 0; i < hits; ++i)
        if (!std::isfinite(keys[i]) || (signs[i] != 1 && signs[i] != -1)) return FE_INVALID;
    for (int i = 0; i < nz; ++i)
        if (!std::isfinite(fractions[i]) || fractions[i] < 0 || fractions[i] > .999) return FE_INVALID;
    double *dk = nullptr, *df = nullptr;
    int32_t *ds = nullptr, *di = nullptr;
    uint8_t *dm = nullptr, *db = nullptr;
    int status = FE_RUNTIME;
    hipError_t e = hipSuccess;
    do {
        if ((e = hipMalloc(&dk, (size_t)std::max(hits, 1) * sizeof(double))) != hipSuccess) break;
        if ((e = hipMalloc(&ds, (size_t)std::max(hits, 1) * sizeof(int32_t))) != hipSuccess) break;
        if ((e = hipMalloc(&di, ((size_t)columns + 1) * sizeof(int32_t))) != hipSuccess) break;
        if ((e = hipMalloc(&df, (size_t)nz * sizeof(double))) != hipSuccess) break;
        if ((e = hipMalloc(&dm, (size_t)size)) != hipSuccess) break;
        if ((e = hipMalloc(&db, (size_t)size)) != hipSuccess) break;
        if (hits) {
            if ((e = hipMemcpy(dk, keys, (size_t)hits * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess) break;
            if ((e = hipMemcpy(ds, signs, (size_t)hits * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess) break;
        }
        if ((e = hipMemcpy(di, offsets, ((size_t)columns + 1) * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess)
            break;
        if ((e = hipMemcpy(df, fractions, (size_t)nz * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess) break;
        const unsigned blocks = (unsigned)((size + MT - 1) / MT);
        hipLaunchKernelGGL(k_classify_csr, dim3(blocks), dim3(MT), 0, nullptr, dk, ds, di, df, nz, (uint32_t)size, dm);
        if ((e = hipGetLastError()) != hipSuccess) break;
        hipLaunchKernelGGL(k_surface, dim3(blocks), dim3(MT), 0, nullptr, dm, db, nx, ny, nz);
        if ((e = hipGetLastError()) != hipSuccess) break;
        if ((e = hipMemcpy(mask, dm, (size_t)size, hipMemcpyDeviceToHost)) != hipSuccess) break;
        if ((e = hipMemcpy(surface, db, (size_t)size, hipMemcpyDeviceToHost)) != hipSuccess) break;
        status = FE_OK;
    } while (0);
    if (e != hipSuccess) fe_note_error(e, "column_mask_surface");
    for (void* p : {(void*)dk, (void*)ds, (void*)di, (void*)df, (void*)dm, (void*)db})
        if (p && hipFree(p) != hipSuccess) status = FE_RUNTIME;
    return status;
}

// FE/mesh_fused_native_v1/field.cu:53-74 / FE/mesh_fused_graph_v1/field.cu:31-63:
// prepare an immutable mesh and grid (corners in world coordinates, xy one
// query per grid column, zs the sample z of each layer) and capture the stage.
// `path` names the reference's ray-tracing program and is not used: the hits
// come from the column scan.  Reference caps: <= 512 per axis, <= 1048576
// voxels, <= 262144 columns; 64 hits per column.
extern "C" int fused_create(const char* path, const double* corners, uint32_t triangles, const double* anchor,
                            const double* xy, int nx, int ny, int nz, double pitch, const double* zs,
                            void** handle) noexcept
{
    (void)path;
    try {
        return mesh_create(corners, triangles, anchor, xy, nx, ny, nz, pitch, zs, 64, 512, 1048576, 262144, handle);
    } catch (...) {
        return FE_RUNTIME;
    }
}

// FE/mesh_fused_graph_v1/field.cu:64-79: replay and copy solid, surface and
// the signed distance out.  Status 4 (hit overflow or a non-finite depth) and
// 1 (the mask lacks a class) latch the handle, as does 2.
extern "C" int fused_query(void* handle, uint8_t* solid, uint8_t* surface, float* distance) noexcept
{
    try {
        return mesh_query(handle, solid, surface, distance);
    } catch (...) {
        return FE_RUNTIME;
    }
}

extern "C" int fused_destroy(void** handle) noexcept
{
    try {
        return mesh_destroy(handle);
    } catch (...) {
        return FE_RUNTIME;
    }
}

// fused_create without the reference's size caps: axes <= 4096, fewer than
// 2^31 voxels, 1..256 hits per column.
extern "C" int fe_mesh_create(const double* corners, uint32_t triangles, const double* anchor, const double* xy, int nx,
                              int ny, int nz, double pitch, const double* zs, int hit_capacity, void** handle) noexcept
{
    try {
        return mesh_create(corners, triangles, anchor, xy, nx, ny, nz, pitch, zs, hit_capacity, 4096,
                           (int64_t)INT32_MAX, (int64_t)INT32_MAX, handle);
    } catch (...) {
        return FE_RUNTIME;
    }
}

// Open-mesh variant: the mask is |generalized winding number| > .5 at the
// given points (nx*ny*nz x 3, in the frame of `corners`), then the same
// surface, paired EDT and signed output.  Query status 5: a point lies on a
// mesh corner.
extern "C" int fe_gwn_create(const double* corners, uint32_t triangles, const double* points, int nx, int ny, int nz,
                             double pitch, void** handle) noexcept
{
    try {
        if (!handle || *handle || !points || nx < 1 || ny < 1 || nz < 1 || nx > 4096 || ny > 4096 || nz > 4096 ||
            (int64_t)nx * ny * nz > INT32_MAX || !std::isfinite(pitch) || pitch < 1e-12 || pitch > 1e12)
            return FE_INVALID;
        double floor_area = 0.0;
        if (!valid_corners(corners, triangles, &floor_area)) return FE_INVALID;
        const int64_t n = (int64_t)nx * ny * nz;
        for (int64_t i = 0; i < 3 * n; ++i)
            if (!std::isfinite(points[i])) return FE_INVALID;
        MeshField* f = new (std::nothrow) MeshField();
        if (!f) return FE_RUNTIME;
        f->gwn = true;
        f->nx = nx;
        f->ny = ny;
        f->nz = nz;
        f->pitch = pitch;
        f->columns = (int64_t)nx * ny;
        f->size = n;
        int status = FE_RUNTIME;
        do {
            if (hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking) != hipSuccess) break;
            if (upload_geometry(f->geo, corners, triangles, floor_area, f->stream) != hipSuccess) break;
            if (hipMalloc(&f->points, (size_t)n * 3 * sizeof(double)) != hipSuccess) break;
            if (hipMemcpyAsync(f->points, points, (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice, f->stream) !=
                hipSuccess)
                break;
            status = mesh_finish(*f);
        } while (0);
        if (status != FE_OK) {
            fe_note_error(hipGetLastError(), "winding field create");
            mesh_release(f);
            return status;
        }
        *handle = f;
        return FE_OK;
    } catch (...) {
        return FE_RUNTIME;
    }
}

extern "C" int fe_mesh_query(void* handle, uint8_t* solid, uint8_t* surface, float* distance) noexcept
{
    return fused_query(handle, solid, surface, distance);
}

extern "C" int fe_mesh_destroy(void** handle) noexcept { return fused_destroy(handle); }

// Replays the captured stage on `stream` (null: the default stream) without
// copying anything out; read the results through fe_mesh_buffers and the
// status through fe_mesh_status after the stream has finished.
extern "C" int fe_mesh_run(void* handle, void* stream) noexcept
{
    if (!handle) return FE_INVALID;
    MeshField& f = *static_cast<MeshField*>(handle);
    if (f.owner != std::this_thread::get_id()) return 3;
    if (f.failed) return FE_RUNTIME;
    FE_TRY(hipGraphLaunch(f.exec, static_cast<hipStream_t>(stream)));
    return FE_OK;
}

// Status of the last completed run, as fused_query reports it (0, 1, 4, 5).
extern "C" int fe_mesh_status(const void* handle) noexcept
{
    if (!handle) return FE_INVALID;
    const int flags = *static_cast<const MeshField*>(handle)->flags_host;
    if (flags & (MF_COUNT | MF_NONFINITE)) return 4;
    if (flags & MF_VERTEX) return FE_VERTEX;
    return (flags & (MF_IN | MF_OUT)) == (MF_IN | MF_OUT) ? FE_OK : FE_INVALID;
}

// Device buffers of a mesh handle: solid and surface (nx*ny*nz bytes), signed
// distance (float32), the squared-distance pair (2 x nx*ny*nz int32), and the
// column hits (counts, and keys / signs at [column*cap + k]).
extern "C" int fe_mesh_buffers(const void* handle, void** solid, void** surface, void** distance, void** pair,
                               void** counts, void** keys, void** signs, int* cap) noexcept
{
    if (!handle) return FE_INVALID;
    const MeshField& f = *static_cast<const MeshField*>(handle);
    if (solid) *solid = f.mask;
    if (surface) *surface = f.surface;
    if (distance) *distance = f.distance;
    if (pair) *pair = f.pair;
    if (counts) *counts = f.counts;
    if (keys) *keys = f.depths;
    if (signs) *signs = f.signs;
    if (cap) *cap = f.cap;
    return FE_OK;
}

// Winding numbers of device points (n x 3) against device corners (ntri x 9);
// w and/or mask (|w| > .5) may be null.  flags: device int receiving bits
// 1/2 (classes seen) and 16 (a point on a corner); the caller clears it.
extern "C" int fe_gwn_device(const double* corners, int64_t ntri, const double* points, int64_t n, double* w,
                             uint8_t* mask, int* flags, void* stream) noexcept
{
    if (!corners || ntri < 1 || !points || n < 1 || n > INT32_MAX || !flags) return FE_INVALID;
    hipLaunchKernelGGL(k_gwn, dim3((unsigned)((n + MT - 1) / MT)), dim3(MT), 0, static_cast<hipStream_t>(stream),
                       corners, ntri, points, (uint32_t)n, w, mask, flags);
    FE_TRY(hipGetLastError());
    return FE_OK;
}
