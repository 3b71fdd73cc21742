// Per-block-pitch (two-level) fields for gfx950.
//
// Reference semantics (FE/faltkarna_v1_multires.py): the brick lattice is the
// block lattice of a coarse field (bricks of B^3 coarse voxels, world side
// L = B*h).  An active brick at level l in {0, 1, 2} stores (B*2^l + 1)^3
// samples at pitch h/2^l over its closed world cube (the "+1" is the one-
// sample overlap band, :84-106, :291-303); inactive bricks store nothing and
// answer +-margin.  The face layer of a brick next to a coarser active brick
// is overwritten by the coarse brick's field interpolated in its own face
// plane (:211-252); queries interpolate in the point's brick at that brick's
// pitch (:405-441).
//
// Here the samples come from a recipe (the fp32 fold of field.hip) instead of
// a mesh: level 0 at the coarse lattice points x0 + f32(i)*h with indices
// past the grid clamped to its last sample (the reference pads its coarse
// field by edge replication), level l at x0 + f32(i)*(h/2^l) on the fine
// lattice, where shared points of two levels coincide bit for bit.  The
// interpolation is scipy.ndimage.map_coordinates(order=1, mode="nearest"):
// each coordinate clamped to [0, n-1], w0 = 1 - t, w1 = 1 - w0, the eight
// products value*wx*wy*wz summed in f64 with the last axis fastest, the sum
// rounded to fp32.
//
// MI355X design: one launch evaluates every brick's tile (a wave per slab of
// about 1024 samples of a brick, the brick found by binary search, the
// culled list of the brick's superblock from the coarse classification
// refined exactly for the slab's box, primitive records read wave-uniformly);
// one launch per refined level writes the bands
// (a workgroup per fine brick, faces in the reference's order); one launch
// answers a batch of queries (a thread per point).
#include "recipe_eval.hpp"
#include <climits>
#include <cmath>

struct MrBrick {          // one active brick
    int32_t bi, bj, bk, level;
    int64_t tile;         // element offset of its (B*2^level + 1)^3 samples in the tile arena
    int64_t item0;        // its first slab in the tile launch (mr_slab_rows rows of samples each)
};
static_assert(sizeof(MrBrick) == 32, "brick record");

struct MrGrid {           // the brick lattice, in f64 (the fp32 origin and pitch widened)
    double lox, loy, loz, h, margin;
    int32_t B, nbx, nby, nbz;
    int32_t nx, ny, nz, pad;
};
static_assert(sizeof(MrGrid) == 72, "grid record");

namespace {

__device__ __forceinline__ int find_brick(const MrBrick* __restrict__ bricks, int n, int64_t item)
{
    int lo = 0, hi = n - 1;  // last brick whose first slab is <= item
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bricks[mid].item0 <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Rows (along the brick's first axis) of one slab of a brick with m samples
// per axis: about 1024 samples per slab.  The host counts slabs by the same
// rule (multires.py, _SLAB_SAMPLES).
__host__ __device__ __forceinline__ int mr_slab_rows(int m) { return max(1, 1024 / (m * m)); }

// Every active brick's samples, one wave per slab.  The wave refines the
// culled list of the brick's superblock (the whole recipe when none was
// stored) for the slab's box with the exact rule of cull_wave, so the dropped
// primitives change no result bit, then folds the refined list at the slab's
// samples, two per lane per pass.  Sample coordinates: level 0 at
// x0 + f32(min(b*B + i, nx - 1)) * h, level l at x0 + f32(b*B*2^l + i) * (h/2^l);
// both are monotone in i, so the slab's extreme samples bound its box.
__global__ __launch_bounds__(256) void k_mr_tiles(const FieldDesc* __restrict__ desc, const MrBrick* __restrict__ bricks,
                                                  int nbricks, int64_t nitems, float* __restrict__ tiles)
{
    __shared__ uint16_t wlist[4][FE_MAX_PRIMS / 2];
    const FieldDesc& D = *desc;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t nw = (int64_t)gridDim.x * 4;
    uint16_t* wl = wlist[wv];
    for (int64_t it = (int64_t)blockIdx.x * 4 + wv; it < nitems; it += nw) {
        const int b = __builtin_amdgcn_readfirstlane(find_brick(bricks, nbricks, it));
        const MrBrick Bk = bricks[b];
        const int n = D.B << Bk.level, m = n + 1, mm = m * m;
        const int i0 = (int)(it - Bk.item0) * mr_slab_rows(m), i1 = min(m, i0 + mr_slab_rows(m));
        const float hl = ldexpf(D.pitch, -Bk.level);  // exact: a power-of-two scaling
        auto coord = [&](float o, int bb, int q, int nmax) {
            return Bk.level == 0 ? o + (float)min(bb * D.B + q, nmax - 1) * D.pitch : o + (float)(bb * n + q) * hl;
        };
        const float lx = coord(D.x0, Bk.bi, i0, D.nx), ux = coord(D.x0, Bk.bi, i1 - 1, D.nx);
        const float ly = coord(D.y0, Bk.bj, 0, D.ny), uy = coord(D.y0, Bk.bj, m - 1, D.ny);
        const float lz = coord(D.z0, Bk.bk, 0, D.nz), uz = coord(D.z0, Bk.bk, m - 1, D.nz);
        const float cx = 0.5f * (lx + ux), cy = 0.5f * (ly + uy), cz = 0.5f * (lz + uz);
        const float ex = ux - lx, ey = uy - ly, ez = uz - lz;
        const float rad = 0.5f * sqrtf(ex * ex + ey * ey + ez * ez) * 1.0001f;
        const int sb = ((Bk.bi / D.SB) * D.nsby + Bk.bj / D.SB) * D.nsbz + Bk.bk / D.SB;
        const int nl = D.list_len[sb];
        const uint16_t* src = nl >= 0 ? D.lists + (int64_t)sb * D.cap : nullptr;
        const int nsrc = nl >= 0 ? nl : D.nprims;
        const uint16_t* list = src;
        int np = nsrc;
        if (nsrc <= FE_MAX_PRIMS / 2) {
            np = cull_wave(D.prims, src, nsrc, cx, cy, cz, rad, D.slack, wl);
            // a second pass over the refined list (in place): its runs join
            // across dropped steps and are no longer cut every 64 entries
            if (np > 8) np = cull_wave(D.prims, wl, np, cx, cy, cz, rad, D.slack, wl);
            list = wl;
        }
        const int S = (i1 - i0) * mm;
        float* out = tiles + Bk.tile + (int64_t)i0 * mm;
        for (int v0 = 0; v0 < S; v0 += 128) {
            float x[2], y[2], z[2], f[2];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int s = min(v0 + k * 64 + lane, S - 1);
                const int i = i0 + s / mm, j = (s / m) % m, q = s % m;
                x[k] = coord(D.x0, Bk.bi, i, D.nx);
                y[k] = coord(D.y0, Bk.bj, j, D.ny);
                z[k] = coord(D.z0, Bk.bk, q, D.nz);
            }
            eval_points<2>(D.prims, list, np, x, y, z, f);
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int s = v0 + k * 64 + lane;
                if (s < S) out[s] = f[k];
            }
        }
    }
}

// map_coordinates(order=1, mode="nearest") of an m^3 fp32 tile at local
// (f64) coordinates c, rounded to fp32.
__device__ __forceinline__ float spline1(const float* __restrict__ t, int m, const double* c)
{
    int lower[3];
    double w[3][2];
    for (int a = 0; a < 3; ++a) {
        double x = c[a];
        const double hi = (double)(m - 1);
        if (x < 0.0) x = 0.0;
        if (x > hi) x = hi;
        const double fl = floor(x);
        lower[a] = (int)fl;
        double tt = x - fl;
        tt -= floor(tt);
        w[a][0] = 1.0 - tt;
        w[a][1] = 1.0 - w[a][0];
    }
    double acc = 0.0;
    for (int h = 0; h < 8; ++h) {
        const int bx = (h >> 2) & 1, by = (h >> 1) & 1, bz = h & 1;
        const int qx = min(lower[0] + bx, m - 1), qy = min(lower[1] + by, m - 1), qz = min(lower[2] + bz, m - 1);
        double coeff = (double)t[((int64_t)qx * m + qy) * m + qz];
        coeff *= 1.0;
        coeff *= w[0][bx];
        coeff *= w[1][by];
        coeff *= w[2][bz];
        acc += coeff;
    }
    return (float)acc;
}

// The overlap band of every brick at `level` (faces in the reference's order:
// axis 0..2, the lower side first; a later face overwrites shared edges).
// Coarser neighbours are not written by this launch.
__global__ __launch_bounds__(256) void k_mr_band(MrGrid G, const MrBrick* __restrict__ bricks,
                                                 const int32_t* __restrict__ slot, const int32_t* __restrict__ fine,
                                                 int nfine, float* __restrict__ tiles)
{
    const double L = (double)G.B * G.h;
    const double lo[3] = {G.lox, G.loy, G.loz};
    const int nb[3] = {G.nbx, G.nby, G.nbz};
    for (int f = blockIdx.x; f < nfine; f += gridDim.x) {
        const MrBrick Bk = bricks[fine[f]];
        const int n = G.B << Bk.level, m = n + 1;
        const double pitch = G.h / (double)(1 << Bk.level);
        const int b[3] = {Bk.bi, Bk.bj, Bk.bk};
        double borg[3];
        for (int a = 0; a < 3; ++a) borg[a] = lo[a] + (double)b[a] * L;
        float* own = tiles + Bk.tile;
        for (int ax = 0; ax < 3; ++ax) {
            for (int side = -1; side <= 1; side += 2) {
                int q[3] = {b[0], b[1], b[2]};
                q[ax] += side;
                if (q[ax] < 0 || q[ax] >= nb[ax]) continue;
                const int s = slot[((int64_t)q[0] * G.nby + q[1]) * G.nbz + q[2]];
                if (s < 0) continue;
                const MrBrick Nk = bricks[s];
                if (Nk.level >= Bk.level) continue;
                const int nn = G.B << Nk.level, mn = nn + 1;
                const double pn = G.h / (double)(1 << Nk.level);
                double norg[3];
                for (int a = 0; a < 3; ++a) norg[a] = lo[a] + (double)q[a] * L;
                const int u = ax == 0 ? 1 : 0, v = ax == 2 ? 1 : 2;  // the face's own axes, ascending
                const int fix = side < 0 ? 0 : n;
                for (int e = threadIdx.x; e < m * m; e += 256) {
                    int idx[3];
                    idx[ax] = fix;
                    idx[u] = e / m;
                    idx[v] = e % m;
                    double c[3];
                    for (int a = 0; a < 3; ++a) {
                        const double p = (double)idx[a] * pitch + borg[a];
                        c[a] = (p - norg[a]) / pn;
                    }
                    own[((int64_t)idx[0] * m + idx[1]) * m + idx[2]] = spline1(tiles + Nk.tile, mn, c);
                }
                __syncthreads();
            }
        }
    }
}

// Signed distance at f64 points: the brick by floor((p - lo) / L), clipped to
// the lattice, a point exactly on a face of an inactive brick moved to the
// active brick below it (axes in order); outside the lattice +margin, an
// inactive brick -margin inside / +margin outside, an active brick the
// interpolated tile at its pitch (fp32, widened).  brick_out (optional):
// the (clipped) brick (3 int32) and its level (-1 inactive) per point; a
// point outside the lattice reports the level of the brick it clipped to, as
// the gradient's step takes it from there.
__global__ __launch_bounds__(256) void k_mr_query(MrGrid G, const MrBrick* __restrict__ bricks,
                                                  const int32_t* __restrict__ slot, const int32_t* __restrict__ kind,
                                                  const float* __restrict__ tiles, const double* __restrict__ pts,
                                                  int64_t n, double* __restrict__ out, int32_t* __restrict__ brick_out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double L = (double)G.B * G.h;
    const double lo[3] = {G.lox, G.loy, G.loz};
    const int nb[3] = {G.nbx, G.nby, G.nbz};
    double p[3], rel[3];
    int64_t b[3];
    bool outside = false;
    for (int a = 0; a < 3; ++a) {
        p[a] = pts[3 * i + a];
        rel[a] = (p[a] - lo[a]) / L;
        b[a] = (int64_t)floor(rel[a]);
        outside = outside || b[a] < 0 || b[a] >= nb[a];
    }
    for (int a = 0; a < 3; ++a) b[a] = b[a] < 0 ? 0 : (b[a] >= nb[a] ? nb[a] - 1 : b[a]);
    auto kind_at = [&](const int64_t* q) { return kind[(q[0] * G.nby + q[1]) * G.nbz + q[2]]; };
    bool inactive = kind_at(b) != 0;
    for (int a = 0; a < 3; ++a) {
        if (rel[a] == floor(rel[a]) && b[a] > 0 && inactive && !outside) {
            int64_t c[3] = {b[0], b[1], b[2]};
            c[a] -= 1;
            if (kind_at(c) == 0) {
                b[a] = c[a];
                inactive = kind_at(b) != 0;
            }
        }
    }
    const int kd = kind_at(b);
    double r = kd < 0 ? -G.margin : G.margin;
    const int level = kd == 0 ? bricks[slot[(b[0] * G.nby + b[1]) * G.nbz + b[2]]].level : -1;
    if (outside) r = G.margin;
    else if (kd == 0) {
        const MrBrick Bk = bricks[slot[(b[0] * G.nby + b[1]) * G.nbz + b[2]]];
        const double pitch = G.h / (double)(1 << Bk.level);
        double c[3];
        for (int a = 0; a < 3; ++a) {
            const double o = lo[a] + (double)b[a] * L;
            c[a] = (p[a] - o) / pitch;
        }
        r = (double)spline1(tiles + Bk.tile, (G.B << Bk.level) + 1, c);
    }
    out[i] = r;
    if (brick_out) {
        brick_out[4 * i] = (int32_t)b[0];
        brick_out[4 * i + 1] = (int32_t)b[1];
        brick_out[4 * i + 2] = (int32_t)b[2];
        brick_out[4 * i + 3] = level;
    }
}

}  // namespace

// ---------------------------------------------------------------- C ABI
// Tiles of `nbricks` active bricks (device table, item0 ascending from 0,
// nitems = total slabs, mr_slab_rows rows of a brick each) from the recipe of
// one classified field descriptor (device; its culled lists as
// fe_field_classify left them).
extern "C" int fe_multires_tiles(const void* desc, const void* bricks, int nbricks, int64_t nitems, float* tiles,
                                 void* stream) noexcept
{
    if (!desc || !bricks || nbricks < 1 || nitems < 1 || !tiles) return FE_INVALID;
    const int64_t blocks = std::min<int64_t>((nitems + 3) / 4, 65536);
    hipLaunchKernelGGL(k_mr_tiles, dim3((unsigned)blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const FieldDesc*>(desc), static_cast<const MrBrick*>(bricks), nbricks, nitems,
                       tiles);
    FE_TRY(hipGetLastError());
    return FE_OK;
}

// Overlap bands of the `nfine` bricks listed in `fine` (device indices into
// the brick table, all at one level); grid: host MrGrid; slot: device int32
// per lattice brick, its table index or -1.
extern "C" int fe_multires_band(const void* grid, const void* bricks, const int32_t* slot, const int32_t* fine,
                                int nfine, float* tiles, void* stream) noexcept
{
    if (!grid || !bricks || !slot || !fine || nfine < 0 || !tiles) return FE_INVALID;
    if (nfine == 0) return FE_OK;
    const MrGrid G = *static_cast<const MrGrid*>(grid);
    hipLaunchKernelGGL(k_mr_band, dim3((unsigned)std::min(nfine, 65536)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), G, static_cast<const MrBrick*>(bricks), slot, fine, nfine,
                       tiles);
    FE_TRY(hipGetLastError());
    return FE_OK;
}

// Distances at n device f64 points (n x 3) into device f64 out; brick_out:
// optional device int32 (n x 4).
extern "C" int fe_multires_query(const void* grid, const void* bricks, const int32_t* slot, const int32_t* kind,
                                 const float* tiles, const double* points, int64_t n, double* out, int32_t* brick_out,
                                 void* stream) noexcept
{
    if (!grid || !bricks || !slot || !kind || !tiles || n < 0 || (n && (!points || !out))) return FE_INVALID;
    if (n == 0) return FE_OK;
    const MrGrid G = *static_cast<const MrGrid*>(grid);
    hipLaunchKernelGGL(k_mr_query, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, static_cast<hipStream_t>(stream), G,
                       static_cast<const MrBrick*>(bricks), slot, kind, tiles, points, n, out, brick_out);
    FE_TRY(hipGetLastError());
    return FE_OK;
}
