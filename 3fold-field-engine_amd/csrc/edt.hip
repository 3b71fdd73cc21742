// Exact integer squared Euclidean distance transform for gfx950.
//
// Mathematics (the reference's, FE/field_edt_integer_v1.py:14-26): with
// f = 0 on site voxels and f = sentinel = sum (n-1)^2 + 1 elsewhere, the
// result is min over s of f(s) + |p - s|^2, computed one axis at a time.  The
// per-axis minimum is exact in integers, so the axis order does not change
// the result; this implementation runs z first (the contiguous axis, straight
// from the uint8 mask) and then y and x (strided axes).
//
//   z pass  one wave per row: 64-bit ballots of the site flags, nearest site
//           left/right from clz/ctz.  Both polarities of a pair come from one
//           read of the mask.  Rows of >= 16 voxels write z summaries (per
//           16-voxel tile: the site bits and the distances to the nearest
//           sites outside the tile, 8 bytes per 16 voxels and plane) that the
//           y pass expands in registers, instead of int32 planes.
//   y,x     a tile of W lines adjacent along z; T threads per line each load
//           a chunk of positions into registers and build the exact lower
//           envelope of the chunk's parabolas as a stack of packed 32-bit
//           entries in LDS (pop test by an exact product comparison, no
//           division).  With every axis <= 1024 (k_edt_cand) each thread then
//           finds a minimiser at its chunk's first position by a pruned
//           search over the chunk envelopes and walks only the envelopes that
//           hold the minimisers of its positions; with a longer axis
//           (k_edt_lines) the chunk envelopes of a line are merged in log2(T)
//           rounds and walked.  Constant lines are returned as they are.
//   signed  FE/edt_persistent_pair_v1/edt.cu:36-42 semantics with IEEE
//           correctly rounded f64 sqrt/mul and f32 sub/mul.
//
// A launch covers a table of volumes (one, or a batch of uneven volumes).
#include "fe_common.hpp"
#include <algorithm>
#include <climits>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>

static thread_local char fe_err_buf[256];

void fe_note_error(hipError_t e, const char* where)
{
    snprintf(fe_err_buf, sizeof fe_err_buf, "%s: %s (%d)", where, hipGetErrorString(e), (int)e);
}

void fe_note_text(const char* msg) { snprintf(fe_err_buf, sizeof fe_err_buf, "%s", msg); }

// Text of the last runtime failure on this host thread ("" if none).
extern "C" const char* fe_last_error() noexcept { return fe_err_buf; }

// One volume of a launch.  A launch covers a table of volumes in two arenas:
// a byte arena of masks and int32 arenas of planes (out and work share the
// layout).  Tiles of all volumes are numbered consecutively per pass; the
// volume of a tile is the last one whose first tile is <= it.
struct EdtVol {
    int64_t mask;       // byte offset of the nx*ny*nz mask
    int64_t data;       // element offset of plane 0
    int64_t plane;      // elements from one plane to the next (>= nx*ny*nz)
    int32_t nx, ny, nz, sentinel;
    uint32_t first[3];  // first tile of the volume in the z, y and x pass
    int32_t planes;     // polarities (1 or 2) of the launch
};
static_assert(sizeof(EdtVol) == 56, "EdtVol layout");

namespace {

// Volume of `tile` in pass `pass` (wave-uniform): the ascending first tiles
// are read 64 at a time and counted with one ballot.
__device__ __forceinline__ int find_vol(const EdtVol* __restrict__ table, int nvol, int pass, uint32_t tile)
{
    const int lane = threadIdx.x & 63;
    for (int base = 0;; base += 64) {
        const int v = base + lane;
        const bool le = v < nvol && table[v].first[pass] <= tile;
        const int cnt = __popcll(__ballot(le));
        if (cnt < 64 || base + 64 >= nvol) return __builtin_amdgcn_readfirstlane(base + cnt - 1);
    }
}

__device__ __forceinline__ EdtVol load_vol(const EdtVol* __restrict__ table, int nvol, const EdtVol& one, int pass,
                                           uint32_t tile)
{
    return nvol ? table[find_vol(table, nvol, pass, tile)] : one;
}

// ---------------------------------------------------------------- z pass
// Site flags of one row (nz <= 4096) as 64-bit words, word s held by lane s:
// plane-0 sites are voxels whose mask equals site_one (1) or is zero (0);
// `valid` marks the positions < nz.
__device__ __forceinline__ void row_sites(const uint8_t* __restrict__ m, int nz, int lane, int site_one, uint64_t& w0,
                                          uint64_t& valid)
{
    const int nseg = (nz + 63) >> 6;
    w0 = 0;
    valid = 0;
    for (int s0 = 0; s0 < nseg; s0 += 8) {  // eight segments' loads in flight before their ballots
        uint8_t by[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int p = ((s0 + u) << 6) + lane;
            by[u] = p < nz ? m[p] : 0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int s = s0 + u;
            const int p = (s << 6) + lane;
            const bool in = p < nz;
            const bool hit = in && by[u] != 0;
            const uint64_t b = __ballot(site_one ? hit : (in && !hit));
            const uint64_t v = __ballot(in);
            if (lane == s) { w0 = b; valid = v; }
        }
    }
}

// Orders this wave's LDS writes before its later LDS reads by other lanes
// (LDS requests of one wave are served in order; the fences keep the
// compiler from moving accesses across).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// 64-bit value of lane `src` (per-lane index; every lane of the wave active).
__device__ __forceinline__ uint64_t lane_fetch64(uint64_t v, int src)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)(uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Z summaries: the z pass in compressed form, read by the y pass in place of
// full int32 planes.  For every row (x, y) and 16-voxel z tile t of it, each
// plane gets two words, A = the tile's 16 site bits | dl << 16 and B = dr,
// where dl = 16 t - (last site before the tile) and dr = (first site after the
// tile) - (16 t + 15), both ZSUM_NONE where there is no such site.  Plane pl's
// words sit at element offset data + pl * plane of the int32 arena, ordered
// (x, t, y) so that a y-pass tile reads them contiguously: A of (x, y, t) at
// (x * tz + t) * ys + y, B at nx * tz * ys further (tz = ceil(nz / 16), ys =
// ny rounded up to a multiple of 4 for 16-byte loads; 2 * nx * tz * ys <=
// nx * ny * nz for nz >= 16).  The squared z distance at
// z = 16 t + j follows from the in-tile bits, or else from dl / dr; with no
// site in the row both distances exceed any axis, so the value clamps to the
// sentinel.
//
// One lane per row, 64 consecutive rows per one-wave block, so that every
// word store covers 64 consecutive y.  The rows are staged through LDS 64
// bytes at a time with coalesced 4-byte loads (any row alignment); A is
// written in ascending tile order with the last site carried forward, then B
// in descending order from the site bits read back from the lane's own A
// words.
#define ZSUM_NONE 0x7fff
#define ZSUM_ROWS 64
// Rows of at least ZSUM_MIN_NZ voxels take the z-summary path (the summaries
// of a plane then fit in its nx*ny*nz words).
#define ZSUM_MIN_NZ 16
#define ZSUM_STRIDE 80  // LDS bytes per staged row: 16-byte aligned, conflict-free 16-byte reads

// Bit k (k < 4) set iff byte k of w is nonzero.
__device__ __forceinline__ uint32_t nonzero_nibble(uint32_t w)
{
    const uint32_t h = (((w & 0x7f7f7f7fu) + 0x7f7f7f7fu) | w) & 0x80808080u;
    return (h * 0x00204081u) >> 28;  // bits 7, 15, 23, 31 land on 28..31 without carries
}

template <bool PAIR>
__global__ __launch_bounds__(ZSUM_ROWS) void k_edt_zsum(const uint8_t* __restrict__ masks, int32_t* __restrict__ outs,
                                                        const EdtVol* __restrict__ table, int nvol, EdtVol one,
                                                        int site_one, uint32_t tiles)
{
    __shared__ __align__(16) uint8_t stage[ZSUM_ROWS * ZSUM_STRIDE];
    const int lane = threadIdx.x;
    // blocks of 64 rows dealt to the XCDs in groups of 64, round-robin (the
    // y pass then reads the summaries 0.04 ms faster at 512^3 than after
    // contiguous ranges per XCD)
    const uint32_t tile = xcd_group_tile<64>(blockIdx.x);  // grid: a multiple of 512
    if (tile >= tiles) return;  // uniform per block
    const EdtVol V = load_vol(table, nvol, one, 0, tile);
    const int64_t rows = (int64_t)V.nx * V.ny;
    const int64_t row0 = (int64_t)(tile - V.first[0]) * ZSUM_ROWS;
    const int nrows = (int)min((int64_t)ZSUM_ROWS, rows - row0);
    const int nz = V.nz, tz = (nz + 15) >> 4;
    const uint8_t* m0 = masks + V.mask + row0 * nz;
    const bool row_ok = lane < nrows;
    const int64_t row = row0 + lane;
    const int64_t x = row / V.ny, y = row - x * V.ny;
    const int64_t ys = (V.ny + 3) & ~3;  // words from one tile of a row to the next
    uint32_t* A0 = reinterpret_cast<uint32_t*>(outs + V.data) + x * tz * ys + y;
    uint32_t* A1 = reinterpret_cast<uint32_t*>(outs + V.data + V.plane) + x * tz * ys + y;
    const int64_t bofs = (int64_t)V.nx * tz * ys;
    // The block's rows are one contiguous byte span.  Each row's chunk
    // [c0, c0 + 64) is staged as the 17 aligned words that cover it (only
    // words that start before the span end, so every load holds a byte of the
    // span) and shifted into place by the row's byte misalignment.
    const uintptr_t span0 = reinterpret_cast<uintptr_t>(m0), span1 = span0 + (uintptr_t)nrows * nz;
    const int mis = (int)((span0 + (uintptr_t)lane * nz) & 3);  // c0 is a multiple of 64
    int last0 = -1, last1 = -1;                                  // last site so far
    for (int c0 = 0; c0 < nz; c0 += 64) {
        if (((span0 | (uintptr_t)nz) & 3) == 0) {  // every row 4-byte aligned: 16 words cover a chunk
            for (int f = lane; f < ZSUM_ROWS * 16; f += 64) {
                const int r = f >> 4, k = f & 15;
                uint32_t w = 0;
                if (r < nrows && c0 + 4 * k < nz)
                    w = *reinterpret_cast<const uint32_t*>(span0 + (uintptr_t)r * nz + c0 + 4 * k);
                reinterpret_cast<uint32_t*>(stage + r * ZSUM_STRIDE)[k] = w;
            }
        } else {
            for (int f = lane; f < ZSUM_ROWS * 17; f += 64) {
                const int r = f / 17, k = f - 17 * r;
                uint32_t w = 0;
                if (r < nrows) {
                    const uintptr_t at = ((span0 + (uintptr_t)r * nz + c0) & ~(uintptr_t)3) + 4 * k;
                    if (at < span1) w = *reinterpret_cast<const uint32_t*>(at);
                }
                reinterpret_cast<uint32_t*>(stage + r * ZSUM_STRIDE)[k] = w;
            }
        }
        __syncthreads();
        if (row_ok) {
            uint32_t w[17], q[16];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint4 u = *reinterpret_cast<const uint4*>(stage + lane * ZSUM_STRIDE + 16 * k);
                w[4 * k] = u.x;
                w[4 * k + 1] = u.y;
                w[4 * k + 2] = u.z;
                w[4 * k + 3] = u.w;
            }
            w[16] = mis ? *reinterpret_cast<const uint32_t*>(stage + lane * ZSUM_STRIDE + 64) : 0u;
#pragma unroll
            for (int k = 0; k < 16; ++k) q[k] = __builtin_amdgcn_alignbyte(w[k + 1], w[k], mis);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int t = (c0 >> 4) + k;
                if (t < tz) {
                    const uint32_t nzb = nonzero_nibble(q[4 * k]) | nonzero_nibble(q[4 * k + 1]) << 4 |
                                         nonzero_nibble(q[4 * k + 2]) << 8 | nonzero_nibble(q[4 * k + 3]) << 12;
                    const int rem = nz - 16 * t;
                    const uint32_t valid = rem >= 16 ? 0xffffu : (1u << rem) - 1u;
                    const uint32_t b0 = (site_one ? nzb : ~nzb) & valid;
                    A0[t * ys] = b0 | (uint32_t)(last0 < 0 ? ZSUM_NONE : 16 * t - last0) << 16;
                    if (b0) last0 = 16 * t + 31 - __builtin_clz(b0);
                    if constexpr (PAIR) {
                        const uint32_t b1 = ~b0 & valid;
                        A1[t * ys] = b1 | (uint32_t)(last1 < 0 ? ZSUM_NONE : 16 * t - last1) << 16;
                        if (b1) last1 = 16 * t + 31 - __builtin_clz(b1);
                    }
                }
            }
        }
        __syncthreads();  // the stage is reused
    }
    if (!row_ok) return;  // no barrier below
    int next0 = -1, next1 = -1;  // first site after the tile
    for (int t = tz - 1; t >= 0; --t) {
        const int rem = nz - 16 * t;
        const uint32_t valid = rem >= 16 ? 0xffffu : (1u << rem) - 1u;
        const uint32_t b0 = A0[t * ys] & 0xffffu;
        A0[bofs + t * ys] = (uint32_t)(next0 < 0 ? ZSUM_NONE : next0 - (16 * t + 15));
        if (b0) next0 = 16 * t + __builtin_ctz(b0);
        if constexpr (PAIR) {
            const uint32_t b1 = ~b0 & valid;
            A1[bofs + t * ys] = (uint32_t)(next1 < 0 ? ZSUM_NONE : next1 - (16 * t + 15));
            if (b1) next1 = 16 * t + __builtin_ctz(b1);
        }
    }
}

// The z pass with full int32 planes (rows shorter than 16 voxels, and the
// single-axis profiling entry); four rows (one per wave) per block; plane 1
// (PAIR) takes the complement of plane 0's sites.
template <bool PAIR>
__global__ __launch_bounds__(256) void k_edt_zpass(const uint8_t* __restrict__ masks, int32_t* __restrict__ outs,
                                                   const EdtVol* __restrict__ table, int nvol, EdtVol one,
                                                   int site_one)
{
    const int lane = threadIdx.x & 63;
    const EdtVol V = load_vol(table, nvol, one, 0, blockIdx.x);
    const int64_t rows = (int64_t)V.nx * V.ny;
    const int64_t row = (int64_t)(blockIdx.x - V.first[0]) * 4 + (threadIdx.x >> 6);
    if (row >= rows) return;  // uniform per wave
    const int nz = V.nz;
    const int64_t plane = V.plane;
    const int32_t sentinel = V.sentinel;
    int32_t* out = outs + V.data;
    const int nseg = (nz + 63) >> 6;
    uint64_t w0, valid;
    row_sites(masks + V.mask + row * nz, nz, lane, site_one, w0, valid);
    const uint64_t w1 = ~w0 & valid;
    const uint64_t seg0 = __ballot(w0 != 0), seg1 = __ballot(w1 != 0);
    int32_t* o0 = out + row * nz;
    int32_t* o1 = o0 + plane;
    const uint64_t low = lane == 63 ? ~0ull : ((2ull << lane) - 1);
    for (int s = 0; s < nseg; ++s) {
        const int p = (s << 6) + lane;
#pragma unroll
        for (int pl = 0; pl < (PAIR ? 2 : 1); ++pl) {
            const uint64_t mw = pl ? w1 : w0, segs = pl ? seg1 : seg0;
            const uint64_t w = readlane64(mw, s);
            // nearest site in earlier / later segments (wave-uniform)
            int carry_l = INT_MIN / 2, carry_r = INT_MAX / 2;
            uint64_t prev = segs & ((1ull << s) - 1);
            if (prev) {
                int j = 63 - __builtin_clzll(prev);
                carry_l = (j << 6) + 63 - __builtin_clzll(readlane64(mw, j));
            }
            uint64_t next = s == 63 ? 0 : (segs >> (s + 1));
            if (next) {
                int j = s + 1 + __builtin_ctzll(next);
                carry_r = (j << 6) + __builtin_ctzll(readlane64(mw, j));
            }
            uint64_t wl = w & low, wr = w >> lane;
            int left = wl ? (s << 6) + 63 - __builtin_clzll(wl) : carry_l;
            int right = wr ? p + __builtin_ctzll(wr) : carry_r;
            int d = min(p - left, right - p);
            int32_t v = (segs == 0) ? sentinel : d * d;
            if (p < nz) (pl ? o1 : o0)[p] = v;
        }
    }
}

// ---------------------------------------------------------------- line passes
// Hull entries are packed into one 32-bit word, F = f(s) + s^2.  Only sites
// with f < sentinel enter a hull.  When every axis is <= 1024 (SMALL),
// F < 3*1023^2 + 1023^2 < 2^22 and the word is (F << 10) | s; otherwise
// (axes <= 4096) F < 3*4095^2 + 4095^2 < 2^26 and the word is
// (F << 6) | (s - chunk start), chunks holding at most 64 positions.
template <bool SMALL>
__device__ __forceinline__ uint32_t pack_entry(int32_t F, int s, int rel)
{
    return SMALL ? (((uint32_t)F << 10) | (uint32_t)s) : (((uint32_t)F << 6) | (uint32_t)rel);
}

// Site b stays on the lower envelope of sites a < b < c iff its parabola is
// strictly lowest somewhere, i.e. the point (b, Fb) lies strictly below the
// segment from (a, Fa) to (c, Fc); exact in int64, no division.
// With every axis <= 1024 (SMALL), F < 2^22 and site differences < 2^10, so
// both products are exact 24 x 24-bit products (full-rate multiplies).
__device__ __forceinline__ int32_t sx24(int32_t x) { return (x << 8) >> 8; }
template <bool SMALL>
__device__ __forceinline__ bool above(int32_t dF1, int ds1, int32_t dF2, int ds2)
{
    if constexpr (SMALL) return (int64_t)sx24(dF1) * sx24(ds1) > (int64_t)sx24(dF2) * sx24(ds2);
    else return (int64_t)dF1 * ds1 > (int64_t)dF2 * ds2;
}
template <bool SMALL>
__device__ __forceinline__ bool keep_mid(int sa, int32_t Fa, int sb, int32_t Fb, int sc, int32_t Fc)
{
    return above<SMALL>(Fc - Fb, sb - sa, Fb - Fa, sc - sb);
}

// Loads positions [p0, p1) of one line into v[0 .. p1-p0) and the sentinel
// into the rest of v (such a site never enters a hull).  ZSUM: the y pass
// over z summaries (k_edt_zsum), rebuilding the squared z distance of (x =
// outer, y = p, z) from the two words of its 16-voxel z tile; otherwise an
// int32 plane read at `ps` with `axis_stride` between positions.
template <int CMAX, bool ZSUM>
__device__ __forceinline__ void load_chunk(int32_t (&v)[CMAX], const int32_t* __restrict__ src, const EdtVol& V,
                                           int plane, int outer, int z, int L, int p0, int p1, int64_t lbase,
                                           int64_t axis_stride)
{
    if constexpr (ZSUM) {
        const int nz = V.nz, j = z & 15;
        const int32_t sentinel = V.sentinel;
        const int tz = (nz + 15) >> 4;
        const int64_t ys = (L + 3) & ~3;
        const uint32_t* pa = reinterpret_cast<const uint32_t*>(src) + V.data + (int64_t)plane * V.plane +
                             ((int64_t)outer * tz + (z >> 4)) * ys + p0;
        const int64_t bofs = (int64_t)V.nx * tz * ys;
        const uint32_t lowm = (2u << j) - 1u;
        auto expand = [&](uint32_t a, uint32_t b) {
            const uint32_t bits = a & 0xffffu;
            const uint32_t lo = bits & lowm, hi = bits >> j;
            const int left = lo ? j - 31 + __builtin_clz(lo) : j + (int)(a >> 16);
            const int right = hi ? __builtin_ctz(hi) : 15 - j + (int)b;
            const int d = min(left, right);
            return min((int32_t)__umul24((uint32_t)d, (uint32_t)d), sentinel);
        };
        // 16-byte loads when both word runs start 16-byte aligned and the
        // chunk length is a multiple of 4 (CMAX is one in every tile shape)
        static_assert(CMAX % 4 == 0, "tile shapes hold whole 4-word groups");
        if (p1 - p0 == CMAX && ((reinterpret_cast<uintptr_t>(pa) & 15) | (bofs & 3)) == 0) {
#pragma unroll
            for (int i = 0; i < CMAX; i += 4) {  // a whole chunk: no per-position test
                const uint4 a = *reinterpret_cast<const uint4*>(pa + i);
                const uint4 b = *reinterpret_cast<const uint4*>(pa + bofs + i);
                v[i] = expand(a.x, b.x);
                v[i + 1] = expand(a.y, b.y);
                v[i + 2] = expand(a.z, b.z);
                v[i + 3] = expand(a.w, b.w);
            }
        } else if (((reinterpret_cast<uintptr_t>(pa) & 15) | ((bofs | (p1 - p0)) & 3)) == 0) {
#pragma unroll
            for (int i = 0; i < CMAX; i += 4) {
                if (p0 + i < p1) {
                    const uint4 a = *reinterpret_cast<const uint4*>(pa + i);
                    const uint4 b = *reinterpret_cast<const uint4*>(pa + bofs + i);
                    v[i] = expand(a.x, b.x);
                    v[i + 1] = expand(a.y, b.y);
                    v[i + 2] = expand(a.z, b.z);
                    v[i + 3] = expand(a.w, b.w);
                } else {
                    v[i] = v[i + 1] = v[i + 2] = v[i + 3] = sentinel;
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < CMAX; ++i) v[i] = p0 + i < p1 ? expand(pa[i], pa[bofs + i]) : sentinel;
        }
    } else {
        const int32_t* ps = src + lbase + (int64_t)p0 * axis_stride;
        if (p1 - p0 == CMAX) {  // a whole chunk: no per-position test
#pragma unroll
            for (int i = 0; i < CMAX; ++i) {
                v[i] = *ps;
                ps += axis_stride;
            }
        } else {
#pragma unroll
            for (int i = 0; i < CMAX; ++i) {
                v[i] = p0 + i < p1 ? *ps : V.sentinel;
                ps += axis_stride;
            }
        }
    }
}

// Streaming stores of v[0 .. p1-p0) to positions [p0, p1) of one line: a
// pass's output is far larger than the caches.
template <int CMAX>
__device__ __forceinline__ void store_chunk(const int32_t (&v)[CMAX], int32_t* __restrict__ dst, int64_t lbase,
                                            int p0, int p1, int64_t axis_stride)
{
    int32_t* pd = dst + lbase + (int64_t)p0 * axis_stride;
    if (p1 - p0 == CMAX) {  // a whole chunk: no per-position test
#pragma unroll
        for (int i = 0; i < CMAX; ++i) {
            __builtin_nontemporal_store(v[i], pd);
            pd += axis_stride;
        }
    } else {
#pragma unroll
        for (int i = 0; i < CMAX; ++i) {
            if (p0 + i < p1) __builtin_nontemporal_store(v[i], pd);
            pd += axis_stride;
        }
    }
}

// One exact lower-envelope pass over W lines x L positions, for launches with
// an axis above 1024 (entries (F << 6) | offset, int64 hull tests).  Thread
// (line, c) owns chunk c (C consecutive positions) of one line; the W lines of
// a tile are adjacent along z, so every load and store of a warp-wide step
// touches W consecutive int32 words per position.
//  1. the thread loads its C values into registers and builds the lower hull
//     of the points (s, f(s)+s^2) of its chunk (the parabola envelope) as a
//     stack of packed entries in LDS, keeping the top two in registers;
//  2. the chunk hulls of a line are merged in log2(T) rounds: in round h the
//     thread at the boundary of two adjacent groups pops the tail of the left
//     group and the head of the right group until the bridge between them is
//     found (merging lower hulls of x-separated point sets), with the entries
//     next to the bridge cached in registers;
//  3. every thread walks the merged envelope over its positions (the minimum
//     is unimodal along the envelope and its site moves right with p),
//     stepping through the chunk stacks, and stores its results from
//     registers.
// A line whose values are all equal is its own result (g = f).  Sites holding
// the sentinel (no site reached yet) are left out of the hulls: such a site
// contributes sentinel + (p-s)^2 >= sentinel, with equality only at its own
// position, so the result is min(envelope of the other sites, sentinel).
// LDS: kmax*W*T packed entries (kmax = the largest chunk length of the launch)
// plus four bytes per thread and eight bytes per line.  ZSUM: the y pass
// reading z summaries (k_edt_zsum) instead of int32 planes.
template <int W, int T, int CMAX, bool ZSUM>
__global__ __launch_bounds__(W* T) void k_edt_lines(const int32_t* __restrict__ src, int32_t* __restrict__ dst,
                                                    const EdtVol* __restrict__ table, int nvol, EdtVol one, int axis,
                                                    uint32_t tiles, int region)
{
    extern __shared__ uint32_t lds[];
    constexpr int NT = W * T;
    const uint32_t tile = xcd_group_tile<32>(blockIdx.x);  // grid: a multiple of 256
    if (tile >= tiles) return;  // uniform per block
    const int pass = axis == 1 ? 1 : 2;
    const EdtVol V = load_vol(table, nvol, one, pass, tile);
    const uint32_t local = tile - V.first[pass];
    const int nz = V.nz;
    const int32_t sentinel = V.sentinel;
    const int64_t yz = (int64_t)V.ny * nz;
    const int L = axis == 1 ? V.ny : V.nx;
    const int64_t axis_stride = axis == 1 ? (int64_t)nz : yz;
    const int64_t outer_stride = axis == 1 ? yz : (int64_t)nz;
    // tiles run z tile fastest, then plane, then the outer axis, and are dealt
    // to the XCDs in groups of 32, round-robin: every XCD gets lines from the
    // whole volume and both planes (plane 1 of a pair holds mostly constant
    // lines, plane 0 the long envelopes; the solid of a part often fills only
    // the middle of its volume), while the z tiles of a row stay in one L2 and
    // the XCDs stream nearby rows together (measured at 512^3: y pass 0.65 ms
    // with contiguous ranges per XCD, 0.53 with pairs, 0.47 with groups of 32;
    // x pass 0.70, 0.84 and 0.65 ms)
    const int ztiles = (nz + W - 1) / W;
    const int zt = local % ztiles;
    const int rest = local / ztiles;
    const int nplanes = V.planes > 1 ? V.planes : 1;
    const int outer = rest / nplanes, plane = rest % nplanes;
    const int64_t base = V.data + (int64_t)plane * V.plane + (int64_t)outer * outer_stride + (int64_t)zt * W;
    const int nlines = min(W, nz - zt * W);

    uint32_t* stk = lds;                                        // [kmax][NT] packed hull entries (region words)
    uint8_t* blo = reinterpret_cast<uint8_t*>(stk + region);  // [NT] surviving range of a chunk hull
    uint8_t* bhi = blo + NT;
    uint8_t* lsb = bhi + NT;                                  // [NT] nearest nonempty chunk <= c
    uint8_t* rsb = lsb + NT;                                  // [NT] nearest nonempty chunk >= c
    int32_t* lval = reinterpret_cast<int32_t*>(rsb + NT);     // [W] first value of the line
    int32_t* lflat = lval + W;                                // [W] line is constant

    const int tid = threadIdx.x;
    const int line = tid % W, c = tid / W;
    const int C = (L + T - 1) / T;
    const int p0 = min(c * C, L), p1 = min(p0 + C, L);
    const bool active = line < nlines;
    // slot of chunk b of this line: rotated within the W-slot row so that the
    // threads of one line (merge) and of one chunk row (hull, walk) spread over banks
    auto slot = [&](int b) { return b * W + (line ^ (b & (W - 1))); };
    auto entry = [&](int k, int b, int& s, int32_t& F) {
        const uint32_t e = stk[k * NT + slot(b)];
        s = b * C + (int)(e & 63u);
        F = (int32_t)(e >> 6);
    };
    const int me = slot(c);
    const int64_t lbase = base + line;

    int32_t v[CMAX];
    if (active) load_chunk<CMAX, ZSUM>(v, src, V, plane, outer, zt * W + line, L, p0, p1, lbase, axis_stride);

    // 0. constant lines (all zero, all at the sentinel, ...) are their own
    // result: found first, their threads skip the hull, merge and output.
    int32_t mn = INT_MAX, mx = INT_MIN;
    if (active) {
#pragma unroll
        for (int i = 0; i < CMAX; ++i)
            if (p0 + i < p1) {
                mn = min(mn, v[i]);
                mx = max(mx, v[i]);
            }
    }
    if (c == 0) {
        lval[line] = mn;
        lflat[line] = 1;
    }
    __syncthreads();
    if (active && p0 < p1 && !(mn == mx && mn == lval[line])) lflat[line] = 0;
    __syncthreads();
    const bool busy = active && !lflat[line];

    // 1. hull of this chunk (positions past the chunk end read as the sentinel,
    // which never enters a hull, so one test covers both)
    int k = -1, s_top = 0, s_sec = 0;
    int32_t F_top = 0, F_sec = 0;
    if (busy) {
        uint32_t* top = stk + me - NT;  // stack slot k lives at stk[k * NT + me]
#pragma unroll
        for (int i = 0; i < CMAX; ++i) {
            const int q = p0 + i;
            const int32_t fq = q < p1 ? v[i] : sentinel;
            if (fq < sentinel) {
                const int32_t Fq = fq + (int32_t)__umul24((uint32_t)q, (uint32_t)q);
                while (k >= 1) {
                    if (above<false>(Fq - F_top, s_top - s_sec, F_top - F_sec, q - s_top)) break;
                    --k;
                    top -= NT;
                    s_top = s_sec;
                    F_top = F_sec;
                    if (k >= 1) entry(k - 1, c, s_sec, F_sec);
                }
                ++k;
                top += NT;
                *top = pack_entry<false>(Fq, q, i);
                s_sec = s_top;
                F_sec = F_top;
                s_top = q;
                F_top = Fq;
            }
        }
    }
    blo[me] = 0;
    bhi[me] = (uint8_t)(k + 1);
    // A wave holds CPW consecutive chunks of each of its lines, so the merge
    // rounds whose groups of 2h chunks fit in one wave read only data of their
    // own wave: those rounds (and the hull stacks before them) need a wave
    // sync, not a block barrier.
    constexpr int CPW = FE_WAVE / W;
    static_assert(CPW >= 2, "a wave holds at least two chunks of a line");
    wave_sync();

    // 2. tree merge of the chunk hulls
    for (int h = 1; h < T; h <<= 1) {
        if (busy && (c & (2 * h - 1)) == h) {
            const int a0 = c - h, a2 = min(c + h, T);
            int ta = c - 1;
            while (ta >= a0 && bhi[slot(ta)] == blo[slot(ta)]) --ta;
            int hb = c;
            while (hb < a2 && bhi[slot(hb)] == blo[slot(hb)]) ++hb;
            if (ta >= a0 && hb < a2) {
                int th = bhi[slot(ta)], tlo = blo[slot(ta)];
                int hl = blo[slot(hb)], hhi = bhi[slot(hb)];
                int s1, s3;
                int32_t F1, F3;
                entry(th - 1, ta, s1, F1);
                entry(hl, hb, s3, F3);
                bool changed = true;
                while (changed) {
                    changed = false;
                    for (;;) {  // tail of the left group
                        int s2, pa = ta;
                        int32_t F2;
                        if (th - tlo >= 2) entry(th - 2, ta, s2, F2);
                        else {
                            pa = ta - 1;
                            while (pa >= a0 && bhi[slot(pa)] == blo[slot(pa)]) --pa;
                            if (pa < a0) break;
                            entry(bhi[slot(pa)] - 1, pa, s2, F2);
                        }
                        if (keep_mid<false>(s2, F2, s1, F1, s3, F3)) break;
                        if (--th == tlo) {
                            bhi[slot(ta)] = (uint8_t)th;
                            ta = pa;
                            th = bhi[slot(pa)];
                            tlo = blo[slot(pa)];
                        }
                        s1 = s2;
                        F1 = F2;
                    }
                    for (;;) {  // head of the right group
                        int s4, nb = hb;
                        int32_t F4;
                        if (hhi - hl >= 2) entry(hl + 1, hb, s4, F4);
                        else {
                            nb = hb + 1;
                            while (nb < a2 && bhi[slot(nb)] == blo[slot(nb)]) ++nb;
                            if (nb >= a2) break;
                            entry(blo[slot(nb)], nb, s4, F4);
                        }
                        if (keep_mid<false>(s1, F1, s3, F3, s4, F4)) break;
                        changed = true;
                        if (++hl == hhi) {
                            blo[slot(hb)] = (uint8_t)hl;
                            hb = nb;
                            hl = blo[slot(nb)];
                            hhi = bhi[slot(nb)];
                        }
                        s3 = s4;
                        F3 = F4;
                    }
                }
                bhi[slot(ta)] = (uint8_t)th;
                blo[slot(hb)] = (uint8_t)hl;
            }
        }
        if (4 * h <= CPW) wave_sync();  // the next round (2h) stays within the wave
        else __syncthreads();
    }
    {
        int b = c;
        while (b >= 0 && bhi[slot(b)] == blo[slot(b)]) --b;
        lsb[me] = (uint8_t)(b < 0 ? 255 : b);
        b = c;
        while (b < T && bhi[slot(b)] == blo[slot(b)]) ++b;
        rsb[me] = (uint8_t)(b >= T ? 255 : b);
    }
    __syncthreads();

    // 3. walk the merged envelope; results replace v
    if (active && p0 < p1 && !lflat[line]) {
        int cb = bhi[me] > blo[me] ? c : (lsb[me] != 255 ? lsb[me] : rsb[me]);
        if (cb == 255) {
#pragma unroll
            for (int i = 0; i < CMAX; ++i) v[i] = sentinel;
        } else {
            int clo = blo[slot(cb)], chi = bhi[slot(cb)];
            int ci = cb >= c ? clo : chi - 1;
            int cs;
            int32_t cF;
            entry(ci, cb, cs, cF);
            // next entry to the right; false at the end of the envelope
            auto next = [&](int& nb, int& ni, int& nlo, int& nhi) {
                nb = cb;
                ni = ci + 1;
                nlo = clo;
                nhi = chi;
                if (ni < chi) return true;
                if (cb + 1 >= T) return false;
                nb = rsb[slot(cb + 1)];
                if (nb == 255) return false;
                nlo = blo[slot(nb)];
                nhi = bhi[slot(nb)];
                ni = nlo;
                return true;
            };
            bool moved = false;
            for (;;) {  // at p0: right while not worse
                int nb, ni, nlo, nhi, ns;
                int32_t nF;
                if (!next(nb, ni, nlo, nhi)) break;
                entry(ni, nb, ns, nF);
                if (nF - 2 * p0 * ns > cF - 2 * p0 * cs) break;
                cb = nb; ci = ni; clo = nlo; chi = nhi; cs = ns; cF = nF;
                moved = true;
            }
            if (!moved) {
                for (;;) {  // left while strictly better
                    int pb = cb, pi = ci - 1, plo = clo, phi = chi;
                    if (pi < clo) {
                        if (cb == 0) break;
                        pb = lsb[slot(cb - 1)];
                        if (pb == 255) break;
                        plo = blo[slot(pb)];
                        phi = bhi[slot(pb)];
                        pi = phi - 1;
                    }
                    int ps;
                    int32_t pF;
                    entry(pi, pb, ps, pF);
                    if (pF - 2 * p0 * ps >= cF - 2 * p0 * cs) break;
                    cb = pb; ci = pi; clo = plo; chi = phi; cs = ps; cF = pF;
                }
            }
#pragma unroll
            for (int i = 0; i < CMAX; ++i) {
                const int p = p0 + i;
                if (p < p1) {
                    for (;;) {
                        int nb, ni, nlo, nhi, ns;
                        int32_t nF;
                        if (!next(nb, ni, nlo, nhi)) break;
                        entry(ni, nb, ns, nF);
                        if (nF - 2 * p * ns > cF - 2 * p * cs) break;
                        cb = nb; ci = ni; clo = nlo; chi = nhi; cs = ns; cF = nF;
                    }
                    v[i] = min(cF - 2 * p * cs + p * p, sentinel);
                }
            }
        }
    }
    if (active) store_chunk<CMAX>(v, dst, lbase, p0, p1, axis_stride);
}

// Tiles per XCD group in the y pass of k_edt_cand (the x pass keeps 32: with
// positions 1 MB apart it prefers each XCD to stream nearby rows).  At 512^3
// groups of 64 read the z summaries faster (y 0.39 -> 0.36 ms); 16 slowed x
// to 0.79 ms.
#ifndef CAND_YG
#define CAND_YG 64
#endif

// The lower-envelope pass without a merge of the chunk hulls (every axis <=
// 1024, so entries are (F << 10) | s).  Tiles, chunks and hulls are those of
// k_edt_lines.  With h(p, s) = f(s) + (p - s)^2:
//  0. the owner of each chunk hull also records the hull's minimum at its
//     chunk's first position p_b and at the next chunk's first position
//     p_{b+1} (along a hull the values at one position fall and then rise, so
//     short scans from its two ends find them), and each line keeps a bit
//     mask of its nonempty chunks;
//  1. each thread finds one minimiser a_c of its line at p_c = c*C, visiting
//     the nonempty chunks outward from its own.  A chunk b > c cannot beat
//     its minimum at p_b plus (p_b - p_c)^2 at p_c, a chunk b < c its minimum
//     at p_{b+1} plus (p_c - p_{b+1})^2 (for s >= p_b >= p,
//     (s - p)^2 >= (s - p_b)^2 + (p_b - p)^2), and no chunk at distance >= d
//     the line's smallest f plus d^2, which ends the walk outward; a visited
//     chunk's minimum at p_c is found by walking its hull from the minimum
//     at its boundary nearest p_c;
//  2. the positions p_c <= p < p_{c+1} of the chunk take the minimum over the
//     sites from a_c to a_{c+1}: for s < a_c and p > p_c,
//     h(p, s) - h(p, a_c) = [h(p_c, s) - h(p_c, a_c)] + 2 (p - p_c)(a_c - s) > 0,
//     and likewise every site beyond a_{c+1} loses to it before p_{c+1}, so
//     every minimiser at p lies in [a_c, a_{c+1}].  The nonempty chunks
//     holding those sites (the first from a_c's entry on, the last up to
//     a_{c+1}'s) are each walked left to right along the positions.
// Four block barriers in all, none in data-dependent rounds.  Constant lines
// are their own result and sentinel sites stay out of the hulls, as in
// k_edt_lines.  LDS: kmax*W*T entries, 12 bytes per thread, 16 per line.
template <int W, int T, int CMAX, bool ZSUM>
__global__ __launch_bounds__(W* T, CMAX <= 16 ? 8 : (CMAX <= 24 ? 6 : 4)) void k_edt_cand(
    const int32_t* __restrict__ src, int32_t* __restrict__ dst, const EdtVol* __restrict__ table, int nvol,
    EdtVol one, int axis, uint32_t tiles, int region)
{
    extern __shared__ uint32_t lds[];
    constexpr int NT = W * T;
    static_assert((W & (W - 1)) == 0 && CMAX <= 32 && T <= 64, "tile shape");
    // a bit per chunk of a line; helpers: lowest and highest set bit
    using Mask = typename std::conditional<(T > 32), unsigned long long, uint32_t>::type;
    constexpr int MB = 8 * (int)sizeof(Mask);
    auto low_bit = [](Mask m) { return MB == 64 ? __builtin_ctzll((unsigned long long)m) : __builtin_ctz((uint32_t)m); };
    auto high_bit = [](Mask m) {
        return MB == 64 ? 63 - __builtin_clzll((unsigned long long)m) : 31 - __builtin_clz((uint32_t)m);
    };
    // groups of CAND_YG tiles per XCD in the y pass, 32 in the x pass (grid: a
    // multiple of 8 times that)
    const uint32_t tile = axis == 1 ? xcd_group_tile<CAND_YG>(blockIdx.x) : xcd_group_tile<32>(blockIdx.x);
    if (tile >= tiles) return;  // uniform per block
    const int pass = axis == 1 ? 1 : 2;
    const EdtVol V = load_vol(table, nvol, one, pass, tile);
    const uint32_t local = tile - V.first[pass];
    const int nz = V.nz;
    const int32_t sentinel = V.sentinel;
    const int64_t yz = (int64_t)V.ny * nz;
    const int L = axis == 1 ? V.ny : V.nx;
    const int64_t axis_stride = axis == 1 ? (int64_t)nz : yz;
    const int64_t outer_stride = axis == 1 ? yz : (int64_t)nz;
    // tile numbering and dealing as in k_edt_lines: z tile fastest, then
    // plane, then outer; groups of 32 tiles round-robin over the XCDs
    const int ztiles = (nz + W - 1) / W;
    const int zt = local % ztiles;
    const int rest = local / ztiles;
    const int nplanes = V.planes > 1 ? V.planes : 1;
    const int outer = rest / nplanes, plane = rest % nplanes;
    const int64_t base = V.data + (int64_t)plane * V.plane + (int64_t)outer * outer_stride + (int64_t)zt * W;
    const int nlines = min(W, nz - zt * W);

    uint32_t* stk = lds;                                       // [kmax][NT] packed hull entries (region words)
    int32_t* lv = reinterpret_cast<int32_t*>(stk + region);   // [NT] hull minimum at p_b
    int32_t* rv = lv + NT;                                     // [NT] hull minimum at p_{b+1}
    int32_t* lmin = rv + NT;                                   // [W] smallest value of the line
    int32_t* lmax = lmin + W;                                  // [W] largest value of the line
    Mask* lmask = reinterpret_cast<Mask*>(lmax + W);          // [W] nonempty chunks of the line (8-byte aligned:
                                                               // region, 2*NT and 2*W words are even)
    uint16_t* info = reinterpret_cast<uint16_t*>(lmask + W);  // [NT] size | argmin at p_b << 6 | at p_{b+1} << 11
    uint16_t* bnd = info + NT;                                 // [NT] minimiser at p_c: chunk << 8 | entry

    const int tid = threadIdx.x;
    const int line = tid % W, c = tid / W;
    const int C = (L + T - 1) / T;
    const int p0 = min(c * C, L), p1 = min(p0 + C, L), np = p1 - p0;
    const bool active = line < nlines;
    // slot of chunk b of this line, rotated within the W-slot row (banks)
    auto slot = [&](int b) { return b * W + (line ^ (b & (W - 1))); };
    const int me = slot(c);
    const int64_t lbase = base + line;
    if (c == 0) {
        lmin[line] = INT_MAX;
        lmax[line] = INT_MIN;
        lmask[line] = 0;
    }

    int32_t v[CMAX];
    if (active) load_chunk<CMAX, ZSUM>(v, src, V, plane, outer, zt * W + line, L, p0, p1, lbase, axis_stride);
    int32_t mn = INT_MAX, mx = INT_MIN;
    if (active && np == CMAX) {
#pragma unroll
        for (int i = 0; i < CMAX; ++i) {
            mn = min(mn, v[i]);
            mx = max(mx, v[i]);
        }
    } else if (active) {
#pragma unroll
        for (int i = 0; i < CMAX; ++i)
            if (i < np) {
                mn = min(mn, v[i]);
                mx = max(mx, v[i]);
            }
    }
    __syncthreads();  // the line records are initialised
    if (active && np > 0) {
        atomicMin(lmin + line, mn);
        atomicMax(lmax + line, mx);
    }
    __syncthreads();
    const int32_t fline = lmin[line];
    const bool busy = active && fline != lmax[line];

    // value of entry e at position p, less p^2
    auto at = [](uint32_t e, int p) { return (int32_t)(e >> 10) - (int32_t)__umul24(2u * (uint32_t)p, e & 1023u); };
    // 0. hull of this chunk and its minima at p_c and p_{c+1}
    if (busy) {
        int k = -1, s_top = 0, s_sec = 0;
        int32_t F_top = 0, F_sec = 0;
        uint32_t* top = stk + me - NT;  // stack slot k lives at stk[k * NT + me]
        // site i of the chunk onto the stack.  CLEAN: every position of the
        // chunk is a site (a full chunk without the sentinel), so no site test
        // is needed, and from the third site on the stack holds two entries
        auto add = [&](const int i, auto clean) {
            constexpr bool CLEAN = decltype(clean)::value;
            const int q = p0 + i;
            const int32_t fq = v[i];  // the sentinel past the chunk end
            if (CLEAN || fq < sentinel) {
                const int32_t Fq = fq + (int32_t)__umul24((uint32_t)q, (uint32_t)q);
                if ((CLEAN ? i >= 2 : k >= 1) && !above<true>(Fq - F_top, s_top - s_sec, F_top - F_sec, q - s_top)) {
                    do {  // pops are rare: the loop stays off the common path
                        --k;
                        top -= NT;
                        s_top = s_sec;
                        F_top = F_sec;
                        if (k >= 1) {
                            const uint32_t e = stk[(k - 1) * NT + me];
                            s_sec = (int)(e & 1023u);
                            F_sec = (int32_t)(e >> 10);
                        }
                    } while (k >= 1 && !above<true>(Fq - F_top, s_top - s_sec, F_top - F_sec, q - s_top));
                }
                ++k;
                top += NT;
                *top = pack_entry<true>(Fq, q, i);
                s_sec = s_top;
                F_sec = F_top;
                s_top = q;
                F_top = Fq;
            }
        };
        if (__all(np == CMAX && mx < sentinel)) {  // wave-uniform
#pragma unroll
            for (int i = 0; i < CMAX; ++i) add(i, std::true_type{});
        } else {
#pragma unroll
            for (int i = 0; i < CMAX; ++i) add(i, std::false_type{});
        }
        const int m = k + 1;
        uint32_t inf = (uint32_t)m;
        if (m > 0) {
            int li = 0;  // minimum at p_c: from the first entry on
            int32_t cur = at(stk[me], p0);
            while (li + 1 < m) {
                const int32_t nxt = at(stk[(li + 1) * NT + me], p0);
                if (nxt >= cur) break;
                ++li;
                cur = nxt;
            }
            lv[me] = cur + p0 * p0;
            int ri = k;  // minimum at p_{c+1}: from the last entry back (the top two are in registers)
            cur = F_top - (int32_t)__umul24(2u * (uint32_t)p1, (uint32_t)s_top);
            if (k >= 1) {
                const int32_t sec = F_sec - (int32_t)__umul24(2u * (uint32_t)p1, (uint32_t)s_sec);
                if (sec <= cur) {
                    ri = k - 1;
                    cur = sec;
                    while (ri > 0) {
                        const int32_t nxt = at(stk[(ri - 1) * NT + me], p1);
                        if (nxt > cur) break;
                        --ri;
                        cur = nxt;
                    }
                }
            }
            rv[me] = cur + p1 * p1;
            inf |= (uint32_t)li << 6 | (uint32_t)ri << 11;
            atomicOr(lmask + line, (Mask)1 << c);
        }
        info[me] = (uint16_t)inf;
    }
    __syncthreads();  // hulls, their minima and the masks published

    const Mask mask = busy ? lmask[line] : (Mask)0;
    // 1. a minimiser at p_c
    if (busy && np > 0) {
        const int p = p0, pp = p * p;
        int32_t best = INT_MAX;  // h(p_c, best site)
        uint32_t sel = 0;
        if (mask >> c & 1) {
            best = lv[me];
            sel = (uint32_t)c << 8 | (uint32_t)(info[me] >> 6 & 31u);
        }
        for (Mask mr = c + 1 < T ? mask >> (c + 1) << (c + 1) : (Mask)0; mr; mr &= mr - 1) {
            const int b = low_bit(mr);  // nonempty chunks to the right, nearest first
            const int d = b * C - p;          // distance to the chunk's first position
            if (fline + d * d >= best) break;
            const int sb = slot(b);
            if (lv[sb] + d * d >= best) continue;
            int i = info[sb] >> 6 & 31;  // from the minimum at p_b leftwards
            int32_t cur = at(stk[i * NT + sb], p);
            while (i > 0) {
                const int32_t nxt = at(stk[(i - 1) * NT + sb], p);
                if (nxt > cur) break;
                --i;
                cur = nxt;
            }
            if (cur + pp < best) {
                best = cur + pp;
                sel = (uint32_t)b << 8 | (uint32_t)i;
            }
        }
        for (Mask ml = mask & (((Mask)1 << c) - 1); ml; ml ^= (Mask)1 << high_bit(ml)) {
            const int b = high_bit(ml);  // nonempty chunks to the left, nearest first
            const int d = p - (b * C + C - 1);     // distance to the chunk's last position
            if (fline + d * d >= best) break;
            const int sb = slot(b);
            const int e = p - (b + 1) * C;  // distance to p_{b+1}
            if (rv[sb] + e * e >= best) continue;
            const uint32_t inf = info[sb];
            const int m = (int)(inf & 63u);
            int i = (int)(inf >> 11 & 31u);  // from the minimum at p_{b+1} rightwards
            int32_t cur = rv[sb] - pp;       // exact when p_{b+1} = p_c
            if (e > 0) {
                cur = at(stk[i * NT + sb], p);
                while (i + 1 < m) {
                    const int32_t nxt = at(stk[(i + 1) * NT + sb], p);
                    if (nxt >= cur) break;
                    ++i;
                    cur = nxt;
                }
            }
            if (cur + pp < best) {
                best = cur + pp;
                sel = (uint32_t)b << 8 | (uint32_t)i;
            }
        }
        bnd[me] = (uint16_t)sel;
    }
    __syncthreads();  // minimisers published

    // 2. the chunk's positions over the hulls holding the sites a_c .. a_{c+1}
    // (a constant line's value is its minimum, so no loaded value stays live).
    // Every hull adds the same p^2 at a position, so the walks take minima of
    // the values less p^2 and p^2 is added once, at the end.
    constexpr int32_t far = INT_MAX - (1 << 21);  // + p^2 (p < 1024) stays finite
#pragma unroll
    for (int j = 0; j < CMAX; ++j) v[j] = busy ? far : fline;
    if (busy && np > 0) {
        const uint32_t s0 = bnd[me];
        const int b0 = (int)(s0 >> 8), i0 = (int)(s0 & 255u);
        int b1 = high_bit(mask), i1 = 255;  // the last chunk of the line: every site from a_c on
        if (p1 < L) {
            const uint32_t s1 = bnd[slot(c + 1)];
            b1 = (int)(s1 >> 8);
            i1 = (int)(s1 & 255u);
        }
        Mask cm = mask & (b1 >= MB - 1 ? ~(Mask)0 : ((Mask)2 << b1) - 1) & ~(((Mask)1 << b0) - 1);
        for (; cm; cm &= cm - 1) {
            const int b = low_bit(cm);
            const int sb = slot(b);
            const uint32_t inf = info[sb];
            const int hi = min(b == b1 ? i1 : 31, (int)(inf & 63u) - 1);
            // first entry: a_c's, else the chunk's minimum at the boundary
            // nearest p_c, else (chunks to the right) its first
            const int i_start = b == b0 ? i0 : (b < c ? (int)(inf >> 11 & 31u) : (b == c ? (int)(inf >> 6 & 31u) : 0));
            int i = i_start;
            const uint32_t e = stk[i * NT + sb];
            int s2 = (int)(e & 1023u) << 1;
            int32_t cur = at(e, p0);
            // the next entry, or none: a value that no position reaches
            // (stays put, as its 2s is 0)
            int sn2 = 0;
            int32_t nxt = INT_MAX;
            if (i < hi) {
                const uint32_t en = stk[(i + 1) * NT + sb];
                sn2 = (int)(en & 1023u) << 1;
                nxt = at(en, p0);
            }
            // 2p as a recurrence: kept opaque so that the compiler does not
            // hoist CMAX precomputed copies out of the chunk loop
            int p2 = 2 * p0;
            asm volatile("" : "+v"(p2));
            // every position of v, also past the chunk end (never stored)
#pragma unroll
            for (int j = 0; j < CMAX; ++j) {
                while (nxt <= cur) {  // the minimum moves right with p
                    s2 = sn2;
                    cur = nxt;
                    sn2 = 0;
                    nxt = INT_MAX;
                    if (++i < hi) {
                        const uint32_t e2 = stk[(i + 1) * NT + sb];
                        sn2 = (int)(e2 & 1023u) << 1;
                        nxt = (int32_t)(e2 >> 10) - (int32_t)__umul24((uint32_t)p2, e2 & 1023u);
                    }
                }
                v[j] = min(v[j], cur);
                cur -= s2;  // values at p + 1
                nxt -= sn2;
                p2 += 2;
            }
        }
        int pp = p0 * p0, p2 = 2 * p0;
#pragma unroll
        for (int j = 0; j < CMAX; ++j) {
            v[j] = min(v[j] + pp, sentinel);
            pp += p2 + 1;
            p2 += 2;
        }
    }
    if (active) store_chunk<CMAX>(v, dst, lbase, p0, p1, axis_stride);
}

// ---------------------------------------------------------------- signed output
__global__ __launch_bounds__(256) void k_edt_signed(const int32_t* __restrict__ pair, float* __restrict__ out,
                                                    int64_t n, double pitch)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float o = (float)(sqrt((double)pair[i]) * pitch);
    const float in = (float)(sqrt((double)pair[n + i]) * pitch);
    const float d = o - in;
    const float sg = d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f);
    const float half = (float)(0.5 * pitch);
    __builtin_nontemporal_store(d - sg * half, out + i);  // written once, read by the caller
}

struct LineCfg {
    int W, T, CMAX;
    bool cand;  // k_edt_cand (every axis <= 1024), else k_edt_lines
};

// Tile shape by line length: W lines (W*4 contiguous bytes per position),
// T threads per line, CMAX >= ceil(L/T) positions per thread (registers).
// compact: every axis of every volume of the launch is <= 1024.  xpass: the
// x pass, whose positions lie a whole plane apart, takes 32 lines up to 512
// positions, so that every access of a position is one 128-byte piece; the
// y pass keeps 16 (its loads are contiguous summaries and its stores 16 lines
// a row apart, where the smaller tile's occupancy pays more).
LineCfg pick_cfg(int L, bool compact, bool xpass)
{
    if (compact) {
        const int w = xpass ? 32 : 16;
        if (L <= 128) return {w, 8, 16, true};
        if (L <= 256) return {w, 16, 16, true};
        if (L <= 512) return {w, 32, 16, true};
        // longer lines: more chunks of 16 rather than longer chunks, which
        // keeps 64 VGPRs and 6-8 waves per SIMD (768^3 thin-feature pair:
        // y 4.0 -> 3.4 ms, x 4.0 -> 3.2 ms against 32 chunks of 24)
        if (L <= 768) return {16, 48, 16, true};
        return {16, 64, 16, true};
    }
    if (L <= 2048) return {8, 64, 32, false};
    return {8, 64, 64, false};
}

template <int W, int T, int CMAX, bool ZSUM>
hipError_t launch_lines_t(const int32_t* src, int32_t* dst, const EdtVol* table, int nvol, const EdtVol& one,
                          int axis, int Lmax, uint64_t tiles, hipStream_t st)
{
    const uint32_t grid = round_up(tiles, 256);  // whole groups of 32 tiles on each XCD
    const int kmax = (Lmax + T - 1) / T;  // <= CMAX by the choice of shape
    const int region = kmax * W * T;      // hull stacks
    const size_t lds = (size_t)region * 4 + (size_t)W * T * 4 + (size_t)W * 8;
    const void* fn = reinterpret_cast<const void*>(&k_edt_lines<W, T, CMAX, ZSUM>);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_edt_lines<W, T, CMAX, ZSUM>), dim3(grid), dim3(W * T), lds, st, src, dst, table, nvol, one,
                       axis, (uint32_t)tiles, region);
    return hipGetLastError();
}

template <int W, int T, int CMAX, bool ZSUM>
hipError_t launch_cand_t(const int32_t* src, int32_t* dst, const EdtVol* table, int nvol, const EdtVol& one, int axis,
                         int Lmax, uint64_t tiles, hipStream_t st)
{
    const uint32_t grid = round_up(tiles, axis == 1 ? 8 * CAND_YG : 256);  // whole groups on each XCD
    const int kmax = (Lmax + T - 1) / T;  // <= CMAX by the choice of shape
    const int region = kmax * W * T;      // hull stacks
    // hulls; minima at both boundaries (8 bytes per thread), line extremes
    // (8 per line), line masks (8 per line), info and minimisers (4 per thread)
    const size_t lds = (size_t)region * 4 + (size_t)W * T * 12 + (size_t)W * 16;
    const void* fn = reinterpret_cast<const void*>(&k_edt_cand<W, T, CMAX, ZSUM>);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_edt_cand<W, T, CMAX, ZSUM>), dim3(grid), dim3(W * T), lds, st, src, dst, table, nvol, one,
                       axis, (uint32_t)tiles, region);
    return hipGetLastError();
}

// One line pass over every volume: Lmax (the longest line of the pass over
// all volumes) picks the tile shape, `tiles` counts the tiles of that shape.
// compact: every axis of every volume is <= 1024 (packed 32-bit envelope);
// zsum (y pass only): src holds z summaries.
hipError_t launch_lines(const int32_t* src, int32_t* dst, const EdtVol* table, int nvol, const EdtVol& one, int axis,
                        int Lmax, uint64_t tiles, bool compact, bool zsum, hipStream_t st)
{
    const LineCfg c = pick_cfg(Lmax, compact, axis == 0);
#define FE_CAND(w, t, cm)                                                                                    \
    if (c.cand && c.W == w && c.T == t && c.CMAX == cm)                                                      \
        return zsum ? launch_cand_t<w, t, cm, true>(src, dst, table, nvol, one, axis, Lmax, tiles, st)      \
                    : launch_cand_t<w, t, cm, false>(src, dst, table, nvol, one, axis, Lmax, tiles, st);
    FE_CAND(32, 8, 16)
    FE_CAND(32, 16, 16)
    FE_CAND(32, 32, 16)
    FE_CAND(16, 8, 16)
    FE_CAND(16, 16, 16)
    FE_CAND(16, 32, 16)
    FE_CAND(16, 48, 16)
    FE_CAND(16, 64, 16)
#undef FE_CAND
#define FE_LINES_Z(w, t, cm)                                                                                 \
    return zsum ? launch_lines_t<w, t, cm, true>(src, dst, table, nvol, one, axis, Lmax, tiles, st)         \
                : launch_lines_t<w, t, cm, false>(src, dst, table, nvol, one, axis, Lmax, tiles, st);
    if (!c.cand && c.W == 8 && c.T == 64 && c.CMAX == 32) { FE_LINES_Z(8, 64, 32) }
    if (!c.cand && c.W == 8 && c.T == 64 && c.CMAX == 64) { FE_LINES_Z(8, 64, 64) }
#undef FE_LINES_Z
    return hipErrorInvalidValue;
}

int64_t edt_sentinel(int nx, int ny, int nz)
{
    return (int64_t)(nx - 1) * (nx - 1) + (int64_t)(ny - 1) * (ny - 1) + (int64_t)(nz - 1) * (nz - 1) + 1;
}

// Tile counts of one volume in the three passes for line tiles of Wy and Wx
// lines (z pass: blocks of ZSUM_ROWS rows for summaries, else of 4 rows).
void vol_tiles(int nx, int ny, int nz, int planes, int Wy, int Wx, bool zsum, uint64_t t[3])
{
    const uint64_t per = zsum ? ZSUM_ROWS : 4;
    t[0] = ((uint64_t)nx * ny + per - 1) / per;
    t[1] = (uint64_t)((nz + Wy - 1) / Wy) * nx * planes;
    t[2] = (uint64_t)((nz + Wx - 1) / Wx) * ny * planes;
}

// z pass (full planes, or summaries when zsum) into `out`, for a table
// (nvol > 0, device table) or one volume (nvol == 0).
void launch_zpass(const uint8_t* masks, int32_t* out, const EdtVol* table, int nvol, const EdtVol& one, int planes,
                  uint64_t tiles, bool zsum, hipStream_t st)
{
    const int site_one = planes == 2 ? 1 : 0;
    const dim3 grid((uint32_t)tiles), block(256);
    if (zsum) {
        const uint32_t gz = round_up(tiles, 512);  // whole groups of 64 blocks on each XCD
        if (planes == 2)
            hipLaunchKernelGGL(k_edt_zsum<true>, dim3(gz), dim3(ZSUM_ROWS), 0, st, masks, out, table, nvol, one,
                               site_one, (uint32_t)tiles);
        else
            hipLaunchKernelGGL(k_edt_zsum<false>, dim3(gz), dim3(ZSUM_ROWS), 0, st, masks, out, table, nvol, one,
                               site_one, (uint32_t)tiles);
    } else {
        if (planes == 2) hipLaunchKernelGGL(k_edt_zpass<true>, grid, block, 0, st, masks, out, table, nvol, one, site_one);
        else hipLaunchKernelGGL(k_edt_zpass<false>, grid, block, 0, st, masks, out, table, nvol, one, site_one);
    }
}

// The three passes over a table (nvol > 0, device table) or one volume (nvol == 0).
// z pass into `out`, y pass into `work`, x pass back into `out`.  planes == 1
// sites: mask == 0; planes == 2, plane 0 sites: mask != 0.  zsum: every
// volume has nz >= ZSUM_MIN_NZ.
hipError_t run_passes(const uint8_t* masks, int32_t* work, int32_t* out, const EdtVol* table, int nvol,
                      const EdtVol& one, int planes, int Lmax_y, int Lmax_x, const uint64_t tiles[3], bool compact,
                      bool zsum, hipStream_t st)
{
    launch_zpass(masks, out, table, nvol, one, planes, tiles[0], zsum, st);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if ((e = launch_lines(out, work, table, nvol, one, 1, Lmax_y, tiles[1], compact, zsum, st)) != hipSuccess)
        return e;
    return launch_lines(work, out, table, nvol, one, 0, Lmax_x, tiles[2], compact, false, st);
}

EdtVol single_vol(int nx, int ny, int nz, int64_t plane, int planes)
{
    EdtVol v{};
    v.planes = planes;
    v.mask = 0;
    v.data = 0;
    v.plane = plane;
    v.nx = nx;
    v.ny = ny;
    v.nz = nz;
    v.sentinel = (int32_t)edt_sentinel(nx, ny, nz);
    return v;
}

}  // namespace

// Largest line length the device passes accept (stack indices are uint16 and
// the widest tile configuration covers 64 chunks of 64 positions).
#define FE_EDT_MAX_AXIS 4096

// Device-resident EDT of `planes` (1 or 2) polarities.  mask: nx*ny*nz bytes
// on the device; work and out: planes*N int32 each; the result lands in out.
// planes == 1: squared distance of every nonzero voxel to the nearest zero
// voxel (FE/edt_native_linear_v1/edt.cu semantics).  planes == 2: plane 0 is
// the squared distance of non-mask voxels to the mask, plane 1 of mask voxels
// to the complement (FE/edt_persistent_pair_v1/edt.cu:32-35 semantics).
extern "C" int fe_edt_device(const uint8_t* mask, int nx, int ny, int nz, int planes, int32_t* work, int32_t* out,
                             void* stream) noexcept
{
    if (!mask || !work || !out || nx < 1 || ny < 1 || nz < 1 || nx > FE_EDT_MAX_AXIS || ny > FE_EDT_MAX_AXIS ||
        nz > FE_EDT_MAX_AXIS || (planes != 1 && planes != 2))
        return FE_INVALID;
    const EdtVol one = single_vol(nx, ny, nz, (int64_t)nx * ny * nz, planes);
    uint64_t tiles[3];
    const bool zsum = nz >= ZSUM_MIN_NZ;
    const bool compact = nx <= 1024 && ny <= 1024 && nz <= 1024;
    vol_tiles(nx, ny, nz, planes, pick_cfg(ny, compact, false).W, pick_cfg(nx, compact, true).W, zsum, tiles);
    FE_TRY(run_passes(mask, work, out, nullptr, 0, one, planes, ny, nx, tiles, compact, zsum,
                      static_cast<hipStream_t>(stream)));
    return FE_OK;
}

// ---------------------------------------------------------------- no-allocation variant
// fe_edt_device with its `work` planes in a file-scope device array instead of
// a caller workspace, for callers that cannot allocate (graph capture, serving
// loops).  The array holds planes * N <= FE_EDT_SCRATCH_WORDS int32 words
// (8 MiB: both planes of the reference EDT's 2^20-voxel cap); every GPU the
// library runs on keeps its own copy until the process ends.  Host code uses
// the per-device address from hipGetSymbolAddress, never the array's name.
// Nothing in it must start at zero (the y pass writes every word it reads
// back).  Calls that overlap on one GPU share the array and race.
#define FE_EDT_SCRATCH_WORDS (1 << 21)
#define FE_MAX_DEVICES 64
static __device__ __align__(256) int32_t scratch[FE_EDT_SCRATCH_WORDS];

namespace {
std::mutex scratch_lock;
int32_t* scratch_ptr[FE_MAX_DEVICES];  // per device, set once with that device current

hipError_t scratch_words(int32_t** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= FE_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> hold(scratch_lock);
    if (!scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        scratch_ptr[dev] = static_cast<int32_t*>(p);
    }
    *out = scratch_ptr[dev];
    return hipSuccess;
}
}  // namespace

extern "C" int fe_edt_device_static(const uint8_t* mask, int nx, int ny, int nz, int planes, int32_t* out,
                                    void* stream) noexcept
{
    if (!mask || !out || nx < 1 || ny < 1 || nz < 1 || nx > FE_EDT_MAX_AXIS || ny > FE_EDT_MAX_AXIS ||
        nz > FE_EDT_MAX_AXIS || (planes != 1 && planes != 2))
        return FE_INVALID;
    if ((int64_t)planes * nx * ny * nz > FE_EDT_SCRATCH_WORDS) {
        snprintf(fe_err_buf, sizeof fe_err_buf, "fe_edt_device_static: %d plane(s) of %lld voxels exceed the %d-word "
                 "scratch", planes, (long long)nx * ny * nz, FE_EDT_SCRATCH_WORDS);
        return FE_INVALID;
    }
    int32_t* work = nullptr;
    FE_TRY(scratch_words(&work));
    return fe_edt_device(mask, nx, ny, nz, planes, work, out, stream);
}

// One axis pass of fe_edt_device, for profiling and per-kernel timing, with
// the intermediate layouts fe_edt_device uses: axis 2 reads the uint8 mask
// `src` and writes the z pass (z summaries when nz >= 16, else int32 planes),
// axis 1 reads that and writes int32 planes, axis 0 reads and writes int32
// planes.
extern "C" int fe_edt_axis_pass(const void* src, int32_t* dst, int nx, int ny, int nz, int planes, int axis,
                                void* stream) noexcept
{
    if (!src || !dst || nx < 1 || ny < 1 || nz < 1 || nx > FE_EDT_MAX_AXIS || ny > FE_EDT_MAX_AXIS ||
        nz > FE_EDT_MAX_AXIS || (planes != 1 && planes != 2) || axis < 0 || axis > 2)
        return FE_INVALID;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const EdtVol one = single_vol(nx, ny, nz, (int64_t)nx * ny * nz, planes);
    uint64_t tiles[3];
    const bool zsum = nz >= ZSUM_MIN_NZ;
    const bool compact = nx <= 1024 && ny <= 1024 && nz <= 1024;
    vol_tiles(nx, ny, nz, planes, pick_cfg(ny, compact, false).W, pick_cfg(nx, compact, true).W, zsum, tiles);
    if (axis == 2) {
        launch_zpass(static_cast<const uint8_t*>(src), dst, nullptr, 0, one, planes, tiles[0], zsum, st);
        FE_TRY(hipGetLastError());
        return FE_OK;
    }
    const int32_t* a = static_cast<const int32_t*>(src);
    if (axis == 1) FE_TRY(launch_lines(a, dst, nullptr, 0, one, 1, ny, tiles[1], compact, zsum, st));
    else FE_TRY(launch_lines(a, dst, nullptr, 0, one, 0, nx, tiles[2], compact, false, st));
    return FE_OK;
}

// ---------------------------------------------------------------- batched volumes
namespace {
struct EdtBatch {
    EdtVol* table = nullptr;  // device
    int nvol = 0, planes = 0, Lmax_y = 0, Lmax_x = 0;
    uint64_t tiles[3] = {0, 0, 0};
    bool compact = true, zsum = true;
};
}  // namespace

// A batch of nvol volumes sharing one launch per axis pass.  dims: host int32
// [nvol][3]; mask_off: byte offset of each mask in the mask arena (mask_bytes
// long); data_off / plane_stride: element offset of each volume's plane 0 and
// the distance between its planes in the int32 arenas (data_elems long; out
// and work use the same layout).  Volumes' plane ranges must not overlap.
// Offsets that are multiples of 4 (and nz % 4 == 0) take the 16-byte vector
// path.  Returns NULL with *status 1 (invalid) or 2 (runtime).
extern "C" void* fe_edt_batch_create(const int32_t* dims, const int64_t* mask_off, const int64_t* data_off,
                                     const int64_t* plane_stride, int nvol, int planes, int64_t mask_bytes,
                                     int64_t data_elems, int* status) noexcept
{
    int dummy = 0;
    int& st = status ? *status : dummy;
    st = FE_INVALID;
    if (!dims || !mask_off || !data_off || !plane_stride || nvol < 1 || nvol > (1 << 20) ||
        (planes != 1 && planes != 2) || mask_bytes < 1 || data_elems < 1)
        return nullptr;
    EdtBatch* b = nullptr;
    try {
        std::vector<EdtVol> host((size_t)nvol);
        std::vector<std::pair<int64_t, int64_t>> ranges((size_t)nvol);
        int Ly = 1, Lx = 1, Lz = 1, zmin = FE_EDT_MAX_AXIS;
        for (int v = 0; v < nvol; ++v) {
            const int nx = dims[3 * v], ny = dims[3 * v + 1], nz = dims[3 * v + 2];
            if (nx < 1 || ny < 1 || nz < 1 || nx > FE_EDT_MAX_AXIS || ny > FE_EDT_MAX_AXIS || nz > FE_EDT_MAX_AXIS)
                return nullptr;
            const int64_t n = (int64_t)nx * ny * nz;
            if (mask_off[v] < 0 || mask_off[v] > mask_bytes - n || data_off[v] < 0 || plane_stride[v] < n ||
                data_off[v] + (planes - 1) * plane_stride[v] > data_elems - n)
                return nullptr;
            host[v] = single_vol(nx, ny, nz, plane_stride[v], planes);
            host[v].mask = mask_off[v];
            host[v].data = data_off[v];
            ranges[v] = {data_off[v], data_off[v] + (planes - 1) * plane_stride[v] + n};
            Ly = std::max(Ly, ny);
            Lx = std::max(Lx, nx);
            Lz = std::max(Lz, nz);
            zmin = std::min(zmin, nz);
        }
        std::sort(ranges.begin(), ranges.end());
        for (size_t i = 1; i < ranges.size(); ++i)
            if (ranges[i].first < ranges[i - 1].second) return nullptr;
        const bool compact = Lx <= 1024 && Ly <= 1024 && Lz <= 1024;
        const int Wy = pick_cfg(Ly, compact, false).W, Wx = pick_cfg(Lx, compact, true).W;
        const bool zsum = zmin >= ZSUM_MIN_NZ;
        uint64_t tot[3] = {0, 0, 0};
        for (int v = 0; v < nvol; ++v) {
            uint64_t t[3];
            vol_tiles(host[v].nx, host[v].ny, host[v].nz, planes, Wy, Wx, zsum, t);
            for (int p = 0; p < 3; ++p) {
                host[v].first[p] = (uint32_t)tot[p];
                tot[p] += t[p];
            }
        }
        for (int p = 0; p < 3; ++p)
            if (tot[p] >= 0x7fffffffull) return nullptr;
        st = FE_RUNTIME;
        b = new EdtBatch();
        b->nvol = nvol;
        b->planes = planes;
        b->Lmax_y = Ly;
        b->Lmax_x = Lx;
        b->compact = compact;
        b->zsum = zsum;
        for (int p = 0; p < 3; ++p) b->tiles[p] = tot[p];
        hipError_t e = hipMalloc(&b->table, sizeof(EdtVol) * (size_t)nvol);
        if (e == hipSuccess) e = hipMemcpy(b->table, host.data(), sizeof(EdtVol) * (size_t)nvol, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            fe_note_error(e, "fe_edt_batch_create");
            if (b->table) (void)hipFree(b->table);
            delete b;
            return nullptr;
        }
        st = FE_OK;
        return b;
    } catch (...) {
        if (b) {
            if (b->table) (void)hipFree(b->table);
            delete b;
        }
        st = FE_RUNTIME;
        return nullptr;
    }
}

// The three passes of every volume of the batch: three kernel launches on
// `stream`, no allocation, copy or synchronisation (capturable in a graph).
extern "C" int fe_edt_batch_run(const void* handle, const uint8_t* masks, int32_t* work, int32_t* out,
                                void* stream) noexcept
{
    const auto* b = static_cast<const EdtBatch*>(handle);
    if (!b || !masks || !work || !out) return FE_INVALID;
    const EdtVol none{};
    FE_TRY(run_passes(masks, work, out, b->table, b->nvol, none, b->planes, b->Lmax_y, b->Lmax_x, b->tiles,
                      b->compact, b->zsum, static_cast<hipStream_t>(stream)));
    return FE_OK;
}

extern "C" void fe_edt_batch_destroy(void* handle) noexcept
{
    auto* b = static_cast<EdtBatch*>(handle);
    if (!b) return;
    if (b->table) (void)hipFree(b->table);
    delete b;
}

// Signed distance from a device pair (2N int32) into N float32.
extern "C" int fe_edt_signed_device(const int32_t* pair, int64_t n, double pitch, float* out, void* stream) noexcept
{
    if (!pair || !out || n < 1 || !std::isfinite(pitch)) return FE_INVALID;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_edt_signed, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, pair, out, n, pitch);
    FE_TRY(hipGetLastError());
    return FE_OK;
}

// ---------------------------------------------------------------- reference-shaped host ABI
namespace {

// Host validation of FE/edt_native_v1/edt.cu:64-69: bytes <= 1, both classes present.
int validate_mask(const uint8_t* mask, int64_t n)
{
    int64_t ones = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (mask[i] > 1) return FE_INVALID;
        ones += mask[i];
    }
    return (ones == 0 || ones == n) ? FE_INVALID : FE_OK;
}

// Plain per-call squared EDT: allocate, upload, three passes, download, free.
int edt_squared_capped(const uint8_t* mask, int nx, int ny, int nz, int32_t* output, int axis_cap, int64_t size_cap)
{
    if (!mask || !output || nx < 1 || ny < 1 || nz < 1 || nx > axis_cap || ny > axis_cap || nz > axis_cap) return FE_INVALID;
    const int64_t n = (int64_t)nx * ny * nz;
    if (n > size_cap) return FE_INVALID;
    if (validate_mask(mask, n)) return FE_INVALID;
    uint8_t* dm = nullptr;
    int32_t *a = nullptr, *b = nullptr;
    int status = FE_RUNTIME;
    hipError_t e = hipSuccess;
    do {
        if ((e = hipMalloc(&dm, (size_t)n)) != hipSuccess) break;
        if ((e = hipMalloc(&a, (size_t)n * 4)) != hipSuccess) break;
        if ((e = hipMalloc(&b, (size_t)n * 4)) != hipSuccess) break;
        if ((e = hipMemcpy(dm, mask, (size_t)n, hipMemcpyHostToDevice)) != hipSuccess) break;
        if (fe_edt_device(dm, nx, ny, nz, 1, b, a, nullptr) != FE_OK) break;
        if ((e = hipMemcpy(output, a, (size_t)n * 4, hipMemcpyDeviceToHost)) != hipSuccess) break;
        status = FE_OK;
    } while (0);
    if (e != hipSuccess) fe_note_error(e, "edt per-call path");
    if (dm && hipFree(dm) != hipSuccess) status = FE_RUNTIME;
    if (a && hipFree(a) != hipSuccess) status = FE_RUNTIME;
    if (b && hipFree(b) != hipSuccess) status = FE_RUNTIME;
    return status;
}

struct PairEDT {
    int nx, ny, nz;
    int64_t size;
    uint8_t* mask = nullptr;
    int32_t* a = nullptr;
    int32_t* b = nullptr;
    float* distance = nullptr;
    hipStream_t stream = nullptr;
    bool failed = false;
};

}  // namespace

// FE/edt_native_v1/edt.cu:18-48 contract (<=128 per axis, <=262144 voxels).
extern "C" int edt_squared(const uint8_t* mask, int nx, int ny, int nz, int32_t* output) noexcept
{
    try {
        return edt_squared_capped(mask, nx, ny, nz, output, 128, 262144);
    } catch (...) {
        return FE_RUNTIME;
    }
}

// FE/edt_native_large_v1/edt.cu:20 and FE/edt_native_linear_v1/edt.cu:33 contract
// (<=512 per axis, <=1048576 voxels).
extern "C" int edt_squared_large(const uint8_t* mask, int nx, int ny, int nz, int32_t* output) noexcept
{
    try {
        return edt_squared_capped(mask, nx, ny, nz, output, 512, 1048576);
    } catch (...) {
        return FE_RUNTIME;
    }
}

// Plain per-call squared EDT without the reference's size caps (axes <= 4096).
extern "C" int fe_edt_squared_host(const uint8_t* mask, int nx, int ny, int nz, int32_t* output) noexcept
{
    try {
        return edt_squared_capped(mask, nx, ny, nz, output, FE_EDT_MAX_AXIS, INT64_MAX);
    } catch (...) {
        return FE_RUNTIME;
    }
}

extern "C" void edt_pair_destroy(void* pointer) noexcept
{
    auto* p = static_cast<PairEDT*>(pointer);
    if (!p) return;
    if (p->mask) (void)hipFree(p->mask);
    if (p->a) (void)hipFree(p->a);
    if (p->b) (void)hipFree(p->b);
    if (p->distance) (void)hipFree(p->distance);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

static void* pair_create(int nx, int ny, int nz, int axis_cap, int64_t size_cap) noexcept
{
    if (nx < 1 || ny < 1 || nz < 1 || nx > axis_cap || ny > axis_cap || nz > axis_cap ||
        (int64_t)nx * ny * nz > size_cap)
        return nullptr;
    PairEDT* p = nullptr;
    try {
        p = new PairEDT();
        p->nx = nx;
        p->ny = ny;
        p->nz = nz;
        p->size = (int64_t)nx * ny * nz;
        if (hipMalloc(&p->mask, (size_t)p->size) != hipSuccess || hipMalloc(&p->a, (size_t)p->size * 8) != hipSuccess ||
            hipMalloc(&p->b, (size_t)p->size * 8) != hipSuccess ||
            hipMalloc(&p->distance, (size_t)p->size * 4) != hipSuccess ||
            hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
            edt_pair_destroy(p);
            return nullptr;
        }
        return p;
    } catch (...) {
        edt_pair_destroy(p);
        return nullptr;
    }
}

// FE/edt_persistent_pair_v1/edt.cu:48-58 contract (<=512 per axis, <=1048576 voxels).
extern "C" void* edt_pair_create(int nx, int ny, int nz) noexcept { return pair_create(nx, ny, nz, 512, 1048576); }

// Same persistent pair without the reference's size caps (axes <= 4096).
extern "C" void* fe_edt_pair_create(int nx, int ny, int nz) noexcept
{
    return pair_create(nx, ny, nz, FE_EDT_MAX_AXIS, INT64_MAX);
}

// FE/edt_persistent_pair_v1/edt.cu:59-78: validate, upload, paired passes,
// signed output on the device, download; status 2 latches the handle.
extern "C" int edt_pair_query(void* pointer, const uint8_t* mask, double pitch, float* output, int32_t* squared) noexcept
{
    auto* p = static_cast<PairEDT*>(pointer);
    if (!p || !mask || !output || !std::isfinite(pitch) || pitch < 1e-12 || pitch > 1e12) return FE_INVALID;
    if (p->failed) return FE_RUNTIME;
    if (validate_mask(mask, p->size)) return FE_INVALID;
#define FE_CHECK(x)                  \
    if ((x) != hipSuccess) {         \
        p->failed = true;            \
        return FE_RUNTIME;           \
    }
    FE_CHECK(hipMemcpyAsync(p->mask, mask, (size_t)p->size, hipMemcpyHostToDevice, p->stream));
    if (fe_edt_device(p->mask, p->nx, p->ny, p->nz, 2, p->b, p->a, p->stream) != FE_OK) {
        p->failed = true;
        return FE_RUNTIME;
    }
    if (fe_edt_signed_device(p->a, p->size, pitch, p->distance, p->stream) != FE_OK) {
        p->failed = true;
        return FE_RUNTIME;
    }
    FE_CHECK(hipMemcpyAsync(output, p->distance, (size_t)p->size * 4, hipMemcpyDeviceToHost, p->stream));
    if (squared) FE_CHECK(hipMemcpyAsync(squared, p->a, (size_t)p->size * 8, hipMemcpyDeviceToHost, p->stream));
    FE_CHECK(hipStreamSynchronize(p->stream));
#undef FE_CHECK
    return FE_OK;
}
