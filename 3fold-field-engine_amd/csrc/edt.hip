// Exact integer squared Euclidean distance transform for gfx950.
//
// Mathematics (the reference's, FE/field_edt_integer_v1.py:14-26): with
// f = 0 on site voxels and f = sentinel = sum (n-1)^2 + 1 elsewhere, the
// result is min over s of f(s) + |p - s|^2, computed one axis at a time.  The
// per-axis minimum is exact in integers, so the axis order does not change
// the result; this implementation runs z first (the contiguous axis, straight
// from the uint8 mask) and then y and x (strided axes, LDS-staged tiles).
//
//   z pass  one wave per row: 64-bit ballots of the site flags, nearest site
//           left/right from clz/ctz, squared 1-D distance.  Both polarities of
//           a pair come from one read of the mask.
//   y,x     W adjacent lines (contiguous along z, so every row of the tile is
//           one coalesced segment) are staged in LDS; T threads share a line,
//           each owns a chunk of positions, builds the exact lower envelope
//           of its chunk's parabolas (Felzenszwalb, pop test by int64 cross
//           multiplication, no division) and then evaluates the envelopes of
//           its own and neighbouring chunks at its positions, visiting a
//           neighbour only while its lower bound min f + gap^2 can still win.
//   signed  FE/edt_persistent_pair_v1/edt.cu:36-42 semantics with IEEE
//           correctly rounded f64 sqrt/mul and f32 sub/mul.
#include "fe_common.hpp"
#include <algorithm>
#include <climits>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

static thread_local char fe_err_buf[256];

void fe_note_error(hipError_t e, const char* where)
{
    snprintf(fe_err_buf, sizeof fe_err_buf, "%s: %s (%d)", where, hipGetErrorString(e), (int)e);
}

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
    uint32_t pad;
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
// Four rows (one per wave) per block; sites of plane 0 are voxels whose mask
// equals site_one (1) or is zero (0); plane 1 (PAIR) takes the complement.
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
    const uint8_t* m = masks + V.mask + row * nz;
    const int nseg = (nz + 63) >> 6;
    uint64_t w0 = 0, valid = 0;
    for (int s = 0; s < nseg; ++s) {
        int p = (s << 6) + lane;
        bool in = p < nz;
        bool hit = in && m[p] != 0;
        uint64_t b = __ballot(site_one ? hit : (in && !hit));
        uint64_t v = __ballot(in);
        if (lane == s) { w0 = b; valid = v; }
    }
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
// Parabola value of site s at position p over the LDS line f.
__device__ __forceinline__ int32_t pval(const int32_t* f, int s, int p) { return f[s] + (p - s) * (p - s); }

// Site b stays on the lower envelope of sites a < b < c iff the parabola of b
// is strictly lowest somewhere, i.e. x(b,c) > x(a,b) for the real
// intersections x; evaluated exactly in int64 without division.
__device__ __forceinline__ bool keep_mid(const int32_t* f, int a, int b, int c)
{
    const int64_t Fa = (int64_t)f[a] + (int64_t)a * a, Fb = (int64_t)f[b] + (int64_t)b * b,
                  Fc = (int64_t)f[c] + (int64_t)c * c;
    return (Fc - Fb) * (int64_t)(b - a) > (Fb - Fa) * (int64_t)(c - b);
}

// One exact lower-envelope pass over W lines x L positions.  Thread (c, line)
// owns chunk c (C positions) of a line.
//  1. every thread builds the lower hull of the points (s, f(s)+s^2) of its
//     chunk (the parabola envelope), keeping the two top entries in registers;
//  2. the chunk hulls of a line are merged in log2(T) rounds: in round h the
//     thread at the boundary of two adjacent groups pops the tail of the left
//     group and the head of the right group until the bridge between them is
//     found (merging lower hulls of x-separated point sets);
//  3. every thread walks the merged envelope over its positions (the minimum
//     is unimodal along the envelope and its site moves right with p);
//     results go back through LDS so that loads and stores are 16-byte vectors
//     along the contiguous z axis.
// A line whose values are all equal is its own result (g = f).  Sites holding
// the sentinel (no site reached yet) are left out of the hulls: such a site
// contributes sentinel + (p-s)^2 >= sentinel, with equality only at its own
// position, so the result is min(envelope of the other sites, sentinel).
// axis 1: lines along y (outer index x); axis 0: lines along x (outer index y).
template <int W, int T, int CMAX>
__global__ __launch_bounds__(W* T) void k_edt_lines(const int32_t* __restrict__ src, int32_t* __restrict__ dst,
                                                    const EdtVol* __restrict__ table, int nvol, EdtVol one, int axis,
                                                    uint32_t tiles, uint32_t grid)
{
    extern __shared__ int32_t lds[];
    const uint32_t tile = xcd_tile(blockIdx.x, grid);
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
    const int outer_count = axis == 1 ? V.nx : V.ny;
    const int ztiles = (nz + W - 1) / W;
    const int zt = local % ztiles;
    const int rest = local / ztiles;
    const int outer = rest % outer_count, plane = rest / outer_count;
    const int64_t base = V.data + (int64_t)plane * V.plane + (int64_t)outer * outer_stride + (int64_t)zt * W;
    const int nlines = min(W, nz - zt * W);
    const int LP = L + 1;
    constexpr int NT = W * T;
    int32_t* f = lds;                                          // [W][LP]
    uint16_t* stk = reinterpret_cast<uint16_t*>(f + W * LP);  // [CMAX][NT]
    uint16_t* blo = stk + CMAX * NT;                           // [NT] surviving range of a chunk hull
    uint16_t* bhi = blo + NT;                                  // [NT]
    uint8_t* lsb = reinterpret_cast<uint8_t*>(bhi + NT);       // [NT] nearest nonempty chunk <= c
    uint8_t* rsb = lsb + NT;                                   // [NT] nearest nonempty chunk >= c
    int32_t* lval = reinterpret_cast<int32_t*>(rsb + NT);      // [W] value of chunk 0's first sample
    int32_t* lflat = lval + W;                                 // [W] line is constant
    const int tid = threadIdx.x;
    // rows of the tile are 16-byte aligned when nz and the volume's offsets are multiples of 4
    const bool vec = ((nz | V.data | V.plane) & 3) == 0;

    if (vec) {
        for (int i = tid; i < (W / 4) * L; i += NT) {
            const int l0 = (i % (W / 4)) * 4, q = i / (W / 4);
            if (l0 < nlines) {
                const int4 v = *reinterpret_cast<const int4*>(src + base + (int64_t)q * axis_stride + l0);
                f[l0 * LP + q] = v.x;
                f[(l0 + 1) * LP + q] = v.y;
                f[(l0 + 2) * LP + q] = v.z;
                f[(l0 + 3) * LP + q] = v.w;
            }
        }
    } else {
        for (int i = tid; i < W * L; i += NT) {
            const int line = i % W, q = i / W;
            if (line < nlines) f[line * LP + q] = src[base + (int64_t)q * axis_stride + line];
        }
    }
    __syncthreads();

    const int line = tid % W, c = tid / W;
    const int C = (L + T - 1) / T;
    const int p0 = min(c * C, L), p1 = min(p0 + C, L);
    const bool active = line < nlines;
    const int32_t* fl = f + line * LP;

    // 1. hull of this chunk
    int k = -1, s_top = 0, s_sec = 0;
    int64_t F_top = 0, F_sec = 0;
    int32_t mn = INT_MAX, mx = INT_MIN;
    if (active) {
        for (int q = p0; q < p1; ++q) {
            const int32_t fq = fl[q];
            mn = min(mn, fq);
            mx = max(mx, fq);
            if (fq >= sentinel) continue;
            const int64_t Fq = (int64_t)fq + (int64_t)q * q;
            while (k >= 1) {
                // top stays iff it is strictly lowest somewhere between sec and q
                if ((Fq - F_top) * (int64_t)(s_top - s_sec) > (F_top - F_sec) * (int64_t)(q - s_top)) break;
                --k;
                s_top = s_sec;
                F_top = F_sec;
                if (k >= 1) {
                    s_sec = stk[(k - 1) * NT + tid];
                    F_sec = (int64_t)fl[s_sec] + (int64_t)s_sec * s_sec;
                }
            }
            ++k;
            stk[k * NT + tid] = (uint16_t)q;
            s_sec = s_top;
            F_sec = F_top;
            s_top = q;
            F_top = Fq;
        }
    }
    blo[tid] = 0;
    bhi[tid] = (uint16_t)(k + 1);
    if (c == 0) {
        lval[line] = mn;
        lflat[line] = 1;
    }
    __syncthreads();
    if (active && p0 < p1 && !(mn == mx && mn == lval[line])) lflat[line] = 0;

    // 2. tree merge of the chunk hulls
    for (int h = 1; h < T; h <<= 1) {
        if (active && (c & (2 * h - 1)) == h) {
            const int a0 = c - h, a2 = min(c + h, T);
            int ta = c - 1;
            while (ta >= a0 && bhi[ta * W + line] == blo[ta * W + line]) --ta;
            int hb = c;
            while (hb < a2 && bhi[hb * W + line] == blo[hb * W + line]) ++hb;
            if (ta >= a0 && hb < a2) {
                bool changed = true;
                while (changed) {
                    changed = false;
                    for (;;) {  // tail of the left group
                        const int tt = ta * W + line;
                        const int th = bhi[tt];
                        const int s1 = stk[(th - 1) * NT + tt];
                        int s2, pa = ta;
                        if (th - blo[tt] >= 2) s2 = stk[(th - 2) * NT + tt];
                        else {
                            pa = ta - 1;
                            while (pa >= a0 && bhi[pa * W + line] == blo[pa * W + line]) --pa;
                            if (pa < a0) break;
                            s2 = stk[(bhi[pa * W + line] - 1) * NT + pa * W + line];
                        }
                        const int hbt = hb * W + line;
                        const int s3 = stk[blo[hbt] * NT + hbt];
                        if (keep_mid(fl, s2, s1, s3)) break;
                        bhi[tt] = (uint16_t)(th - 1);
                        if (th - 1 == blo[tt]) ta = pa;
                    }
                    for (;;) {  // head of the right group
                        const int ht = hb * W + line;
                        const int hl = blo[ht];
                        const int s1 = stk[hl * NT + ht];
                        int s2, nb = hb;
                        if (bhi[ht] - hl >= 2) s2 = stk[(hl + 1) * NT + ht];
                        else {
                            nb = hb + 1;
                            while (nb < a2 && bhi[nb * W + line] == blo[nb * W + line]) ++nb;
                            if (nb >= a2) break;
                            s2 = stk[blo[nb * W + line] * NT + nb * W + line];
                        }
                        const int tt = ta * W + line;
                        const int s0 = stk[(bhi[tt] - 1) * NT + tt];
                        if (keep_mid(fl, s0, s1, s2)) break;
                        blo[ht] = (uint16_t)(hl + 1);
                        changed = true;
                        if (hl + 1 == bhi[ht]) hb = nb;
                    }
                }
            }
        }
        __syncthreads();
    }
    {
        int b = c;
        while (b >= 0 && bhi[b * W + line] == blo[b * W + line]) --b;
        lsb[tid] = (uint8_t)(b < 0 ? 255 : b);
        b = c;
        while (b < T && bhi[b * W + line] == blo[b * W + line]) ++b;
        rsb[tid] = (uint8_t)(b >= T ? 255 : b);
    }
    __syncthreads();

    // 3. walk the merged envelope
    int32_t out[CMAX];
    if (active && p0 < p1) {
        if (lflat[line]) {
#pragma unroll
            for (int i = 0; i < CMAX; ++i)
                if (p0 + i < p1) out[i] = fl[p0 + i];
        } else {
            int cb = (bhi[tid] > blo[tid]) ? c : (lsb[tid] != 255 ? lsb[tid] : rsb[tid]);
            int ct = cb * W + line;
            int ci = cb == c ? blo[ct] : (cb < c ? bhi[ct] - 1 : blo[ct]);
            int cs = stk[ci * NT + ct];
            int32_t cf = fl[cs];
            {
                int32_t cv = cf + (p0 - cs) * (p0 - cs);
                bool moved = false;
                for (;;) {  // right while not worse
                    int nb = cb, ni = ci + 1;
                    if (ni >= bhi[ct]) {
                        if (cb + 1 >= T) break;
                        nb = rsb[(cb + 1) * W + line];
                        if (nb == 255) break;
                        ni = blo[nb * W + line];
                    }
                    const int nt = nb * W + line;
                    const int ns = stk[ni * NT + nt];
                    const int32_t nf = fl[ns];
                    const int32_t nv = nf + (p0 - ns) * (p0 - ns);
                    if (nv > cv) break;
                    cb = nb; ct = nt; ci = ni; cs = ns; cf = nf; cv = nv; moved = true;
                }
                if (!moved) {
                    for (;;) {  // left while strictly better
                        int pb = cb, pi = ci - 1;
                        if (pi < blo[ct]) {
                            if (cb == 0) break;
                            pb = lsb[(cb - 1) * W + line];
                            if (pb == 255) break;
                            pi = bhi[pb * W + line] - 1;
                        }
                        const int pt = pb * W + line;
                        const int ps = stk[pi * NT + pt];
                        const int32_t pf = fl[ps];
                        const int32_t pv = pf + (p0 - ps) * (p0 - ps);
                        if (pv >= cv) break;
                        cb = pb; ct = pt; ci = pi; cs = ps; cf = pf; cv = pv;
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < CMAX; ++i) {
                const int p = p0 + i;
                if (p < p1) {
                    int32_t cv = cf + (p - cs) * (p - cs);
                    for (;;) {
                        int nb = cb, ni = ci + 1;
                        if (ni >= bhi[ct]) {
                            if (cb + 1 >= T) break;
                            nb = rsb[(cb + 1) * W + line];
                            if (nb == 255) break;
                            ni = blo[nb * W + line];
                        }
                        const int nt = nb * W + line;
                        const int ns = stk[ni * NT + nt];
                        const int32_t nf = fl[ns];
                        const int32_t nv = nf + (p - ns) * (p - ns);
                        if (nv > cv) break;
                        cb = nb; ct = nt; ci = ni; cs = ns; cf = nf; cv = nv;
                    }
                    out[i] = min(cv, sentinel);
                }
            }
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int i = 0; i < CMAX; ++i)
            if (p0 + i < p1) f[line * LP + p0 + i] = out[i];
    }
    __syncthreads();
    if (vec) {
        for (int i = tid; i < (W / 4) * L; i += NT) {
            const int l0 = (i % (W / 4)) * 4, q = i / (W / 4);
            if (l0 < nlines)
                *reinterpret_cast<int4*>(dst + base + (int64_t)q * axis_stride + l0) =
                    make_int4(f[l0 * LP + q], f[(l0 + 1) * LP + q], f[(l0 + 2) * LP + q], f[(l0 + 3) * LP + q]);
        }
    } else {
        for (int i = tid; i < W * L; i += NT) {
            const int ln = i % W, q = i / W;
            if (ln < nlines) dst[base + (int64_t)q * axis_stride + ln] = f[ln * LP + q];
        }
    }
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
    out[i] = d - sg * half;
}

struct LineCfg {
    int W, T, CMAX;
};

// Tile shape by line length (W lines, T threads per line, CMAX >= L/T): LDS
// holds W*(L+1) int32 values plus CMAX*W*T uint16 hull slots.  Measured on
// MI355X at L = 512: 16 lines x 32 threads beats 16 x 16 and 16 x 64 (fewer
// merge rounds vs. shorter serial chunks) and 32-line tiles (LDS occupancy).
LineCfg pick_cfg(int L)
{
    if (L <= 128) return {32, 8, 16};
    if (L <= 512) return {16, 32, 16};
    if (L <= 1024) return {8, 32, 32};
    return {4, 64, 64};
}

template <int W, int T, int CMAX>
hipError_t launch_lines_t(const int32_t* src, int32_t* dst, const EdtVol* table, int nvol, const EdtVol& one,
                          int axis, int Lmax, uint64_t tiles, hipStream_t st)
{
    const uint32_t grid = round8(tiles);
    const size_t lds = (size_t)W * (Lmax + 1) * 4 + (size_t)CMAX * W * T * 2 + (size_t)W * T * 6 + (size_t)W * 8;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_edt_lines<W, T, CMAX>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_edt_lines<W, T, CMAX>), dim3(grid), dim3(W * T), lds, st, src, dst, table, nvol, one, axis,
                       (uint32_t)tiles, grid);
    return hipGetLastError();
}

// One line pass over every volume: Lmax (the longest line of the pass over
// all volumes) picks the tile shape, `tiles` counts the tiles of that shape.
hipError_t launch_lines(const int32_t* src, int32_t* dst, const EdtVol* table, int nvol, const EdtVol& one, int axis,
                        int Lmax, uint64_t tiles, hipStream_t st)
{
    LineCfg c = pick_cfg(Lmax);
    if (c.W == 32) return launch_lines_t<32, 8, 16>(src, dst, table, nvol, one, axis, Lmax, tiles, st);
    if (c.W == 16) return launch_lines_t<16, 32, 16>(src, dst, table, nvol, one, axis, Lmax, tiles, st);
    if (c.W == 8) return launch_lines_t<8, 32, 32>(src, dst, table, nvol, one, axis, Lmax, tiles, st);
    return launch_lines_t<4, 64, 64>(src, dst, table, nvol, one, axis, Lmax, tiles, st);
}

int64_t edt_sentinel(int nx, int ny, int nz)
{
    return (int64_t)(nx - 1) * (nx - 1) + (int64_t)(ny - 1) * (ny - 1) + (int64_t)(nz - 1) * (nz - 1) + 1;
}

// Tile counts of one volume in the three passes for line tiles of Wy and Wx lines.
void vol_tiles(int nx, int ny, int nz, int planes, int Wy, int Wx, uint64_t t[3])
{
    t[0] = ((uint64_t)nx * ny + 3) / 4;
    t[1] = (uint64_t)((nz + Wy - 1) / Wy) * nx * planes;
    t[2] = (uint64_t)((nz + Wx - 1) / Wx) * ny * planes;
}

// The three passes over a table (nvol > 0, device table) or one volume (nvol == 0).
// z pass into `out`, y pass into `work`, x pass back into `out`.  planes == 1
// sites: mask == 0; planes == 2, plane 0 sites: mask != 0.
hipError_t run_passes(const uint8_t* masks, int32_t* work, int32_t* out, const EdtVol* table, int nvol,
                      const EdtVol& one, int planes, int Lmax_y, int Lmax_x, const uint64_t tiles[3], hipStream_t st)
{
    const int site_one = planes == 2 ? 1 : 0;
    if (planes == 2)
        hipLaunchKernelGGL(k_edt_zpass<true>, dim3((uint32_t)tiles[0]), dim3(256), 0, st, masks, out, table, nvol, one,
                           site_one);
    else
        hipLaunchKernelGGL(k_edt_zpass<false>, dim3((uint32_t)tiles[0]), dim3(256), 0, st, masks, out, table, nvol,
                           one, site_one);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if ((e = launch_lines(out, work, table, nvol, one, 1, Lmax_y, tiles[1], st)) != hipSuccess) return e;
    return launch_lines(work, out, table, nvol, one, 0, Lmax_x, tiles[2], st);
}

EdtVol single_vol(int nx, int ny, int nz, int64_t plane)
{
    EdtVol v{};
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
    const EdtVol one = single_vol(nx, ny, nz, (int64_t)nx * ny * nz);
    uint64_t tiles[3];
    vol_tiles(nx, ny, nz, planes, pick_cfg(ny).W, pick_cfg(nx).W, tiles);
    FE_TRY(run_passes(mask, work, out, nullptr, 0, one, planes, ny, nx, tiles, static_cast<hipStream_t>(stream)));
    return FE_OK;
}

// One axis pass of fe_edt_device, for profiling and per-kernel timing:
// axis 2 reads the uint8 mask `src` (z pass), axes 1 and 0 read int32 planes.
extern "C" int fe_edt_axis_pass(const void* src, int32_t* dst, int nx, int ny, int nz, int planes, int axis,
                                void* stream) noexcept
{
    if (!src || !dst || nx < 1 || ny < 1 || nz < 1 || nx > FE_EDT_MAX_AXIS || ny > FE_EDT_MAX_AXIS ||
        nz > FE_EDT_MAX_AXIS || (planes != 1 && planes != 2) || axis < 0 || axis > 2)
        return FE_INVALID;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const EdtVol one = single_vol(nx, ny, nz, (int64_t)nx * ny * nz);
    uint64_t tiles[3];
    vol_tiles(nx, ny, nz, planes, pick_cfg(ny).W, pick_cfg(nx).W, tiles);
    if (axis == 2) {
        const uint8_t* m = static_cast<const uint8_t*>(src);
        if (planes == 2)
            hipLaunchKernelGGL(k_edt_zpass<true>, dim3((uint32_t)tiles[0]), dim3(256), 0, st, m, dst, nullptr, 0, one, 1);
        else
            hipLaunchKernelGGL(k_edt_zpass<false>, dim3((uint32_t)tiles[0]), dim3(256), 0, st, m, dst, nullptr, 0, one, 0);
        FE_TRY(hipGetLastError());
        return FE_OK;
    }
    const int32_t* a = static_cast<const int32_t*>(src);
    if (axis == 1) FE_TRY(launch_lines(a, dst, nullptr, 0, one, 1, ny, tiles[1], st));
    else FE_TRY(launch_lines(a, dst, nullptr, 0, one, 0, nx, tiles[2], st));
    return FE_OK;
}

// ---------------------------------------------------------------- batched volumes
namespace {
struct EdtBatch {
    EdtVol* table = nullptr;  // device
    int nvol = 0, planes = 0, Lmax_y = 0, Lmax_x = 0;
    uint64_t tiles[3] = {0, 0, 0};
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
        int Ly = 1, Lx = 1;
        for (int v = 0; v < nvol; ++v) {
            const int nx = dims[3 * v], ny = dims[3 * v + 1], nz = dims[3 * v + 2];
            if (nx < 1 || ny < 1 || nz < 1 || nx > FE_EDT_MAX_AXIS || ny > FE_EDT_MAX_AXIS || nz > FE_EDT_MAX_AXIS)
                return nullptr;
            const int64_t n = (int64_t)nx * ny * nz;
            if (mask_off[v] < 0 || mask_off[v] > mask_bytes - n || data_off[v] < 0 || plane_stride[v] < n ||
                data_off[v] + (planes - 1) * plane_stride[v] > data_elems - n)
                return nullptr;
            host[v] = single_vol(nx, ny, nz, plane_stride[v]);
            host[v].mask = mask_off[v];
            host[v].data = data_off[v];
            ranges[v] = {data_off[v], data_off[v] + (planes - 1) * plane_stride[v] + n};
            Ly = std::max(Ly, ny);
            Lx = std::max(Lx, nx);
        }
        std::sort(ranges.begin(), ranges.end());
        for (size_t i = 1; i < ranges.size(); ++i)
            if (ranges[i].first < ranges[i - 1].second) return nullptr;
        const int Wy = pick_cfg(Ly).W, Wx = pick_cfg(Lx).W;
        uint64_t tot[3] = {0, 0, 0};
        for (int v = 0; v < nvol; ++v) {
            uint64_t t[3];
            vol_tiles(host[v].nx, host[v].ny, host[v].nz, planes, Wy, Wx, t);
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
                      static_cast<hipStream_t>(stream)));
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
