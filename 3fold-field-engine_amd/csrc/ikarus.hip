// IKARUS expression trees on gfx950: an fp32 interpreter for the compiled
// post-order program of a tree (sphere, box, capped cylinder, cone, halfspace,
// rectangular frustum and trilinear grid leaves; affine transforms, axis
// projection, uniform scale, offset; hard and smooth union / intersection /
// difference), and the block-sparse build of a field from one (eight-corner
// classification, active-block tiles).
//
// The reference generates one kernel per tree with the parameters baked in as
// fp32 literals (FE/ikarus_v1/eval_warp.py:160-320, eval_warp_uniform_v2.py:
// 100-270).  Here one kernel serves every tree: the host compiles the tree to
// 16-byte instructions and an fp32 parameter vector, every lane of a wave runs
// the same instruction (the opcode and parameter reads are wave-uniform, so
// they are scalar loads and the dispatch does not diverge), the current point
// lives in registers and the value stack in LDS with its top in a register.
// Each leaf re-applies its ancestors' point maps from the world point, in the
// order the recursive evaluation applies them, so every fp32 operation rounds
// as in the per-node functions; operations are the reference's, in its order,
// without contraction (the library is built with -ffp-contract=off), with
// Warp's min/max (a<b?a:b, a>b?a:b), clamp (min(max(lo, x), hi)) and lerp
// (a*(1-t) + b*t).
#include "fe_common.hpp"
#include <cstdio>
#include <mutex>

namespace {

constexpr int IK_T = 256;      // threads per block
constexpr int IK_STACK = 16;   // value-stack slots per thread (LDS) below the register top

// Opcodes (the host compiler in ikarus.py emits the same numbers).
enum : int32_t {
    IK_WORLD = 0,      // current point := world point
    IK_XFORM = 1,      // p := M p, M = 3 x 4 inverse transform (12 params)
    IK_PROJ = 2,       // p[a] := ref (1 param)
    IK_SCALE_IN = 3,   // p := p / s (1 param)
    IK_HALFSPACE = 10, // nx ny nz off
    IK_SPHERE = 11,    // cx cy cz r
    IK_BOX = 12,       // cx cy cz hx hy hz
    IK_CYL = 13,       // a = axis; cx cy cz r h/2
    IK_CONE = 14,      // a = axis; cx cy cz h/2 h r1 r2
    IK_FRUSTUM = 15,   // z0 z1 nz_xhi c_xhi nz_xlo c_xlo nz_yhi c_yhi nz_ylo c_ylo
    IK_GRID = 16,      // a = grid leaf index
    IK_OFFSET = 20,    // v - d
    IK_SCALE_OUT = 21, // v * s
    IK_UNION = 22,     // a = 1 when the second operand was evaluated first; k
    IK_INTERSECT = 23,
    IK_SUBTRACT = 24,
};

struct IkIns {
    int32_t op, a, pofs, pad;
};

// One trilinear grid leaf (FE/ikarus_v1/eval_warp.py:92-148): samples at
// origin + i*pitch, C order; outside [0, n-1] the leaf reads 1e6.
struct IkGrid {
    int64_t off;
    int32_t nx, ny, nz, is_sdf;
    float ox, oy, oz, pitch, iso, pad;
};

__device__ __forceinline__ float wmin(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float wmax(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float wclamp(float x, float lo, float hi) { return wmin(wmax(lo, x), hi); }
__device__ __forceinline__ float wlerp(float a, float b, float t) { return a * (1.0f - t) + b * t; }

__device__ float ik_grid(const IkGrid& g, const float* __restrict__ data, float px, float py, float pz)
{
    const float fx = (px - g.ox) / g.pitch, fy = (py - g.oy) / g.pitch, fz = (pz - g.oz) / g.pitch;
    if (fx < 0.0f || fy < 0.0f || fz < 0.0f || fx > (float)(g.nx - 1) || fy > (float)(g.ny - 1) ||
        fz > (float)(g.nz - 1))
        return 1.0e6f;
    const int ix = (int)floorf(fx), iy = (int)floorf(fy), iz = (int)floorf(fz);
    const int ix1 = min(ix + 1, g.nx - 1), iy1 = min(iy + 1, g.ny - 1), iz1 = min(iz + 1, g.nz - 1);
    const float tx = fx - (float)ix, ty = fy - (float)iy, tz = fz - (float)iz;
    const float* d = data + g.off;
    auto at = [&](int i, int j, int k) { return d[((int64_t)i * g.ny + j) * g.nz + k]; };
    const float c00 = wlerp(at(ix, iy, iz), at(ix1, iy, iz), tx);
    const float c10 = wlerp(at(ix, iy1, iz), at(ix1, iy1, iz), tx);
    const float c01 = wlerp(at(ix, iy, iz1), at(ix1, iy, iz1), tx);
    const float c11 = wlerp(at(ix, iy1, iz1), at(ix1, iy1, iz1), tx);
    const float c0 = wlerp(c00, c10, ty), c1 = wlerp(c01, c11, ty);
    const float val = wlerp(c0, c1, tz);
    return g.is_sdf == 1 ? val : (g.iso - val) * g.pitch;
}

// Radial and axial coordinates of a capped cylinder or cone about axis a.
__device__ __forceinline__ void axial(int a, float lx, float ly, float lz, float& rad, float& ax)
{
    if (a == 2) {
        rad = sqrtf(lx * lx + ly * ly);
        ax = lz;
    } else if (a == 1) {
        rad = sqrtf(lx * lx + lz * lz);
        ax = ly;
    } else {
        rad = sqrtf(ly * ly + lz * lz);
        ax = lx;
    }
}

__device__ __forceinline__ float round_box2(float qx, float qy)
{
    const float ox = wmax(qx, 0.0f), oy = wmax(qy, 0.0f);
    return sqrtf(ox * ox + oy * oy) + wmin(wmax(qx, qy), 0.0f);
}

// Set (never cleared on the device) when a program underflows or overflows
// the value stack or holds an unknown opcode; fe_ikarus_fault reads and
// clears it.  One word per GPU, shared by every call on that GPU.
__device__ unsigned int ik_fault;

__device__ __noinline__ float ik_bad()
{
    atomicOr(&ik_fault, 1u);
    return __builtin_nanf("");
}

// Evaluates the program at world point (wx, wy, wz); stk is this thread's
// column of the block's LDS stack (stride IK_T).  A malformed program (a
// binary operation without two values, a push past IK_STACK slots, an
// unknown opcode) stops at that instruction: the value is NaN and the fault
// word is set, and no LDS outside this thread's column is touched.
__device__ float ik_eval(const IkIns* __restrict__ prog, int n, const float* __restrict__ par,
                         const IkGrid* __restrict__ grids, const float* __restrict__ gdata, float wx, float wy,
                         float wz, float* stk)
{
    float px = wx, py = wy, pz = wz;
    float top = 0.0f;
    int sp = 0;  // values below the top held in LDS
    bool have = false;
    for (int i = 0; i < n; ++i) {
        const IkIns in = prog[i];
        const float* q = par + in.pofs;
        float v;
        switch (in.op) {
            case IK_WORLD:
                px = wx;
                py = wy;
                pz = wz;
                continue;
            case IK_XFORM: {
                const float lx = q[0] * px + q[1] * py + q[2] * pz + q[3];
                const float ly = q[4] * px + q[5] * py + q[6] * pz + q[7];
                const float lz = q[8] * px + q[9] * py + q[10] * pz + q[11];
                px = lx;
                py = ly;
                pz = lz;
                continue;
            }
            case IK_PROJ:
                if (in.a == 0) px = q[0];
                else if (in.a == 1) py = q[0];
                else pz = q[0];
                continue;
            case IK_SCALE_IN:
                px = px / q[0];
                py = py / q[0];
                pz = pz / q[0];
                continue;
            case IK_HALFSPACE:
                v = px * q[0] + py * q[1] + pz * q[2] - q[3];
                break;
            case IK_SPHERE: {
                const float dx = px - q[0], dy = py - q[1], dz = pz - q[2];
                v = sqrtf(dx * dx + dy * dy + dz * dz) - q[3];
                break;
            }
            case IK_BOX: {
                const float qx = fabsf(px - q[0]) - q[3], qy = fabsf(py - q[1]) - q[4], qz = fabsf(pz - q[2]) - q[5];
                const float ox = wmax(qx, 0.0f), oy = wmax(qy, 0.0f), oz = wmax(qz, 0.0f);
                v = sqrtf(ox * ox + oy * oy + oz * oz) + wmin(wmax(qx, wmax(qy, qz)), 0.0f);
                break;
            }
            case IK_CYL: {
                float rad, ax;
                axial(in.a, px - q[0], py - q[1], pz - q[2], rad, ax);
                v = round_box2(rad - q[3], fabsf(ax) - q[4]);
                break;
            }
            case IK_CONE: {
                float rad, ax;
                axial(in.a, px - q[0], py - q[1], pz - q[2], rad, ax);
                float t = (ax + q[3]) / q[4];
                t = wclamp(t, 0.0f, 1.0f);
                const float r_at = q[5] + (q[6] - q[5]) * t;
                v = round_box2(rad - r_at, fabsf(ax) - q[3]);
                break;
            }
            case IK_FRUSTUM: {
                const float d_zlo = q[0] - pz, d_zhi = pz - q[1];
                const float d_xhi = (px + q[2] * pz - q[3]) / sqrtf(1.0f * 1.0f + q[2] * q[2]);
                const float d_xlo = (-px + q[4] * pz - q[5]) / sqrtf(-1.0f * -1.0f + q[4] * q[4]);
                const float d_yhi = (py + q[6] * pz - q[7]) / sqrtf(1.0f * 1.0f + q[6] * q[6]);
                const float d_ylo = (-py + q[8] * pz - q[9]) / sqrtf(-1.0f * -1.0f + q[8] * q[8]);
                const float m1 = wmax(d_zlo, d_zhi), m2 = wmax(d_xhi, d_xlo), m3 = wmax(d_yhi, d_ylo);
                v = wmax(m1, wmax(m2, m3));
                break;
            }
            case IK_GRID:
                v = ik_grid(grids[in.a], gdata, px, py, pz);
                break;
            case IK_OFFSET:
                if (!have) return ik_bad();
                top = top - q[0];
                continue;
            case IK_SCALE_OUT:
                if (!have) return ik_bad();
                top = top * q[0];
                continue;
            case IK_UNION:
            case IK_INTERSECT:
            case IK_SUBTRACT: {  // the second-evaluated operand is the top, the first is in LDS
                if (sp <= 0 || !have) return ik_bad();
                const float below = stk[(--sp) * IK_T];
                const float da = in.a ? top : below, db = in.a ? below : top;
                const float k = q[0];
                if (in.op == IK_UNION) {
                    if (k <= 0.0f) {
                        top = wmin(da, db);
                    } else {
                        const float h = wmax(0.0f, k - fabsf(da - db)) / k;
                        top = wmin(da, db) - h * h * k * 0.25f;
                    }
                } else if (in.op == IK_INTERSECT) {
                    if (k <= 0.0f) {
                        top = wmax(da, db);
                    } else {
                        const float h = wmax(0.0f, k - fabsf(da - db)) / k;
                        top = wmax(da, db) + h * h * k * 0.25f;
                    }
                } else {
                    if (k <= 0.0f) {
                        top = wmax(da, -db);
                    } else {
                        const float h = wmax(0.0f, k - fabsf(da + db)) / k;
                        top = wmax(da, -db) + h * h * k * 0.25f;
                    }
                }
                continue;
            }
            default:
                return ik_bad();
        }
        // a leaf value: push the old top, keep the new one in the register
        if (have) {
            if (sp >= IK_STACK) return ik_bad();
            stk[(sp++) * IK_T] = top;
        }
        top = v;
        have = true;
    }
    return top;
}

struct IkProgram {
    const IkIns* ins;
    const float* par;
    const IkGrid* grids;
    const float* gdata;
    int n;
};

__global__ __launch_bounds__(IK_T) void k_ik_points(IkProgram P, const float* __restrict__ pts, int64_t n,
                                                     float* __restrict__ out)
{
    __shared__ float stk[IK_STACK * IK_T];
    const int64_t i = (int64_t)blockIdx.x * IK_T + threadIdx.x;
    if (i >= n) return;
    out[i] = ik_eval(P.ins, P.n, P.par, P.grids, P.gdata, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2],
                     stk + threadIdx.x);
}

struct IkGridMeta {
    float x0, y0, z0, pitch, margin;
    int nbx, nby, nbz, B;
};

// Eight-corner classification of every block (FE/faltkarna_v1.py:128-164):
// corners at origin + {0, B*h} per axis; mixed signs (v >= 0 is positive) ->
// surface, else min|v| <= margin -> surface, else all negative -> interior,
// else exterior.  Eight lanes per block.
__global__ __launch_bounds__(IK_T) void k_ik_classify(IkProgram P, IkGridMeta G, int32_t* __restrict__ kind,
                                                       float* __restrict__ cma)
{
    __shared__ float stk[IK_STACK * IK_T];
    const int64_t nb = (int64_t)G.nbx * G.nby * G.nbz;
    const int64_t b = ((int64_t)blockIdx.x * IK_T + threadIdx.x) >> 3;
    const int c = threadIdx.x & 7;
    const bool valid = b < nb;
    const int64_t bb = valid ? b : 0;
    const int bi = (int)(bb / ((int64_t)G.nby * G.nbz)), bj = (int)((bb / G.nbz) % G.nby), bk = (int)(bb % G.nbz);
    const float side = (float)G.B * G.pitch;
    const float ox = G.x0 + (float)(bi * G.B) * G.pitch;
    const float oy = G.y0 + (float)(bj * G.B) * G.pitch;
    const float oz = G.z0 + (float)(bk * G.B) * G.pitch;
    const float x = ox + (float)(c & 1) * side, y = oy + (float)((c >> 1) & 1) * side,
                z = oz + (float)((c >> 2) & 1) * side;
    const float v = ik_eval(P.ins, P.n, P.par, P.grids, P.gdata, x, y, z, stk + threadIdx.x);
    float mn = fabsf(v);
    int pos = v >= 0.0f, neg = !pos;
#pragma unroll
    for (int s = 1; s < 8; s <<= 1) {
        mn = wmin(mn, __shfl_xor(mn, s));
        pos |= __shfl_xor(pos, s);
        neg |= __shfl_xor(neg, s);
    }
    mn = wmin(1.0e30f, mn);
    if (valid && c == 0) {
        cma[b] = mn;
        kind[b] = (pos && neg) ? 0 : (mn <= G.margin ? 0 : (neg ? -1 : 1));
    }
}

// Tiles of the active blocks `ids` ([n_active][B][B][B], C order), sample
// (i, j, k) of block (bi, bj, bk) at origin + (float)(bi*B + i)*h per axis
// (FE/faltkarna_v1.py:166-190).
__global__ __launch_bounds__(IK_T) void k_ik_tiles(IkProgram P, IkGridMeta G, const int32_t* __restrict__ ids,
                                                    int64_t n_active, float* __restrict__ tiles)
{
    __shared__ float stk[IK_STACK * IK_T];
    const int64_t B3 = (int64_t)G.B * G.B * G.B;
    const int64_t t = (int64_t)blockIdx.x * IK_T + threadIdx.x;
    if (t >= n_active * B3) return;
    const int64_t a = t / B3, r = t - a * B3;
    const int li = (int)(r / (G.B * G.B)), lj = (int)((r / G.B) % G.B), lk = (int)(r % G.B);
    const int64_t b = ids[a];
    const int bi = (int)(b / ((int64_t)G.nby * G.nbz)), bj = (int)((b / G.nbz) % G.nby), bk = (int)(b % G.nbz);
    const float x = G.x0 + (float)(bi * G.B + li) * G.pitch;
    const float y = G.y0 + (float)(bj * G.B + lj) * G.pitch;
    const float z = G.z0 + (float)(bk * G.B + lk) * G.pitch;
    tiles[t] = ik_eval(P.ins, P.n, P.par, P.grids, P.gdata, x, y, z, stk + threadIdx.x);
}

// Block-level sums of two counts into 64-bit global counters (one atomic
// each per block).
__device__ __forceinline__ void block_count(int a, int b, unsigned long long* out)
{
    __shared__ int sa[IK_T / FE_WAVE], sb[IK_T / FE_WAVE];
    const int lane = threadIdx.x & (FE_WAVE - 1), wave = threadIdx.x / FE_WAVE;
    const int wa = __popcll(__ballot(a != 0)), wb = __popcll(__ballot(b != 0));
    if (lane == 0) {
        sa[wave] = wa;
        sb[wave] = wb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int ta = 0, tb = 0;
        for (int w = 0; w < IK_T / FE_WAVE; ++w) {
            ta += sa[w];
            tb += sb[w];
        }
        if (ta) atomicAdd(out, (unsigned long long)ta);
        if (tb) atomicAdd(out + 1, (unsigned long long)tb);
    }
}

// Regular-grid pass of a hierarchical query (FE/ikarus_v1/queries.py:120-131,
// 211-223): sample i = (lo + pitch*i) per axis in f64, rounded to fp32 (what
// NumPy's grid and eval_batch's fp32 cast produce), C order; counts[0] +=
// samples with v < below, counts[1] += samples with |v| <= band; the indices
// of the band samples are appended to `cells` (order unspecified) through
// cursor (cells may be null to count only).
__global__ __launch_bounds__(IK_T) void k_ik_grid(IkProgram P, double lx, double ly, double lz, double pitch, int nx,
                                                   int ny, int nz, float below, float band,
                                                   unsigned long long* counts, int32_t* cells,
                                                   unsigned int* cursor)
{
    __shared__ float stk[IK_STACK * IK_T];
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t t = (int64_t)blockIdx.x * IK_T + threadIdx.x;
    bool in = false, bd = false;
    if (t < n) {
        const int i = (int)(t / ((int64_t)ny * nz)), j = (int)((t / nz) % ny), k = (int)(t % nz);
        const float x = (float)(lx + pitch * (double)i), y = (float)(ly + pitch * (double)j),
                    z = (float)(lz + pitch * (double)k);
        const float v = ik_eval(P.ins, P.n, P.par, P.grids, P.gdata, x, y, z, stk + threadIdx.x);
        in = v < below;
        bd = fabsf(v) <= band;
    }
    if (cells) {
        const uint64_t m = __ballot(bd);
        const int lane = threadIdx.x & (FE_WAVE - 1);
        unsigned int base = 0;
        if (lane == 0 && m) base = atomicAdd(cursor, (unsigned int)__popcll(m));
        base = __shfl(base, 0);
        if (bd) cells[base + __popcll(m & ((1ull << lane) - 1))] = (int32_t)t;
    }
    block_count(in, bd, counts);
}

// Refinement of band cells (FE/ikarus_v1/queries.py:234-249): cell c (grid
// index) gets r^3 samples at centre(c) + ((a + 0.5)/r * coarse - half) per axis
// (f64, a-major order), rounded to fp32; counts[0] += samples with v < 0,
// counts[1] += samples with |v| <= band.
__global__ __launch_bounds__(IK_T) void k_ik_refine(IkProgram P, double lx, double ly, double lz, double pitch, int ny,
                                                     int nz, const int32_t* __restrict__ cells, int64_t ncells, int r,
                                                     double half, float band, unsigned long long* counts)
{
    __shared__ float stk[IK_STACK * IK_T];
    const int64_t r3 = (int64_t)r * r * r;
    const int64_t t = (int64_t)blockIdx.x * IK_T + threadIdx.x;
    bool in = false, bd = false;
    if (t < ncells * r3) {
        const int64_t c = cells[t / r3];
        const int sub = (int)(t % r3);
        const int a = sub / (r * r), b = (sub / r) % r, e = sub % r;
        const int i = (int)(c / ((int64_t)ny * nz)), j = (int)((c / nz) % ny), k = (int)(c % nz);
        const double ox = ((double)a + 0.5) / (double)r * pitch - half;
        const double oy = ((double)b + 0.5) / (double)r * pitch - half;
        const double oz = ((double)e + 0.5) / (double)r * pitch - half;
        const float x = (float)((lx + pitch * (double)i) + ox), y = (float)((ly + pitch * (double)j) + oy),
                    z = (float)((lz + pitch * (double)k) + oz);
        const float v = ik_eval(P.ins, P.n, P.par, P.grids, P.gdata, x, y, z, stk + threadIdx.x);
        in = v < 0.0f;
        bd = fabsf(v) <= band;
    }
    block_count(in, bd, counts);
}

// ---------------------------------------------------------------- grid-leaf pairs
// Hard intersection of two grid leaves, the pair chosen per point
// (FE/ikarus_v1/eval_warp.py:504-560): max(leaf a, leaf b) with the selection
// max.  Out-of-range pair or leaf indices give NaN.
__global__ __launch_bounds__(IK_T) void k_ik_pairs(const IkGrid* __restrict__ grids, int nleaves,
                                                    const float* __restrict__ gdata, const float* __restrict__ pts,
                                                    const int32_t* __restrict__ pair, const int32_t* __restrict__ la,
                                                    const int32_t* __restrict__ lb, int npairs, int64_t n,
                                                    float* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * IK_T + threadIdx.x;
    if (i >= n) return;
    const int p = pair[i];
    const int a = (p >= 0 && p < npairs) ? la[p] : -1, b = (p >= 0 && p < npairs) ? lb[p] : -1;
    if (a < 0 || a >= nleaves || b < 0 || b >= nleaves) {
        out[i] = __builtin_nanf("");
        return;
    }
    const float x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
    out[i] = wmax(ik_grid(grids[a], gdata, x, y, z), ik_grid(grids[b], gdata, x, y, z));
}

// Band tree with the plan as data (FE/ikarus_v1/eval_warp.py:567-700): the
// minimum over bands i of max(leaf(p with p[axis] := zs[i]), max(p[axis] -
// z1[i], z0[i] - p[axis])), folded from 1e30 with the selection min/max --
// the interpreter's arithmetic on the same tree (a unit-axis halfspace adds
// only signed zeros to p[axis]).
__global__ __launch_bounds__(IK_T) void k_ik_band(const IkGrid* __restrict__ grids, const float* __restrict__ gdata,
                                                   int axis, const float* __restrict__ zs, const float* __restrict__ z0,
                                                   const float* __restrict__ z1, int nbands,
                                                   const float* __restrict__ pts, int64_t n, float* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * IK_T + threadIdx.x;
    if (i >= n) return;
    const float x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
    const float pax = axis == 0 ? x : (axis == 1 ? y : z);
    const IkGrid g = grids[0];
    float best = 1.0e30f;
    for (int b = 0; b < nbands; ++b) {
        const float s = zs[b];
        const float body = ik_grid(g, gdata, axis == 0 ? s : x, axis == 1 ? s : y, axis == 2 ? s : z);
        const float slab = wmax(pax - z1[b], z0[b] - pax);
        best = wmin(best, wmax(body, slab));
    }
    out[i] = best;
}

bool valid_program(const void* ins, int n_ins, const void* par)
{
    return ins && par && n_ins > 0 && n_ins <= (1 << 20);
}

// fp32 parameters an instruction reads (-1: unknown opcode).
int ik_param_count(int op)
{
    switch (op) {
        case IK_WORLD: case IK_GRID: return 0;
        case IK_XFORM: return 12;
        case IK_PROJ: case IK_SCALE_IN: case IK_OFFSET: case IK_SCALE_OUT: return 1;
        case IK_UNION: case IK_INTERSECT: case IK_SUBTRACT: return 1;
        case IK_HALFSPACE: case IK_SPHERE: return 4;
        case IK_BOX: return 6;
        case IK_CYL: return 5;
        case IK_CONE: return 7;
        case IK_FRUSTUM: return 10;
        default: return -1;
    }
}

std::mutex fault_lock;
unsigned int* fault_ptr[64];  // per device, from hipGetSymbolAddress with that device current

hipError_t fault_word(unsigned int** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> hold(fault_lock);
    if (!fault_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(ik_fault));
        if (e != hipSuccess) return e;
        fault_ptr[dev] = static_cast<unsigned int*>(p);
    }
    *out = fault_ptr[dev];
    return hipSuccess;
}

}  // namespace

// Checks a compiled program on the host before upload: 1 <= n_ins <= 2^20,
// known opcodes, every instruction's parameters inside params[0, n_params),
// grid leaves below n_grids, axis arguments 0..2, the binary order flag 0/1,
// and the value stack: a unary or binary operation finds its operands, no
// push goes past 16 LDS slots, and the program ends with one value.
// FE_INVALID names the first bad instruction in fe_last_error().
extern "C" int fe_ikarus_check(const void* ins_host, int n_ins, int64_t n_params, int n_grids) noexcept
{
    char msg[160];
    if (!ins_host || n_ins < 1 || n_ins > (1 << 20) || n_params < 0 || n_grids < 0) {
        fe_note_text("fe_ikarus_check: a program of 1..2^20 instructions and non-negative sizes required");
        return FE_INVALID;
    }
    const IkIns* in = static_cast<const IkIns*>(ins_host);
    int sp = 0;
    bool have = false;
    for (int i = 0; i < n_ins; ++i) {
        const int op = in[i].op, a = in[i].a;
        const int np = ik_param_count(op);
        const char* bad = nullptr;
        if (np < 0) bad = "unknown opcode";
        else if (np > 0 && (in[i].pofs < 0 || (int64_t)in[i].pofs + np > n_params)) bad = "parameters out of range";
        else if (op == IK_GRID && (a < 0 || a >= n_grids)) bad = "grid leaf out of range";
        else if ((op == IK_PROJ || op == IK_CYL || op == IK_CONE) && (a < 0 || a > 2)) bad = "axis out of range";
        else if ((op == IK_UNION || op == IK_INTERSECT || op == IK_SUBTRACT) && (a < 0 || a > 1)) bad = "bad order flag";
        else if ((op == IK_OFFSET || op == IK_SCALE_OUT) && !have) bad = "unary operation without a value";
        else if (op == IK_UNION || op == IK_INTERSECT || op == IK_SUBTRACT) {
            if (!have || sp < 1) bad = "binary operation without two values";
            else --sp;
        } else if (op >= IK_HALFSPACE && op <= IK_GRID) {
            if (have && sp >= IK_STACK) bad = "value stack deeper than 17";
            else if (have) ++sp;
            have = true;
        }
        if (bad) {
            snprintf(msg, sizeof msg, "fe_ikarus_check: instruction %d (opcode %d): %s", i, op, bad);
            fe_note_text(msg);
            return FE_INVALID;
        }
    }
    if (!have || sp != 0) {
        fe_note_text("fe_ikarus_check: the program does not end with exactly one value");
        return FE_INVALID;
    }
    return FE_OK;
}

// Waits for `stream`, then reads and clears this GPU's fault word:
// FE_INVALID if a program evaluated since the last call was malformed (its
// values are NaN), else FE_OK.
extern "C" int fe_ikarus_fault(void* stream) noexcept
{
    unsigned int* word = nullptr;
    FE_TRY(fault_word(&word));
    FE_TRY(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    unsigned int v = 0;
    FE_TRY(hipMemcpy(&v, word, sizeof v, hipMemcpyDeviceToHost));
    if (!v) return FE_OK;
    const unsigned int zero = 0;
    FE_TRY(hipMemcpy(word, &zero, sizeof zero, hipMemcpyHostToDevice));
    fe_note_text("fe_ikarus_fault: a program under- or overflowed the value stack or held an unknown opcode");
    return FE_INVALID;
}

// Values of a compiled tree at device points (n x 3 fp32) into out (n fp32).
// ins: n_ins 16-byte instructions {op, a, param offset, 0}; params: fp32;
// grids: 48-byte grid-leaf records {int64 offset; int32 nx, ny, nz, is_sdf;
// float ox, oy, oz, pitch, iso, pad} over the fp32 samples `grid_data` (both
// may be null without grid leaves).  The program's value stack must fit 17
// slots (the compiler checks).
extern "C" int fe_ikarus_eval(const void* ins, int n_ins, const float* params, const void* grids,
                              const float* grid_data, const float* points, int64_t n, float* out, void* stream) noexcept
{
    if (!valid_program(ins, n_ins, params) || !points || !out || n < 0) return FE_INVALID;
    if (n == 0) return FE_OK;
    const IkProgram P{static_cast<const IkIns*>(ins), params, static_cast<const IkGrid*>(grids), grid_data, n_ins};
    hipLaunchKernelGGL(k_ik_points, dim3((unsigned)((n + IK_T - 1) / IK_T)), dim3(IK_T), 0,
                       static_cast<hipStream_t>(stream), P, points, n, out);
    FE_TRY(hipGetLastError());
    return FE_OK;
}

// Block classification of a compiled tree over the grid (x0, y0, z0, pitch,
// margin as fp32; nbx x nby x nbz blocks of edge B): kind int32 and
// corner_min_abs fp32 per block, C order.
extern "C" int fe_ikarus_classify(const void* ins, int n_ins, const float* params, const void* grids,
                                  const float* grid_data, const float* origin_pitch_margin, const int32_t* blocks,
                                  int B, int32_t* kind, float* cma, void* stream) noexcept
{
    if (!valid_program(ins, n_ins, params) || !origin_pitch_margin || !blocks || !kind || !cma || B < 1 || B > 64)
        return FE_INVALID;
    const IkGridMeta G{origin_pitch_margin[0], origin_pitch_margin[1], origin_pitch_margin[2], origin_pitch_margin[3],
                       origin_pitch_margin[4], blocks[0], blocks[1], blocks[2], B};
    if (G.nbx < 1 || G.nby < 1 || G.nbz < 1) return FE_INVALID;
    const int64_t nb = (int64_t)G.nbx * G.nby * G.nbz;
    const IkProgram P{static_cast<const IkIns*>(ins), params, static_cast<const IkGrid*>(grids), grid_data, n_ins};
    hipLaunchKernelGGL(k_ik_classify, dim3((unsigned)((nb * 8 + IK_T - 1) / IK_T)), dim3(IK_T), 0,
                       static_cast<hipStream_t>(stream), P, G, kind, cma);
    FE_TRY(hipGetLastError());
    return FE_OK;
}

// Tiles [n_active][B][B][B] of the ascending active block ids.
extern "C" int fe_ikarus_tiles(const void* ins, int n_ins, const float* params, const void* grids,
                               const float* grid_data, const float* origin_pitch_margin, const int32_t* blocks, int B,
                               const int32_t* ids, int64_t n_active, float* tiles, void* stream) noexcept
{
    if (!valid_program(ins, n_ins, params) || !origin_pitch_margin || !blocks || !ids || !tiles || B < 1 || B > 64 ||
        n_active < 0)
        return FE_INVALID;
    if (n_active == 0) return FE_OK;
    const IkGridMeta G{origin_pitch_margin[0], origin_pitch_margin[1], origin_pitch_margin[2], origin_pitch_margin[3],
                       origin_pitch_margin[4], blocks[0], blocks[1], blocks[2], B};
    const IkProgram P{static_cast<const IkIns*>(ins), params, static_cast<const IkGrid*>(grids), grid_data, n_ins};
    const int64_t total = n_active * B * B * B;
    hipLaunchKernelGGL(k_ik_tiles, dim3((unsigned)((total + IK_T - 1) / IK_T)), dim3(IK_T), 0,
                       static_cast<hipStream_t>(stream), P, G, ids, n_active, tiles);
    FE_TRY(hipGetLastError());
    return FE_OK;
}

// One regular-grid pass of a hierarchical query (see k_ik_grid): lo and pitch
// in f64 (host values), nx*ny*nz < 2^31 samples; counts: device uint64[2]
// (added to); cells/cursor: device int32[nx*ny*nz] and uint32 (zeroed by the
// caller), or both null.
extern "C" int fe_ikarus_grid(const void* ins, int n_ins, const float* params, const void* grids, const float* grid_data,
                              const double* lo, double pitch, const int32_t* shape, float below, float band,
                              unsigned long long* counts, int32_t* cells, unsigned int* cursor, void* stream) noexcept
{
    if (!valid_program(ins, n_ins, params) || !lo || !shape || !counts || !(pitch > 0) || (!cells) != (!cursor))
        return FE_INVALID;
    const int64_t n = (int64_t)shape[0] * shape[1] * shape[2];
    if (shape[0] < 1 || shape[1] < 1 || shape[2] < 1 || n >= INT32_MAX) return FE_INVALID;
    const IkProgram P{static_cast<const IkIns*>(ins), params, static_cast<const IkGrid*>(grids), grid_data, n_ins};
    hipLaunchKernelGGL(k_ik_grid, dim3((unsigned)((n + IK_T - 1) / IK_T)), dim3(IK_T), 0,
                       static_cast<hipStream_t>(stream), P, lo[0], lo[1], lo[2], pitch, shape[0], shape[1], shape[2],
                       below, band, counts, cells, cursor);
    FE_TRY(hipGetLastError());
    return FE_OK;
}

// Refinement pass (see k_ik_refine) over ncells band cells of the grid (lo,
// pitch, shape) with r^3 samples each.
extern "C" int fe_ikarus_refine(const void* ins, int n_ins, const float* params, const void* grids,
                                const float* grid_data, const double* lo, double pitch, const int32_t* shape,
                                const int32_t* cells, int64_t ncells, int r, float band, unsigned long long* counts,
                                void* stream) noexcept
{
    if (!valid_program(ins, n_ins, params) || !lo || !shape || !counts || !(pitch > 0) || r < 1 || r > 64 ||
        ncells < 0 || (ncells && !cells))
        return FE_INVALID;
    const int64_t total = ncells * r * r * r;
    if (total == 0) return FE_OK;
    const IkProgram P{static_cast<const IkIns*>(ins), params, static_cast<const IkGrid*>(grids), grid_data, n_ins};
    hipLaunchKernelGGL(k_ik_refine, dim3((unsigned)((total + IK_T - 1) / IK_T)), dim3(IK_T), 0,
                       static_cast<hipStream_t>(stream), P, lo[0], lo[1], lo[2], pitch, shape[1], shape[2], cells,
                       ncells, r, pitch / 2.0, band, counts);
    FE_TRY(hipGetLastError());
    return FE_OK;
}

// Grid-leaf pairs at device points (FE/ikarus_v1/eval_warp.py:545-560):
// out[i] = max(leaf la[pair[i]], leaf lb[pair[i]]) at point i (n x 3 fp32);
// grids: nleaves 48-byte records over `grid_data`; pair ids index la / lb
// (npairs each), out-of-range indices give NaN.
extern "C" int fe_ikarus_pairs_eval(const void* grids, int nleaves, const float* grid_data, const float* points,
                                    const int32_t* pair_of_point, const int32_t* leaf_a, const int32_t* leaf_b,
                                    int npairs, int64_t n, float* out, void* stream) noexcept
{
    if (!grids || !grid_data || nleaves < 1 || !points || !pair_of_point || !leaf_a || !leaf_b || npairs < 1 ||
        n < 0 || !out)
        return FE_INVALID;
    if (n == 0) return FE_OK;
    hipLaunchKernelGGL(k_ik_pairs, dim3((unsigned)((n + IK_T - 1) / IK_T)), dim3(IK_T), 0,
                       static_cast<hipStream_t>(stream), static_cast<const IkGrid*>(grids), nleaves, grid_data, points,
                       pair_of_point, leaf_a, leaf_b, npairs, n, out);
    FE_TRY(hipGetLastError());
    return FE_OK;
}

// Band tree of one grid leaf (grids[0]) at device points: the plan is nbands
// fp32 triples (zs, z0, z1) in three device arrays; axis 0..2.
extern "C" int fe_ikarus_band_eval(const void* grids, const float* grid_data, int axis, const float* zs,
                                   const float* z0, const float* z1, int nbands, const float* points, int64_t n,
                                   float* out, void* stream) noexcept
{
    if (!grids || !grid_data || axis < 0 || axis > 2 || nbands < 0 || (nbands && (!zs || !z0 || !z1)) || !points ||
        n < 0 || !out)
        return FE_INVALID;
    if (n == 0) return FE_OK;
    hipLaunchKernelGGL(k_ik_band, dim3((unsigned)((n + IK_T - 1) / IK_T)), dim3(IK_T), 0,
                       static_cast<hipStream_t>(stream), static_cast<const IkGrid*>(grids), grid_data, axis, zs, z0, z1,
                       nbands, points, n, out);
    FE_TRY(hipGetLastError());
    return FE_OK;
}

