// Recipe records, field descriptors, the fp32 recipe evaluation and the
// exact wave-level culling of primitive lists, shared by the block-sparse
// build (field.hip) and the per-block-pitch field (multires.hip).
#pragma once
#include "fe_common.hpp"

#define FE_BIG 1.0e6f
#define FE_MAX_PRIMS 4096
#define FE_CHUNK 1024

struct Prim {
    int32_t kind, op;
    float cx, cy, cz, a, b, c;
};
static_assert(sizeof(Prim) == 32, "prim layout");

// One field to build (160 bytes).  Pointers are device addresses.
struct FieldDesc {
    float x0, y0, z0, pitch;
    float margin, slack;
    int32_t nprims, B;
    int32_t nx, ny, nz, SB;
    int32_t nbx, nby, nbz, cap;
    int32_t nsbx, nsby, nsbz, corrupt;
    const Prim* prims;
    int32_t* kind;
    float* cma;
    uint16_t* lists;
    int32_t* list_len;
    int32_t* ids;
    int32_t* slot;
    float* tiles;
    int32_t* counts;  // [0] n_active, [1] overflow (active blocks beyond tile_cap)
    int64_t tile_cap;
};
static_assert(sizeof(FieldDesc) == 160, "descriptor layout");

// ---------------------------------------------------------------- primitives
// Same operation order as oracle/fe_oracle.c or_prim_eval (IKARUS emitted
// fp32 source, FE/ikarus_v1/eval_warp.py:182-224; FE/faltkarna_v1.py:91-104).
// Square roots are correctly rounded (sqrt_rn, the bits of sqrtf).  FAST:
// the bare v_sqrt_f32 (within one ulp) instead; used only for the centres of
// culling intervals, whose slack is far wider than that error.
template <bool FAST = false>
__device__ __forceinline__ float root(float x) { return FAST ? __builtin_amdgcn_sqrtf(x) : sqrt_rn(x); }
template <bool FAST = false>
__device__ __forceinline__ float prim_eval(int kind, float cx, float cy, float cz, float a, float b, float c, float x,
                                           float y, float z)
{
    float lx = x - cx, ly = y - cy, lz = z - cz;
    switch (kind) {
    case 0: {
        float s = lx * lx + ly * ly;
        s = s + lz * lz;
        return root<FAST>(s) - a;
    }
    case 1: {
        float qx = fabsf(lx) - a, qy = fabsf(ly) - b, qz = fabsf(lz) - c;
        float ox = sel_max(qx, 0.0f), oy = sel_max(qy, 0.0f), oz = sel_max(qz, 0.0f);
        float s = ox * ox + oy * oy;
        s = s + oz * oz;
        return root<FAST>(s) + sel_min(sel_max(qx, sel_max(qy, qz)), 0.0f);
    }
    case 2:
    case 3:
    case 4: {
        float r2, ax;
        if (kind == 4) { r2 = lx * lx + ly * ly; ax = lz; }
        else if (kind == 3) { r2 = lx * lx + lz * lz; ax = ly; }
        else { r2 = ly * ly + lz * lz; ax = lx; }
        float qx = root<FAST>(r2) - a, qy = fabsf(ax) - b;
        float ox = sel_max(qx, 0.0f), oy = sel_max(qy, 0.0f);
        return root<FAST>(ox * ox + oy * oy) + sel_min(sel_max(qx, qy), 0.0f);
    }
    case 5: {
        float qx = root<FAST>(lx * lx + ly * ly) - a;
        return root<FAST>(qx * qx + lz * lz) - b;
    }
    default:  // 6: infinite z-cylinder
        return root<FAST>(lx * lx + ly * ly) - a;
    }
}

__device__ __forceinline__ float fold(int op, float f, float v)
{
    return op == 0 ? sel_min(f, v) : (op == 1 ? sel_max(f, -v) : sel_max(f, v));
}

// Evaluates K points over a primitive list (list == nullptr: all primitives).
// Primitive parameters are wave-uniform scalar loads shared by the K points:
// the list (wave-uniform, n entries) is read 64 entries at a time, one per
// lane, and broadcast with readlane; the records are read through the
// constant address space, so uniform indices become scalar loads and the
// kind switch scalar branches, and the next record's load is issued before
// the current record is evaluated.
typedef const __attribute__((address_space(4))) Prim ConstPrim;
__device__ __forceinline__ Prim load_prim(ConstPrim* cp, int idx)
{
    ConstPrim* q = cp + idx;
    Prim p;
    p.kind = q->kind;
    p.op = q->op;
    p.cx = q->cx;
    p.cy = q->cy;
    p.cz = q->cz;
    p.a = q->a;
    p.b = q->b;
    p.c = q->c;
    return p;
}
template <int K>
__device__ __forceinline__ void eval_points(const Prim* __restrict__ prims, const uint16_t* list, int n, const float* x,
                                            const float* y, const float* z, float* f)
{
    ConstPrim* cp = (ConstPrim*)prims;
    const int lane = threadIdx.x & 63;
    n = __builtin_amdgcn_readfirstlane(n);
    for (int jb = 0; jb < n; jb += 64) {
        const int cnt = min(64, n - jb);
        const int mine = list ? (lane < cnt ? (int)list[jb + lane] : 0) : jb + lane;
        Prim pn = load_prim(cp, __builtin_amdgcn_readlane(mine, 0));
        int t = 0;
        if (jb == 0) {  // the first primitive starts the fold
            const Prim p = pn;
            pn = load_prim(cp, __builtin_amdgcn_readlane(mine, min(1, cnt - 1)));
#pragma unroll
            for (int k = 0; k < K; ++k) f[k] = prim_eval(p.kind, p.cx, p.cy, p.cz, p.a, p.b, p.c, x[k], y[k], z[k]);
            t = 1;
        }
        // the next record is read ahead without a branch (the last one twice)
        for (; t < cnt; ++t) {
            const Prim p = pn;
            pn = load_prim(cp, __builtin_amdgcn_readlane(mine, min(t + 1, cnt - 1)));
#pragma unroll
            for (int k = 0; k < K; ++k)
                f[k] = fold(p.op, f[k], prim_eval(p.kind, p.cx, p.cy, p.cz, p.a, p.b, p.c, x[k], y[k], z[k]));
        }
    }
}

// ---------------------------------------------------------------- culling scan
// A clamp x -> min(max(x,l),u).  c2(c1(x)) is the clamp (c2(l1), c2(u1)).
struct Clamp2 {
    float ll, lu, hl, hu;  // clamps of the lower and upper fold bounds
};
__device__ __forceinline__ float capply(float x, float l, float u) { return fminf(fmaxf(x, l), u); }
__device__ __forceinline__ Clamp2 ccompose(const Clamp2& a, const Clamp2& b)  // a first, then b
{
    return {capply(a.ll, b.ll, b.lu), capply(a.lu, b.ll, b.lu), capply(a.hl, b.hl, b.hu), capply(a.hu, b.hl, b.hu)};
}
__device__ __forceinline__ Clamp2 cshfl_up(const Clamp2& v, int d)
{
    return {__shfl_up(v.ll, d), __shfl_up(v.lu, d), __shfl_up(v.hl, d), __shfl_up(v.hu, d)};
}

// Interval of one primitive's fp32 values over a box given by centre and
// half diagonal; `slack` (1e-4 of the largest coordinate or size magnitude)
// covers fp32 rounding of both evaluations and the few-ulp error of the
// centre value evaluated with bare hardware square roots.
__device__ __forceinline__ void prim_interval(const Prim& p, float cx, float cy, float cz, float rad, float slack,
                                              float& lo, float& hi)
{
    float v = prim_eval<true>(p.kind, p.cx, p.cy, p.cz, p.a, p.b, p.c, cx, cy, cz);
    lo = v - rad - slack;
    hi = v + rad + slack;
}

// Culling rule.  Consecutive steps of one op form a run, and the fold over a
// run is a selection over its operands (min for union, max for subtraction
// and intersection, plus the incoming value), which returns the extreme value
// and, among operands equal to it, the last.  An operand that is strictly
// worse than some other operand of its run everywhere in the box is never
// that value, so dropping it leaves every result bit unchanged; a run's
// interval after its last step bounds the best other operand.  Runs are cut at
// wave boundaries (each cut run is itself a run) and primitive 0, which
// starts the fold, is its own run and always kept.

// Inclusive scan value at the end of this lane's run within the wave.
__device__ __forceinline__ Clamp2 run_end(const Clamp2& inc, int op, bool first, bool active, bool has_next)
{
    const int lane = threadIdx.x & 63;
    const int me = first ? 3 : op;
    const int nx = __shfl_down(me, 1);
    const bool end = active && (lane == 63 || !has_next || nx != me || me == 3);
    const uint64_t ends = __ballot(end);
    const uint64_t after = ends >> lane;
    const int e = after ? lane + __builtin_ctzll(after) : lane;
    return {__shfl(inc.ll, e), __shfl(inc.lu, e), __shfl(inc.hl, e), __shfl(inc.hu, e)};
}

// Keep an operand with value interval [lo, hi] whose run ends with the fold
// interval [rlo, rhi]: union needs lo <= rhi, subtraction -lo >= rlo,
// intersection hi >= rlo.
__device__ __forceinline__ bool keep_step(int op, bool first, float lo, float hi, float rlo, float rhi)
{
    if (first) return true;
    if (op == 0) return !(lo > rhi);
    if (op == 1) return !(-lo < rlo);
    return !(hi < rlo);
}

// Wave-level refinement of a primitive list for a smaller box (a group of
// blocks or one block).  src: the parent list (its first entry is primitive 0)
// or nullptr for all primitives; out: per-wave LDS array (may be src: entry j
// is read before any entry at or past the write position is written).  The
// exact rule above, scanned 64 candidates at a time (runs of the list,
// which join across dropped steps of other ops).
static __device__ int cull_wave(const Prim* __restrict__ prims, const uint16_t* src, int nsrc, float cx, float cy, float cz,
                         float rad, float slack, uint16_t* out)
{
    const int lane = threadIdx.x & 63;
    float Flo = 0.0f, Fhi = 0.0f;
    int n = 0;
    for (int base = 0; base < nsrc; base += 64) {
        const int j = base + lane;
        Clamp2 cl = {-INFINITY, INFINITY, -INFINITY, INFINITY};
        float lo = 0.0f, hi = 0.0f;
        int op = 0, idx = 0;
        if (j < nsrc) {
            idx = src ? (int)src[j] : j;
            const Prim p = prims[idx];
            prim_interval(p, cx, cy, cz, rad, slack, lo, hi);
            op = p.op;
            if (j == 0) cl = {lo, lo, hi, hi};
            else if (op == 0) cl = {-INFINITY, lo, -INFINITY, hi};
            else if (op == 1) cl = {-hi, INFINITY, -lo, INFINITY};
            else cl = {lo, INFINITY, hi, INFINITY};
        }
        Clamp2 inc = cl;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            Clamp2 o = cshfl_up(inc, d);
            if (lane >= d) inc = ccompose(o, inc);
        }
        const Clamp2 run = run_end(inc, op, j == 0, j < nsrc, j + 1 < nsrc);
        const float rlo = capply(Flo, run.ll, run.lu), rhi = capply(Fhi, run.hl, run.hu);
        const bool keep = j < nsrc && keep_step(op, j == 0, lo, hi, rlo, rhi);
        const uint64_t bal = __ballot(keep);
        if (keep) out[n + __popcll(bal & ((1ull << lane) - 1))] = (uint16_t)idx;
        n += __popcll(bal);
        const float all_ll = __shfl(inc.ll, 63), all_lu = __shfl(inc.lu, 63);
        const float all_hl = __shfl(inc.hl, 63), all_hu = __shfl(inc.hu, 63);
        Flo = capply(Flo, all_ll, all_lu);
        Fhi = capply(Fhi, all_hl, all_hu);
    }
    return n;
}
