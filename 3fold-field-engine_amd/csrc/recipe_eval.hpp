// Recipe records, field descriptors and the fp32 recipe evaluation shared by
// the block-sparse build (field.hip) and the per-block-pitch field
// (multires.hip).
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
        for (int t = 0; t < cnt; ++t) {
            const Prim p = pn;
            if (t + 1 < cnt) pn = load_prim(cp, __builtin_amdgcn_readlane(mine, t + 1));
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const float v = prim_eval(p.kind, p.cx, p.cy, p.cz, p.a, p.b, p.c, x[k], y[k], z[k]);
                f[k] = jb + t == 0 ? v : fold(p.op, f[k], v);
            }
        }
    }
}

