// Shared helpers of the gfx950 field-engine library (wave64, HIP runtime only).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define FE_WAVE 64

// Status codes of the reference C ABI (FE/native_edt_v1.py:18-33):
// 0 success, 1 invalid input (not latched), 2 runtime failure (latched by callers).
enum { FE_OK = 0, FE_INVALID = 1, FE_RUNTIME = 2 };

// Records the last runtime failure (file:line and the HIP error string) for
// fe_last_error(); defined in edt.hip.
void fe_note_error(hipError_t e, const char* where);
// Records a message for fe_last_error() (invalid input explained).
void fe_note_text(const char* msg);

#define FE_STR2(x) #x
#define FE_STR(x) FE_STR2(x)
#define FE_TRY(expr)                                              \
    do {                                                          \
        hipError_t fe_e_ = (expr);                                \
        if (fe_e_ != hipSuccess) {                                \
            fe_note_error(fe_e_, __FILE__ ":" FE_STR(__LINE__));  \
            return FE_RUNTIME;                                    \
        }                                                         \
    } while (0)

// Selection min/max with the NumPy/Warp tie rule (a tie returns the second
// operand, which matters for signed zeros).
__host__ __device__ __forceinline__ float sel_min(float a, float b) { return a < b ? a : b; }
__host__ __device__ __forceinline__ float sel_max(float a, float b) { return a > b ? a : b; }

// Correctly rounded fp32 square root (the same bits as sqrtf), branch-free:
// inputs below 2^-96 are scaled by 2^32 (result by 2^-16); v_sqrt_f32 is
// within one ulp, so the correctly rounded root is s or a neighbouring float,
// and the exact fma residuals x - s_dn*s <= 0 and x - s_up*s > 0 pick it (the
// root lies below the midpoint (s_dn + s)/2 exactly when x - s_dn*s <= 0, as
// that residual is a multiple of ulp^2).  This is the residual test of the
// compiler's own correctly rounded lowering without its final class check,
// which returns x for +-0 and +inf: those already come out as x here (for 0
// the lower neighbour is a NaN pattern and the upper residual is +0; for +inf
// both residuals are NaN), and NaN stays NaN.
__device__ __forceinline__ float sqrt_rn(float x)
{
    const bool scale = x < 0x1p-96f;
    const float xs = scale ? x * 0x1p32f : x;
    const float s = __builtin_amdgcn_sqrtf(xs);
    const float s_dn = __uint_as_float(__float_as_uint(s) - 1u);
    const float s_up = __uint_as_float(__float_as_uint(s) + 1u);
    float r = __builtin_fmaf(-s_dn, s, xs) <= 0.0f ? s_dn : s;
    r = __builtin_fmaf(-s_up, s, xs) > 0.0f ? s_up : r;
    return scale ? r * 0x1p-16f : r;
}

// Uniform 64-bit lane read (idx must be wave-uniform).
__device__ __forceinline__ uint64_t readlane64(uint64_t v, int idx)
{
    uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, idx);
    uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), idx);
    return ((uint64_t)hi << 32) | lo;
}

// Blocks are dispatched round-robin over the 8 XCDs of MI355X.  Remapping the
// hardware block id so that each XCD walks a contiguous range of logical tiles
// keeps neighbouring tiles (which share cache lines) in one L2.  `grid` must be
// a multiple of 8; logical ids >= the real tile count are idle.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t bid, uint32_t grid)
{
    uint32_t per = grid >> 3;
    return (bid & 7u) * per + (bid >> 3);
}

static inline uint32_t round8(uint64_t n) { return (uint32_t)((n + 7) & ~(uint64_t)7); }

// Logical tiles dealt to the XCDs in groups of G consecutive tiles,
// round-robin: every XCD gets work from the whole tile range, which keeps the
// XCDs balanced when the work per tile varies along the range (the solid of a
// part often fills only its middle), while G neighbouring tiles, which may
// share cache lines, stay in one L2.  `grid` must be a multiple of 8*G.
template <uint32_t G>
__device__ __forceinline__ uint32_t xcd_group_tile(uint32_t bid)
{
    const uint32_t q = bid >> 3, x = bid & 7u;
    return (q / G) * (8u * G) + x * G + q % G;
}

static inline uint32_t round_up(uint64_t n, uint32_t m) { return (uint32_t)((n + m - 1) / m * m); }
