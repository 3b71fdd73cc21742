// Lossless leaf-box packing of a dense fp32 field on gfx950, plus leaf
// unpacking, point lookups and the trilinear query on packed fields.
//
// Reference semantics (FE/adaptive_sample_blocks_v1.py:19-58 and its native
// twin FE/adaptive_pack_native_v1/pack.cpp:18-55): a node is a box; along
// each axis it is "invariant" when every sample's bit pattern equals the one
// in the node's first slab, and then stores one slab; its direct cost is
// 52 + 4*prod(stored).  A node with prod(stored) <= 64 or at most
// `leaf_samples` samples is a leaf; otherwise it is bisected at shape//2 along
// the largest non-invariant axis (first on ties) and kept split only when
// 52 + cost(left) + cost(right) < direct cost.  The output is the pre-order
// (left first) node table of 52-byte records and the payload of every leaf's
// stored samples in C order.
//
// GPU formulation: the candidate split tree is built top-down one level per
// pass (invariance flags by wave-ballot AND reductions over each candidate's
// samples, split decisions and child creation per node), the keep-or-collapse
// decisions and subtree sizes bottom-up, the pre-order indices and payload
// offsets top-down, and a final pass writes records and copies payloads.
// Every decision depends only on exact bit comparisons and integer costs, so
// the table is identical to the reference's.
#include "fe_common.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

namespace {

struct PNode {
    int32_t lo[3], shape[3];
    int32_t inv;  // bit a set: axis a invariant
    int32_t stored[3];
    int32_t split;  // kept split
    int64_t child;  // arena index of the left child (right = child + 1), -1: none
    int64_t reduced, direct, cost, nn, pay, idx, off;
    int32_t reach, pad;
};

__device__ __forceinline__ int64_t node_count(const PNode& p) { return (int64_t)p.shape[0] * p.shape[1] * p.shape[2]; }

// AND-reduction of the three invariance flags over the nodes of one level
// ([lv.x, lv.x + lv.y) of the arena), S parts per node, grid-stride.
__global__ __launch_bounds__(256) void k_pack_inv(PNode* __restrict__ arena, const int64_t* __restrict__ lv, int S,
                                                  const uint32_t* __restrict__ d, int64_t sx, int64_t sy)
{
    const int64_t start = lv[0], n = lv[1];
    for (int64_t it = blockIdx.x; it < n * S; it += gridDim.x) {
        PNode& P = arena[start + it / S];
        const int part = (int)(it % S);
        const int64_t cnt = node_count(P);
        const int64_t b = cnt * part / S, e = cnt * (part + 1) / S;
        const int hy = P.shape[1], hz = P.shape[2];
        const int64_t o = (int64_t)P.lo[0] * sx + (int64_t)P.lo[1] * sy + P.lo[2];
        bool ex = true, ey = true, ez = true;
        for (int64_t l = b + threadIdx.x; l < e; l += 256) {
            const int z = (int)(l % hz);
            const int64_t r = l / hz;
            const int y = (int)(r % hy), x = (int)(r / hy);
            const uint32_t v = d[o + x * sx + y * sy + z];
            ex = ex && v == d[o + y * sy + z];
            ey = ey && v == d[o + x * sx + z];
            ez = ez && v == d[o + x * sx + y * sy];
        }
        const int m = (__ballot(!ex) ? 0 : 1) | (__ballot(!ey) ? 0 : 2) | (__ballot(!ez) ? 0 : 4);
        if ((threadIdx.x & 63) == 0 && m != 7) atomicAnd(&P.inv, m);
    }
}

// Run lengths of equal bits along each axis: r_a(p) = how many samples from p
// on (along +a) hold p's bits, so a box is invariant along a exactly when
// every sample of its lower a-face has r_a >= the box's extent along a.
// uint16 (axes <= 4096).  z: a wave per (x, y) row, 64 z per ballot, rows
// walked from the end with the next break carried.
__global__ __launch_bounds__(256) void k_runs_z(const uint32_t* __restrict__ d, int nx, int ny, int nz,
                                                uint16_t* __restrict__ rz, unsigned long long* __restrict__ nonfinite)
{
    const int lane = threadIdx.x & 63;
    const int64_t rows = (int64_t)nx * ny;
    const int64_t nw = (int64_t)gridDim.x * 4;
    bool bad = false;  // a non-finite sample (this pass reads every sample once)
    for (int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); row < rows; row += nw) {
        const uint32_t* line = d + row * nz;
        uint16_t* out = rz + row * nz;
        int carry = nz - 1;  // first break at or after the chunk after this one
        for (int c0 = ((nz - 1) >> 6) << 6; c0 >= 0; c0 -= 64) {
            const int z = c0 + lane;
            bool brk = false;
            if (z < nz) {
                const uint32_t v = line[z];
                bad = bad || (v & 0x7f800000u) == 0x7f800000u;
                brk = z == nz - 1 || v != line[z + 1];
            }
            const uint64_t m = __ballot(brk);
            const uint64_t mine = lane ? m >> lane : m;
            const int nb = mine ? z + __builtin_ctzll(mine) : carry;
            if (z < nz) out[z] = (uint16_t)(nb - z + 1);
            if (m) carry = c0 + __builtin_ctzll(m);
        }
    }
    if (__ballot(bad) && lane == 0) atomicOr(nonfinite, 1ull);
}

// y and x: a thread per column, walked from the end (adjacent threads hold
// adjacent z, so every step's loads and stores are coalesced).
__global__ __launch_bounds__(256) void k_runs_yx(const uint32_t* __restrict__ d, int n_outer, int n_line, int nz,
                                                 int64_t s_outer, int64_t s_line, uint16_t* __restrict__ r)
{
    const int64_t cols = (int64_t)n_outer * nz;
    for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < cols; c += (int64_t)gridDim.x * 256) {
        const int64_t base = (c / nz) * s_outer + c % nz;
        int64_t o = base + (int64_t)(n_line - 1) * s_line;
        uint32_t next = d[o];
        int run = 1;
        r[o] = 1;
#pragma unroll 4
        for (int i = n_line - 2; i >= 0; --i) {
            o -= s_line;
            const uint32_t v = d[o];
            run = v == next ? run + 1 : 1;
            next = v;
            r[o] = (uint16_t)run;
        }
    }
}

// The three invariance flags of every node of one level from its faces'
// run lengths: z from r_z on the face z = lo, y from r_y on y = lo, x from
// r_x on x = lo.  S parts per node over the concatenated faces.
__global__ __launch_bounds__(256) void k_pack_inv_faces(PNode* __restrict__ arena, const int64_t* __restrict__ lv, int S,
                                                        const uint16_t* __restrict__ rx, const uint16_t* __restrict__ ry,
                                                        const uint16_t* __restrict__ rz, int64_t sx, int64_t sy)
{
    const int64_t start = lv[0], n = lv[1];
    for (int64_t it = blockIdx.x; it < n * S; it += gridDim.x) {
        PNode& P = arena[start + it / S];
        const int part = (int)(it % S);
        const int hx = P.shape[0], hy = P.shape[1], hz = P.shape[2];
        const int64_t fz = (int64_t)hx * hy, fy = (int64_t)hx * hz, fx = (int64_t)hy * hz;
        const int64_t cnt = fz + fy + fx;
        const int64_t b = cnt * part / S, e = cnt * (part + 1) / S;
        const int64_t o = (int64_t)P.lo[0] * sx + (int64_t)P.lo[1] * sy + P.lo[2];
        bool ex = true, ey = true, ez = true;
        // face positions stay below 3 * 4096^2 < 2^31: 32-bit division
        for (uint32_t l = (uint32_t)b + threadIdx.x; l < (uint32_t)e; l += 256) {
            if (l < (uint32_t)fz) {
                const uint32_t x = l / (uint32_t)hy, y = l - x * (uint32_t)hy;
                ez = ez && rz[o + x * sx + y * sy] >= hz;
            } else if (l < (uint32_t)(fz + fy)) {
                const uint32_t k = l - (uint32_t)fz;
                const uint32_t x = k / (uint32_t)hz, z = k - x * (uint32_t)hz;
                ey = ey && ry[o + x * sx + z] >= hy;
            } else {
                const uint32_t k = l - (uint32_t)(fz + fy);
                const uint32_t y = k / (uint32_t)hz, z = k - y * (uint32_t)hz;
                ex = ex && rx[o + y * sy + z] >= hx;
            }
        }
        const int m = (__ballot(!ex) ? 0 : 1) | (__ballot(!ey) ? 0 : 2) | (__ballot(!ez) ? 0 : 4);
        if ((threadIdx.x & 63) == 0 && m != 7) atomicAnd(&P.inv, m);
    }
}

__device__ void init_node(PNode& c, const int* lo, const int* shape)
{
    for (int a = 0; a < 3; ++a) {
        c.lo[a] = lo[a];
        c.shape[a] = shape[a];
        c.stored[a] = 0;
    }
    c.inv = 7;
    c.child = -1;
    c.split = 0;
    c.reach = 0;
    c.reduced = c.direct = c.cost = c.nn = c.pay = c.idx = c.off = 0;
}

// Leaf-or-candidate decision of every node of level l; a candidate's two
// children are appended at the arena cursor (ctl[0]); ctl[1] is set when the
// arena is full (the caller then packs again with a larger one).
__global__ __launch_bounds__(256) void k_pack_decide(PNode* __restrict__ arena, const int64_t* __restrict__ lv, int leaf,
                                                     int64_t cap, unsigned long long* __restrict__ ctl)
{
    const int64_t start = lv[0], n = lv[1];
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        PNode& P = arena[start + i];
        int64_t reduced = 1;
        for (int a = 0; a < 3; ++a) {
            P.stored[a] = ((P.inv >> a) & 1) ? 1 : P.shape[a];
            reduced *= P.stored[a];
        }
        P.reduced = reduced;
        P.direct = 52 + 4 * reduced;
        P.child = -1;
        if (reduced <= 64 || node_count(P) <= leaf) continue;
        int axis = -1;
        for (int a = 0; a < 3; ++a)
            if (P.stored[a] > 1 && (axis < 0 || P.shape[a] > P.shape[axis])) axis = a;
        if (axis < 0) continue;
        const unsigned long long c = atomicAdd(&ctl[0], 2ull);
        if ((int64_t)c + 2 > cap) {
            atomicOr(&ctl[1], 1ull);
            continue;
        }
        int ls[3], rs[3], rl[3];
        for (int a = 0; a < 3; ++a) {
            ls[a] = rs[a] = P.shape[a];
            rl[a] = P.lo[a];
        }
        ls[axis] = P.shape[axis] / 2;
        rs[axis] -= ls[axis];
        rl[axis] += ls[axis];
        init_node(arena[c], P.lo, ls);
        init_node(arena[c + 1], rl, rs);
        P.child = (int64_t)c;
    }
}

// Range of level l + 1: the children appended after level l.
__global__ void k_pack_level(int64_t* __restrict__ lv, const unsigned long long* __restrict__ ctl, int64_t cap)
{
    const int64_t s = lv[0] + lv[1];
    lv[2] = s;
    lv[3] = max((int64_t)0, min((int64_t)ctl[0], cap) - s);
}

__global__ __launch_bounds__(256) void k_pack_bottom(PNode* __restrict__ arena, const int64_t* __restrict__ lv)
{
    const int64_t start = lv[0], n = lv[1];
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        PNode& P = arena[start + i];
        P.split = 0;
        P.cost = P.direct;
        P.nn = 1;
        P.pay = P.reduced;
        if (P.child < 0) continue;
        const PNode& L = arena[P.child];
        const PNode& R = arena[P.child + 1];
        const int64_t s = 52 + L.cost + R.cost;
        if (s < P.direct) {
            P.split = 1;
            P.cost = s;
            P.nn = 1 + L.nn + R.nn;
            P.pay = L.pay + R.pay;
        }
    }
}

__global__ __launch_bounds__(256) void k_pack_top(PNode* __restrict__ arena, const int64_t* __restrict__ lv)
{
    const int64_t start = lv[0], n = lv[1];
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const PNode& P = arena[start + i];
        if (!P.reach || !P.split) continue;
        PNode& L = arena[P.child];
        PNode& R = arena[P.child + 1];
        L.reach = R.reach = 1;
        L.idx = P.idx + 1;
        L.off = P.off;
        R.idx = P.idx + 1 + L.nn;
        R.off = P.off + L.pay;
    }
}

// Copies rows (x, y) of a final leaf's stored box into the payload: lanes run
// along z in groups of the next power of two >= the stored z extent (at most
// 256), the groups of workgroup slot g (of gstep) over the rows, so a lane
// divides once per row it copies (stored extents <= 4096).
__device__ __forceinline__ void emit_rows(const PNode& P, uint32_t* __restrict__ payload, const uint32_t* __restrict__ d,
                                          int64_t sx, int64_t sy, uint32_t g, uint32_t gstep)
{
    const uint32_t s1 = (uint32_t)P.stored[1], s2 = (uint32_t)P.stored[2];
    const uint32_t rows = (uint32_t)P.stored[0] * s1;
    const int lg = s2 >= 256 ? 8 : (s2 <= 1 ? 0 : 32 - __clz(s2 - 1));
    const uint32_t zl = threadIdx.x & ((1u << lg) - 1u), G = 256u >> lg;
    const int64_t o = (int64_t)P.lo[0] * sx + (int64_t)P.lo[1] * sy + P.lo[2];
    uint32_t* dst = payload + P.off;
    for (uint32_t r = g * G + (threadIdx.x >> lg); r < rows; r += gstep * G) {
        const uint32_t x = r / s1, y = r - x * s1;
        const uint32_t* src = d + o + x * sx + y * sy;
        for (uint32_t z = zl; z < s2; z += 1u << lg) dst[(int64_t)r * s2 + z] = src[z];
    }
}

#define EMIT_BIG 4096  // stored samples: larger leaves are copied by k_pack_emit_big

// Writes the 52-byte record (as 13 words) of every node of the final tree in
// one level and copies each final leaf's stored samples into the payload (a
// leaf of more than EMIT_BIG stored samples is listed in big, counted by
// *nbig, for k_pack_emit_big).
__global__ __launch_bounds__(256) void k_pack_emit(const PNode* __restrict__ arena, const int64_t* __restrict__ lv,
                                                   uint32_t* __restrict__ rec, uint32_t* __restrict__ payload,
                                                   const uint32_t* __restrict__ d, int64_t sx, int64_t sy,
                                                   int64_t* __restrict__ big, unsigned long long* __restrict__ nbig)
{
    const int64_t start = lv[0], n = lv[1];
    for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
        const PNode& P = arena[start + i];
        if (!P.reach) continue;
        if (threadIdx.x == 0) {
            uint32_t* r = rec + P.idx * 13;
            for (int a = 0; a < 3; ++a) {
                r[a] = (uint32_t)P.lo[a];
                r[3 + a] = (uint32_t)P.shape[a];
                r[6 + a] = P.split ? 0u : (uint32_t)P.stored[a];
            }
            int64_t off = -1;
            if (P.split) {
                const PNode& L = arena[P.child];
                r[9] = (uint32_t)(P.idx + 1);
                r[10] = (uint32_t)(P.idx + 1 + L.nn);
            } else {
                r[9] = r[10] = 0xffffffffu;
                off = P.off;
            }
            r[11] = (uint32_t)((uint64_t)off & 0xffffffffu);
            r[12] = (uint32_t)((uint64_t)off >> 32);
            if (!P.split && P.reduced > EMIT_BIG) big[atomicAdd(nbig, 1ull)] = start + i;
        }
        if (P.split || P.reduced > EMIT_BIG) continue;
        emit_rows(P, payload, d, sx, sy, 0, 1);
    }
}

// The listed large leaves: K = grid / n workgroups per leaf (at least one),
// so a few huge leaves spread over the whole grid and many moderate ones run
// side by side.
__global__ __launch_bounds__(256) void k_pack_emit_big(const PNode* __restrict__ arena, const int64_t* __restrict__ big,
                                                       const unsigned long long* __restrict__ nbig,
                                                       uint32_t* __restrict__ payload, const uint32_t* __restrict__ d,
                                                       int64_t sx, int64_t sy)
{
    const unsigned long long n = *nbig;
    if (n == 0) return;
    const uint32_t K = (uint32_t)max(1ull, (unsigned long long)gridDim.x / n);
    const uint32_t lanes = gridDim.x / K;  // leaves in flight at once
    for (unsigned long long i = blockIdx.x / K; i < n && blockIdx.x / K < lanes; i += lanes)
        emit_rows(arena[big[i]], payload, d, sx, sy, blockIdx.x % K, K);
}

__global__ __launch_bounds__(256) void k_nonfinite(const uint32_t* __restrict__ d, int64_t n,
                                                   unsigned long long* __restrict__ flag)
{
    bool bad = false;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        bad = bad || (d[i] & 0x7f800000u) == 0x7f800000u;
    if (__ballot(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1ull);
}

struct Packed {
    int64_t n_nodes = 0, n_samples = 0;
    uint32_t* nodes = nullptr;    // device, 13 words per record
    uint32_t* payload = nullptr;  // device
    ~Packed()
    {
        if (nodes) (void)hipFree(nodes);
        if (payload) (void)hipFree(payload);
    }
};

// Work buffers of pack_device kept between calls (per device, grown as
// needed): the candidate-node arena, the level ranges and three counters.
constexpr size_t PACK_RUNS_MAX = size_t(8) << 30;  // bytes of run lengths (6 per sample) held for a pack

struct PackArena {
    char* p = nullptr;
    size_t n = 0;
};
std::mutex arena_lock;
PackArena arenas[64];

hipError_t pack_arena(size_t need, char** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    PackArena& a = arenas[dev];
    if (a.n < need) {
        if (a.p && (e = hipFree(a.p)) != hipSuccess) return e;
        a.p = nullptr;
        a.n = 0;
        if ((e = hipMalloc(&a.p, need)) != hipSuccess) return e;
        a.n = need;
    }
    *out = a.p;
    return hipSuccess;
}

int ceil_log2(int n)
{
    int k = 0;
    while ((1 << k) < n) ++k;
    return k;
}

// Device pack of `data` (device, C order, nx*ny*nz fp32).  The candidate
// split tree is built level by level in one node arena with the level ranges
// kept on the device, so the levels run back to back: the host reads the
// tree's sizes (and the non-finite flag) once before allocating the output,
// then waits once for the records and payload.  A tree larger than the arena
// (rare: the first arena holds 4 nodes per leaf-sized piece of the grid) is
// built again in one large enough for any tree.
Packed* pack_device(const float* data, int nx, int ny, int nz, int leaf, hipStream_t st, int* status)
{
    *status = FE_RUNTIME;
    const int64_t N = (int64_t)nx * ny * nz;
    const uint32_t* d = reinterpret_cast<const uint32_t*>(data);
    const int64_t sx = (int64_t)ny * nz, sy = nz;
    // every split halves one axis, so no chain of splits is longer than this
    const int depth = ceil_log2(nx) + ceil_log2(ny) + ceil_log2(nz) + 1;
    const int64_t pieces = (N + leaf) / (leaf + 1) + 1;
    std::lock_guard<std::mutex> hold(arena_lock);
    for (int attempt = 0; attempt < 2; ++attempt) {
        const int64_t cap = attempt == 0 ? 4 * pieces + 1024 : 2 * (int64_t)depth * pieces + 1024;
        const size_t lv_bytes = (size_t)(depth + 2) * 16, ctl_bytes = 32;
        const size_t nodes_bytes = ((size_t)cap * sizeof(PNode) + 255) & ~(size_t)255;
        // leaves above EMIT_BIG stored samples are disjoint: at most N / EMIT_BIG
        const size_t big_bytes = ((size_t)(N / EMIT_BIG + 2) * 8 + 255) & ~(size_t)255;
        // run lengths (2 bytes per sample and axis) when they fit; else each
        // level scans its nodes' samples
        const bool runs = (size_t)6 * (size_t)N <= PACK_RUNS_MAX;
        char* base = nullptr;
        if (hipError_t e = pack_arena(ctl_bytes + lv_bytes + nodes_bytes + big_bytes + (runs ? (size_t)6 * N : 0),
                                      &base)) {
            fe_note_error(e, "pack arena");
            return nullptr;
        }
        unsigned long long* ctl = reinterpret_cast<unsigned long long*>(base);  // [0] cursor [1] overflow [2] non-finite
                                                                                 // [3] large leaves listed
        int64_t* lv = reinterpret_cast<int64_t*>(base + ctl_bytes);            // per level {start, count}
        PNode* arena = reinterpret_cast<PNode*>(base + ctl_bytes + lv_bytes);
        int64_t* big = reinterpret_cast<int64_t*>(base + ctl_bytes + lv_bytes + nodes_bytes);
        uint16_t* rz = runs ? reinterpret_cast<uint16_t*>(base + ctl_bytes + lv_bytes + nodes_bytes + big_bytes) : nullptr;
        uint16_t* ry = runs ? rz + N : nullptr;
        uint16_t* rx = runs ? ry + N : nullptr;
        PNode root;
        std::memset(&root, 0, sizeof root);
        root.shape[0] = nx;
        root.shape[1] = ny;
        root.shape[2] = nz;
        root.inv = 7;
        root.child = -1;
        root.reach = 1;
        const unsigned long long ctl0[4] = {1, 0, 0, 0};
        std::vector<int64_t> lv0((size_t)(depth + 2) * 2, 0);
        lv0[1] = 1;
        hipError_t e = hipSuccess;
        auto ok = [&](hipError_t r, const char* where) {
            if (r != hipSuccess && e == hipSuccess) {
                e = r;
                fe_note_error(r, where);
            }
            return e == hipSuccess;
        };
        if (!ok(hipMemcpyAsync(ctl, ctl0, sizeof ctl0, hipMemcpyHostToDevice, st), "pack init") ||
            !ok(hipMemcpyAsync(lv, lv0.data(), lv0.size() * 8, hipMemcpyHostToDevice, st), "pack init") ||
            !ok(hipMemcpyAsync(arena, &root, sizeof root, hipMemcpyHostToDevice, st), "pack init"))
            return nullptr;
        const unsigned grid_nodes = 2048;
        if (!runs)
            hipLaunchKernelGGL(k_nonfinite, dim3((unsigned)std::min<int64_t>((N + 255) / 256, 4096)), dim3(256), 0, st, d,
                               N, ctl + 2);
        if (runs) {  // the z run lengths also flag non-finite samples
            hipLaunchKernelGGL(k_runs_z, dim3((unsigned)std::min<int64_t>(((int64_t)nx * ny + 3) / 4, 65536)), dim3(256),
                               0, st, d, nx, ny, nz, rz, ctl + 2);
            const int64_t cy = (int64_t)nx * nz, cx = (int64_t)ny * nz;
            hipLaunchKernelGGL(k_runs_yx, dim3((unsigned)std::min<int64_t>((cy + 255) / 256, 65536)), dim3(256), 0, st, d,
                               nx, ny, nz, sx, sy, ry);
            hipLaunchKernelGGL(k_runs_yx, dim3((unsigned)std::min<int64_t>((cx + 255) / 256, 65536)), dim3(256), 0, st, d,
                               ny, nx, nz, sy, sx, rx);
        }
        for (int l = 0; l < depth; ++l) {
            // parts per node: about 2 K face samples (with run lengths; else
            // 16 K samples) each at this level's node size, so the top levels'
            // few large nodes still fill the GPU
            const int64_t per = std::max<int64_t>(1, N >> std::min(l, 62));
            if (runs) {
                const double face = 3.0 * std::pow((double)per, 2.0 / 3.0);
                const int S = (int)std::min(4096.0, std::max(1.0, std::ceil(face / 2048.0)));
                hipLaunchKernelGGL(k_pack_inv_faces, dim3(grid_nodes * 4), dim3(256), 0, st, arena, lv + 2 * l, S, rx, ry,
                                   rz, sx, sy);
            } else {
                const int S = (int)std::min<int64_t>(4096, (per + 16383) / 16384);
                hipLaunchKernelGGL(k_pack_inv, dim3(grid_nodes * 4), dim3(256), 0, st, arena, lv + 2 * l, S, d, sx, sy);
            }
            hipLaunchKernelGGL(k_pack_decide, dim3(grid_nodes), dim3(256), 0, st, arena, lv + 2 * l, leaf, cap, ctl);
            hipLaunchKernelGGL(k_pack_level, dim3(1), dim3(1), 0, st, lv + 2 * l, ctl, cap);
        }
        for (int l = depth - 1; l >= 0; --l)
            hipLaunchKernelGGL(k_pack_bottom, dim3(grid_nodes), dim3(256), 0, st, arena, lv + 2 * l);
        int64_t sizes[2];
        unsigned long long flags[4];
        if (!ok(hipGetLastError(), "pack levels") ||
            !ok(hipMemcpyAsync(sizes, &arena[0].nn, 16, hipMemcpyDeviceToHost, st), "pack sizes") ||
            !ok(hipMemcpyAsync(flags, ctl, sizeof flags, hipMemcpyDeviceToHost, st), "pack sizes") ||
            !ok(hipStreamSynchronize(st), "pack sizes"))
            return nullptr;
        if (flags[2]) {
            *status = FE_INVALID;
            return nullptr;
        }
        if (flags[1]) continue;  // the arena was too small: again with the bound
        Packed* out = new (std::nothrow) Packed();
        if (!out) return nullptr;
        out->n_nodes = sizes[0];
        out->n_samples = sizes[1];
        if (!ok(hipMalloc(&out->nodes, (size_t)out->n_nodes * 52), "pack output") ||
            !ok(hipMalloc(&out->payload, (size_t)std::max<int64_t>(1, out->n_samples) * 4), "pack output")) {
            delete out;
            return nullptr;
        }
        for (int l = 0; l + 1 < depth; ++l)
            hipLaunchKernelGGL(k_pack_top, dim3(grid_nodes), dim3(256), 0, st, arena, lv + 2 * l);
        for (int l = 0; l < depth; ++l)
            hipLaunchKernelGGL(k_pack_emit, dim3(grid_nodes * 4), dim3(256), 0, st, arena, lv + 2 * l, out->nodes,
                               out->payload, d, sx, sy, big, ctl + 3);
        hipLaunchKernelGGL(k_pack_emit_big, dim3(grid_nodes), dim3(256), 0, st, arena, big, ctl + 3, out->payload, d,
                           sx, sy);
        if (!ok(hipGetLastError(), "pack emit") || !ok(hipStreamSynchronize(st), "pack emit")) {
            delete out;
            return nullptr;
        }
        *status = FE_OK;
        return out;
    }
    fe_note_error(hipErrorOutOfMemory, "pack arena bound");
    return nullptr;
}

// ---------------------------------------------------------------- lookups
struct Node52 {
    int32_t lo[3], shape[3], stored[3], children[2];
    int32_t off_lo, off_hi;
};
static_assert(sizeof(Node52) == 52, "node record");

__device__ __forceinline__ int64_t node_offset(const Node52& n)
{
    return (int64_t)(((uint64_t)(uint32_t)n.off_hi << 32) | (uint32_t)n.off_lo);
}

// Leaf of (x, y, z) by tree descent (FE/adaptive_sample_blocks_v1.py:103-128):
// into the first child when its box holds the point, else the second.
__device__ __forceinline__ int leaf_of(const Node52* __restrict__ nodes, int x, int y, int z)
{
    int id = 0;
    for (;;) {
        const Node52& nd = nodes[id];
        if (nd.children[0] < 0) return id;
        const Node52& l = nodes[nd.children[0]];
        const bool in = x >= l.lo[0] && x < l.lo[0] + l.shape[0] && y >= l.lo[1] && y < l.lo[1] + l.shape[1] &&
                        z >= l.lo[2] && z < l.lo[2] + l.shape[2];
        id = in ? nd.children[0] : nd.children[1];
    }
}

// Bits of (x, y, z) in leaf nd (an axis stored once reads its one sample).
__device__ __forceinline__ uint32_t leaf_bits(const Node52& nd, const uint32_t* __restrict__ payload, int x, int y, int z)
{
    const int64_t lx = nd.stored[0] == 1 ? 0 : x - nd.lo[0];
    const int64_t ly = nd.stored[1] == 1 ? 0 : y - nd.lo[1];
    const int64_t lz = nd.stored[2] == 1 ? 0 : z - nd.lo[2];
    return payload[node_offset(nd) + (lx * nd.stored[1] + ly) * nd.stored[2] + lz];
}

__device__ __forceinline__ uint32_t lookup(const Node52* __restrict__ nodes, const uint32_t* __restrict__ payload, int x,
                                           int y, int z)
{
    return leaf_bits(nodes[leaf_of(nodes, x, y, z)], payload, x, y, z);
}

__global__ __launch_bounds__(256) void k_samples_at(const Node52* __restrict__ nodes, const uint32_t* __restrict__ payload,
                                                    const int64_t* __restrict__ idx, int64_t n, uint32_t* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    out[i] = lookup(nodes, payload, (int)idx[3 * i], (int)idx[3 * i + 1], (int)idx[3 * i + 2]);
}

// Trilinear query (FE/adaptive_sample_blocks_v1.py:130-151): f64 grid
// coordinates clipped to [0, n-1], eight clamped neighbours, and the order-1
// spline sum of scipy.ndimage.map_coordinates (w0 = 1-t, w1 = 1-w0; each term
// value*w_x*w_y*w_z, summed with the last axis fastest), rounded to fp32.
__global__ __launch_bounds__(256) void k_query(const Node52* __restrict__ nodes, const uint32_t* __restrict__ payload,
                                               int nx, int ny, int nz, double ox, double oy, double oz, double pitch,
                                               const double* __restrict__ pts, int64_t n, float* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int shape[3] = {nx, ny, nz};
    const double org[3] = {ox, oy, oz};
    int lower[3];
    double w[3][2];
    for (int a = 0; a < 3; ++a) {
        double c = (pts[3 * i + a] - org[a]) / pitch;
        const double hi = (double)(shape[a] - 1);
        if (c < 0.0) c = 0.0;
        if (c > hi) c = hi;
        const double fl = floor(c);
        lower[a] = (int)fl;
        double t = c - fl;
        t -= floor(t);
        w[a][0] = 1.0 - t;
        w[a][1] = 1.0 - w[a][0];
    }
    // one descent for the lower corner; a corner inside that leaf's box reads
    // it directly (the pack's tree splits every node into two children that
    // partition it, so the corner's own descent would end in the same leaf)
    const Node52& base = nodes[leaf_of(nodes, lower[0], lower[1], lower[2])];
    double acc = 0.0;
    for (int h = 0; h < 8; ++h) {
        const int bx = (h >> 2) & 1, by = (h >> 1) & 1, bz = h & 1;
        const int qx = min(lower[0] + bx, nx - 1), qy = min(lower[1] + by, ny - 1), qz = min(lower[2] + bz, nz - 1);
        const bool same = qx >= base.lo[0] && qx < base.lo[0] + base.shape[0] && qy >= base.lo[1] &&
                          qy < base.lo[1] + base.shape[1] && qz >= base.lo[2] && qz < base.lo[2] + base.shape[2];
        const uint32_t bits = same ? leaf_bits(base, payload, qx, qy, qz) : lookup(nodes, payload, qx, qy, qz);
        double coeff = (double)__uint_as_float(bits);
        coeff *= 1.0;
        coeff *= w[0][bx];
        coeff *= w[1][by];
        coeff *= w[2][bz];
        acc += coeff;
    }
    out[i] = (float)acc;
}

// Writes voxels [j0, j1) (z fastest) of a leaf's clipped box (hx, hy, hz),
// threads striding by `step`.
// Rows (x, y) of a leaf's box (clipped to hx x hy x hz) to the dense grid:
// lanes run along z in groups of the next power of two >= hz (at most 256),
// the groups of workgroup slot g (of gstep) over the rows, so a lane divides
// once per row it writes (every extent <= 4096).
__device__ __forceinline__ void unpack_rows(const Node52& nd, const uint32_t* __restrict__ payload, int hx, int hy,
                                            int hz, int ny, int nz, uint32_t g, uint32_t gstep,
                                            uint32_t* __restrict__ dense)
{
    const int64_t off = node_offset(nd);
    const int lg = hz >= 256 ? 8 : (hz <= 1 ? 0 : 32 - __clz(hz - 1));
    const uint32_t zl = threadIdx.x & ((1u << lg) - 1u), G = 256u >> lg;
    const uint32_t rows = (uint32_t)hx * (uint32_t)hy;
    const int s1 = nd.stored[1], s2 = nd.stored[2];
    const bool bx = nd.stored[0] == 1, by = s1 == 1, bz = s2 == 1;
    for (uint32_t r = g * G + (threadIdx.x >> lg); r < rows; r += gstep * G) {
        const uint32_t x = r / (uint32_t)hy, y = r - x * (uint32_t)hy;
        uint32_t* dst = dense + ((int64_t)(nd.lo[0] + (int)x) * ny + nd.lo[1] + (int)y) * nz + nd.lo[2];
        const uint32_t* src = payload + off + ((int64_t)(bx ? 0 : x) * s1 + (by ? 0 : y)) * s2;
        for (uint32_t z = zl; z < (uint32_t)hz; z += 1u << lg) dst[z] = src[bz ? 0 : z];
    }
}

#define UNPACK_BIG 8192       // voxels: larger leaves go to k_unpack_big
#define UNPACK_HUGE (1 << 20)  // voxels: larger leaves each take the whole grid there

// Dense reconstruction of a packed field: one workgroup per leaf writes its
// box; leaves of more than UNPACK_BIG voxels are listed in big[1 + i]
// (big[0] counts them) for k_unpack_big.
__global__ __launch_bounds__(256) void k_unpack(const Node52* __restrict__ nodes, int64_t count,
                                                const uint32_t* __restrict__ payload, int nx, int ny, int nz,
                                                uint32_t* __restrict__ dense, unsigned int* __restrict__ big)
{
    const int64_t id = blockIdx.x;
    if (id >= count) return;
    const Node52& nd = nodes[id];
    if (nd.children[0] >= 0) return;
    const int hx = min(nd.shape[0], nx - nd.lo[0]), hy = min(nd.shape[1], ny - nd.lo[1]), hz = min(nd.shape[2], nz - nd.lo[2]);
    if (hx <= 0 || hy <= 0 || hz <= 0) return;
    const int64_t cnt = (int64_t)hx * hy * hz;
    if (cnt > UNPACK_BIG) {
        // two lists of count + 1 words each: [0] = length, then leaf ids
        unsigned int* list = cnt > UNPACK_HUGE ? big + count + 1 : big;
        if (threadIdx.x == 0) list[1 + atomicAdd(list, 1u)] = (unsigned int)id;
        return;
    }
    unpack_rows(nd, payload, hx, hy, hz, ny, nz, 0, 1, dense);
}

// The listed leaves: the huge ones (second list) one after another, each
// over every workgroup; then the large ones with K = grid / n workgroups per
// leaf (at least one), side by side.
__global__ __launch_bounds__(256) void k_unpack_big(const Node52* __restrict__ nodes, int64_t count,
                                                    const uint32_t* __restrict__ payload, int nx, int ny, int nz,
                                                    uint32_t* __restrict__ dense, const unsigned int* __restrict__ big)
{
    auto rows = [&](unsigned int id, uint32_t g, uint32_t gstep) {
        const Node52& nd = nodes[id];
        const int hx = min(nd.shape[0], nx - nd.lo[0]), hy = min(nd.shape[1], ny - nd.lo[1]),
                  hz = min(nd.shape[2], nz - nd.lo[2]);
        unpack_rows(nd, payload, hx, hy, hz, ny, nz, g, gstep, dense);
    };
    const unsigned int* huge = big + count + 1;
    for (unsigned int i = 0; i < huge[0]; ++i) rows(huge[1 + i], blockIdx.x, gridDim.x);
    const unsigned int n = big[0];
    if (n == 0) return;
    const uint32_t K = max(1u, gridDim.x / n);
    const uint32_t lanes = gridDim.x / K;  // leaves in flight at once
    if (blockIdx.x / K >= lanes) return;
    for (unsigned int i = blockIdx.x / K; i < n; i += lanes) rows(big[1 + i], blockIdx.x % K, K);
}

struct HostPacked {
    std::vector<uint8_t> nodes;
    std::vector<uint32_t> payload;
};

}  // namespace

// ---------------------------------------------------------------- device ABI
extern "C" void* fe_pack_device(const float* data, int nx, int ny, int nz, int leaf, int* status, void* stream) noexcept
{
    int st_local = FE_INVALID;
    int* s = status ? status : &st_local;
    *s = FE_INVALID;
    if (!data || nx < 1 || ny < 1 || nz < 1 || nx > 4096 || ny > 4096 || nz > 4096 || leaf < 8 || leaf > 4096)
        return nullptr;
    try {
        return pack_device(data, nx, ny, nz, leaf, static_cast<hipStream_t>(stream), s);
    } catch (...) {
        *s = FE_RUNTIME;
        return nullptr;
    }
}

extern "C" int64_t fe_pack_nodes(const void* h) noexcept { return h ? static_cast<const Packed*>(h)->n_nodes : 0; }
extern "C" int64_t fe_pack_samples(const void* h) noexcept { return h ? static_cast<const Packed*>(h)->n_samples : 0; }

// Copies a device pack into caller buffers (device or host; hipMemcpyDefault).
extern "C" int fe_pack_copy(const void* h, void* nodes, int64_t count, void* payload, int64_t samples, void* stream) noexcept
{
    const Packed* p = static_cast<const Packed*>(h);
    if (!p || !nodes || !payload || count != p->n_nodes || samples != p->n_samples) return FE_INVALID;
    hipStream_t st = static_cast<hipStream_t>(stream);
    FE_TRY(hipMemcpyAsync(nodes, p->nodes, (size_t)count * 52, hipMemcpyDefault, st));
    if (samples) FE_TRY(hipMemcpyAsync(payload, p->payload, (size_t)samples * 4, hipMemcpyDefault, st));
    FE_TRY(hipStreamSynchronize(st));
    return FE_OK;
}

extern "C" void fe_pack_free(void* h) noexcept { delete static_cast<Packed*>(h); }

// Packed-field lookups on the device: nodes (count x 52 bytes) and payload are device pointers.
extern "C" int fe_samples_at(const void* nodes, const uint32_t* payload, const int64_t* idx, int64_t n, uint32_t* out,
                             void* stream) noexcept
{
    if (!nodes || !payload || !idx || !out || n < 0) return FE_INVALID;
    if (n == 0) return FE_OK;
    hipLaunchKernelGGL(k_samples_at, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const Node52*>(nodes), payload, idx, n, out);
    FE_TRY(hipGetLastError());
    return FE_OK;
}

extern "C" int fe_query_packed(const void* nodes, const uint32_t* payload, int nx, int ny, int nz, const double* origin,
                               double pitch, const double* pts, int64_t n, float* out, void* stream) noexcept
{
    if (!nodes || !payload || !origin || !pts || !out || n < 0 || nx < 1 || ny < 1 || nz < 1) return FE_INVALID;
    if (n == 0) return FE_OK;
    hipLaunchKernelGGL(k_query, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const Node52*>(nodes), payload, nx, ny, nz, origin[0], origin[1], origin[2], pitch,
                       pts, n, out);
    FE_TRY(hipGetLastError());
    return FE_OK;
}

extern "C" int fe_unpack(const void* nodes, int64_t count, const uint32_t* payload, int nx, int ny, int nz, float* dense,
                         void* stream) noexcept
{
    if (!nodes || !payload || !dense || count < 1 || count > INT32_MAX || nx < 1 || ny < 1 || nz < 1)
        return FE_INVALID;
    hipStream_t st = static_cast<hipStream_t>(stream);
    // the two large-leaf lists (at most one entry per leaf each) live in the pack arena,
    // held for the whole call
    std::lock_guard<std::mutex> hold(arena_lock);
    char* base = nullptr;
    FE_TRY(pack_arena((size_t)(count + 1) * 8, &base));
    unsigned int* big = reinterpret_cast<unsigned int*>(base);
    FE_TRY(hipMemsetAsync(big, 0, 4, st));
    FE_TRY(hipMemsetAsync(big + count + 1, 0, 4, st));
    hipLaunchKernelGGL(k_unpack, dim3((unsigned)count), dim3(256), 0, st, static_cast<const Node52*>(nodes), count,
                       payload, nx, ny, nz, reinterpret_cast<uint32_t*>(dense), big);
    hipLaunchKernelGGL(k_unpack_big, dim3(2048), dim3(256), 0, st, static_cast<const Node52*>(nodes), count, payload, nx,
                       ny, nz, reinterpret_cast<uint32_t*>(dense), big);
    FE_TRY(hipGetLastError());
    // the arena may be reused by the next call once this stream is done
    FE_TRY(hipStreamSynchronize(st));
    return FE_OK;
}

// ---------------------------------------------------------------- reference-shaped host ABI
// FE/adaptive_pack_native_v1/pack.cpp:57-79: pack a host field, keep the
// result on the host until adaptive_copy, free with adaptive_free.
extern "C" void* adaptive_pack(const void* data, const int32_t* shape, int threshold) noexcept
{
    try {
        if (!data || !shape || threshold < 8 || threshold > 4096) return nullptr;
        int64_t total = 1;
        for (int a = 0; a < 3; ++a) {
            if (shape[a] < 1 || shape[a] > 4096) return nullptr;
            total *= shape[a];
        }
        if (total > 16777216) return nullptr;
        float* d = nullptr;
        if (hipMalloc(&d, (size_t)total * 4) != hipSuccess) return nullptr;
        void* result = nullptr;
        if (hipMemcpy(d, data, (size_t)total * 4, hipMemcpyHostToDevice) == hipSuccess) {
            int status = 0;
            Packed* p = pack_device(d, shape[0], shape[1], shape[2], threshold, nullptr, &status);
            if (p) {
                HostPacked* h = new HostPacked();
                h->nodes.resize((size_t)p->n_nodes * 52);
                h->payload.resize((size_t)p->n_samples);
                if (hipMemcpy(h->nodes.data(), p->nodes, h->nodes.size(), hipMemcpyDeviceToHost) == hipSuccess &&
                    (h->payload.empty() ||
                     hipMemcpy(h->payload.data(), p->payload, h->payload.size() * 4, hipMemcpyDeviceToHost) == hipSuccess))
                    result = h;
                else
                    delete h;
                delete p;
            }
        }
        (void)hipFree(d);
        return result;
    } catch (...) {
        return nullptr;
    }
}

extern "C" int64_t adaptive_nodes(const void* pointer) noexcept
{
    return pointer ? (int64_t)static_cast<const HostPacked*>(pointer)->nodes.size() / 52 : 0;
}

extern "C" int64_t adaptive_samples(const void* pointer) noexcept
{
    return pointer ? (int64_t)static_cast<const HostPacked*>(pointer)->payload.size() : 0;
}

extern "C" int adaptive_copy(const void* pointer, void* nodes, int64_t count, void* payload, int64_t samples) noexcept
{
    if (!pointer || !nodes || !payload) return 1;
    const HostPacked& h = *static_cast<const HostPacked*>(pointer);
    if (count != (int64_t)h.nodes.size() / 52 || samples != (int64_t)h.payload.size()) return 1;
    std::memcpy(nodes, h.nodes.data(), h.nodes.size());
    std::memcpy(payload, h.payload.data(), h.payload.size() * 4);
    return 0;
}

extern "C" void adaptive_free(void* pointer) noexcept { delete static_cast<HostPacked*>(pointer); }
