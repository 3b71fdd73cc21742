// Tiled min/max CSG over leaf-box packed fields on gfx950, and the bounded
// leaf-box tile reader.
//
// Reference semantics (FE/csg_batch_native_v1/execute.cpp:10-54,
// FE/field_csg_tiled_v1.py:31-99): instruction i produces value F+i; union is
// a<b?a:b, difference a>(-b)?a:(-b), offset f32(f64(a)-d) (ties return the
// second operand, as NumPy's minimum/maximum); any non-finite result is an
// error.  Results are elementwise, so the tile shape does not change them;
// only the status code depends on the reference's (16,32,16) tile order, and
// that is reproduced exactly: the first failing tile in x,y,z order decides,
// and inside a tile the operand reads (status 3) precede the instructions
// (status 4).
//
// MI355X design: one wave per output tile of 1x8x64 samples.  The host
// compiles the program into micro-ops over LDS slots (linear-scan liveness,
// every input read once right before its first use).  Before the program, a
// wave per leaf appends the leaf to the lists of the tiles it overlaps; an
// input read then copies the parts of its tile's listed leaves into the slot
// with coalesced reads along z (a tile whose list overflows, or a launch whose
// lists would not fit, walks the field's box tree breadth-first with the whole
// workgroup instead).  Each instruction then runs over the slot arrays in
// LDS; only the final value is written to HBM.
#include "fe_common.hpp"
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <functional>
#include <mutex>
#include <new>
#include <vector>

namespace {

struct Node52 {
    int32_t lo[3], shape[3], stored[3], children[2];
    int32_t off_lo, off_hi;
};
static_assert(sizeof(Node52) == 52, "node record");

__host__ __device__ __forceinline__ int64_t noff(const Node52& n)
{
    return (int64_t)(((uint64_t)(uint32_t)n.off_hi << 32) | (uint32_t)n.off_lo);
}

struct FieldRef {  // one packed operand on the device
    const Node52* nodes;
    const uint32_t* payload;
    int64_t count, samples;
};

struct MOp {  // micro-op of the compiled program
    int32_t kind;  // 0 read input a into slot dst (b: the read's index), 1 union, 2 difference, 3 offset
    int32_t dst, a, b;
    double off;
};
static_assert(sizeof(MOp) == 24, "micro-op");

// Output tiles of 1 x 8 x 64 samples, one wave each (many tiles in flight per
// CU hide the dependent loads of a read); leaf lists and start nodes belong to
// list tiles of LX = 4 output tiles along x.
constexpr int TX = 1, TY = 8, TZ = 64, TS = TX * TY * TZ;  // 512 samples per tile
constexpr int LX = 4;                                       // list tile: LX x TY x TZ samples
constexpr int NTH = 64;                                     // threads per workgroup
constexpr int FRONT = 256;                                  // BFS frontier / leaf list capacity

// A load through a pointer known to be in device memory (global, not generic,
// instructions for pointers read out of descriptor tables)
template <class T>
__device__ __forceinline__ T gld(const T* p)
{
    return *(const __attribute__((address_space(1))) T*)p;
}

__device__ __forceinline__ bool box_overlap(const int* lo, const int* shape, const int* tlo, const int* thi)
{
    return lo[0] < thi[0] && lo[0] + shape[0] > tlo[0] && lo[1] < thi[1] && lo[1] + shape[1] > tlo[1] &&
           lo[2] < thi[2] && lo[2] + shape[2] > tlo[2];
}

__device__ uint32_t descend(const Node52* __restrict__ nodes, const uint32_t* __restrict__ payload, int x, int y, int z)
{
    int id = 0;
    for (;;) {
        const Node52& nd = nodes[id];
        if (nd.children[0] < 0) {
            const int64_t lx = nd.stored[0] == 1 ? 0 : x - nd.lo[0];
            const int64_t ly = nd.stored[1] == 1 ? 0 : y - nd.lo[1];
            const int64_t lz = nd.stored[2] == 1 ? 0 : z - nd.lo[2];
            return payload[noff(nd) + (lx * nd.stored[1] + ly) * nd.stored[2] + lz];
        }
        const Node52& l = nodes[nd.children[0]];
        const bool in = x >= l.lo[0] && x < l.lo[0] + l.shape[0] && y >= l.lo[1] && y < l.lo[1] + l.shape[1] &&
                        z >= l.lo[2] && z < l.lo[2] + l.shape[2];
        id = in ? nd.children[0] : nd.children[1];
    }
}

// Deepest node of a tree whose subtree holds every leaf overlapping the box
// [tlo, thi): from the root, down while exactly one child's box overlaps it.
// Child indices are checked (in range, after the parent) before a program
// runs, so the walk ends.
__device__ int tile_start(const Node52* __restrict__ nodes, const int* tlo, const int* thi)
{
    int id = 0;
    for (;;) {
        const Node52& nd = nodes[id];
        if (nd.children[0] < 0) return id;
        const Node52& c0 = nodes[nd.children[0]];
        const Node52& c1 = nodes[nd.children[1]];
        const bool o0 = box_overlap(c0.lo, c0.shape, tlo, thi), o1 = box_overlap(c1.lo, c1.shape, tlo, thi);
        if (o0 == o1) return id;
        id = o0 ? nd.children[0] : nd.children[1];
    }
}

// Leaf lists of every (read, list tile), LCAP words each: word 0 counts the
// leaves overlapping the list tile, words 1 .. LCAP-1 hold the first LCAP-1 of
// them (in no particular order: the leaves of a checked operand are
// disjoint).  A count >= LCAP marks a list that did not fit; the reads of its
// tiles walk the tree.
constexpr int LCAP = 64;

// Start nodes of every (read, list tile): one thread each, before the program
// runs, so a tile's breadth-first read starts below the levels that hold it
// whole.  With leaf lists (built first) only the tiles whose list overflowed
// walk the tree.
__global__ __launch_bounds__(256) void k_csg_starts(const FieldRef* __restrict__ fields, const int* __restrict__ reads,
                                                    int nreads, int nx, int ny, int nz, int gty, int gtz, int nlt,
                                                    const int* __restrict__ lists, int* __restrict__ starts)
{
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= (int64_t)nreads * nlt) return;
    if (lists && lists[g * LCAP] < LCAP) return;  // never read
    const int r = (int)(g / nlt), t = (int)(g % nlt);
    const int tz = t % gtz, ty = (t / gtz) % gty, tx = t / (gtz * gty);
    const int tlo[3] = {tx * LX, ty * TY, tz * TZ};
    const int thi[3] = {min(tlo[0] + LX, nx), min(tlo[1] + TY, ny), min(tlo[2] + TZ, nz)};
    starts[g] = tile_start(fields[reads[r]].nodes, tlo, thi);
}

// Zeroes the count word of every leaf list.
__global__ __launch_bounds__(256) void k_csg_list_clear(int* __restrict__ lists, int64_t n)
{
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g < n) lists[g * LCAP] = 0;
}

// One wave per (read, node record): a leaf appends itself to the lists of the
// list tiles its box (clipped to the grid) overlaps, one list tile per lane (a
// leaf of the exterior may cover thousands of them).
__global__ __launch_bounds__(256) void k_csg_lists(const FieldRef* __restrict__ fields, const int* __restrict__ reads,
                                                   int nx, int ny, int nz, int gty, int gtz, int nlt,
                                                   int* __restrict__ lists)
{
    const int r = blockIdx.y;
    const FieldRef F = fields[reads[r]];
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= F.count) return;
    const Node52& nd = F.nodes[i];
    if (nd.children[0] >= 0) return;
    const int n[3] = {nx, ny, nz}, T[3] = {LX, TY, TZ};
    int t0[3], nt[3];
    for (int a = 0; a < 3; ++a) {
        const int64_t lo = max((int64_t)nd.lo[a], (int64_t)0);
        const int64_t hi = min((int64_t)nd.lo[a] + nd.shape[a], (int64_t)n[a]);
        if (hi <= lo) return;
        t0[a] = (int)(lo / T[a]);
        nt[a] = (int)((hi - 1) / T[a]) - t0[a] + 1;
    }
    int* base = lists + (int64_t)r * nlt * LCAP;
    const int total = nt[0] * nt[1] * nt[2];  // <= 1024 * 512 * 64
    for (int k = lane; k < total; k += 64) {
        const int tz = t0[2] + k % nt[2], ty = t0[1] + (k / nt[2]) % nt[1], tx = t0[0] + k / (nt[2] * nt[1]);
        int* L = base + ((int64_t)(tx * gty + ty) * gtz + tz) * LCAP;
        const int c = atomicAdd(L, 1);
        if (c < LCAP - 1) L[1 + c] = (int)i;
    }
}

// Copies the part of leaf nd inside the tile [tlo, thi) into the slot.  A part
// has at most TX*TY = 8 rows (x, y) of at most TZ = 64 samples along z: lanes
// run along z in groups of the next power of two >= the row length, the
// groups over the rows, so the only division is row / ey (ey <= 8, by a
// multiply and shift, exact for rows < 2^11).
__device__ __forceinline__ void copy_leaf_part(const Node52& nd, const uint32_t* __restrict__ payload, const int* tlo,
                                               const int* thi, float* slot)
{
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = max(tlo[a], nd.lo[a]);
        hi[a] = min(thi[a], nd.lo[a] + nd.shape[a]);
    }
    const int ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    if (ex <= 0 || ey <= 0 || ez <= 0) return;
    const int lg = ez <= 1 ? 0 : 32 - __clz(ez - 1);  // group width 2^lg >= ez
    const int zl = (int)threadIdx.x & ((1 << lg) - 1);
    if (zl >= ez) return;
    const int s1 = nd.stored[1], s2 = nd.stored[2];
    const bool bx = nd.stored[0] == 1, by = s1 == 1;
    const int z = lo[2] + zl;
    const int64_t sz = s2 == 1 ? 0 : z - nd.lo[2];
    const uint32_t* src = payload + noff(nd) + sz;
    float* dst = slot + (z - tlo[2]);
    const uint32_t inv = (65536u + (uint32_t)ey - 1u) / (uint32_t)ey;  // ceil(2^16 / ey)
    const int rows = ex * ey;
    for (int r = (int)threadIdx.x >> lg; r < rows; r += NTH >> lg) {
        const int rx = (int)(((uint32_t)r * inv) >> 16), ry = r - rx * ey;
        const int x = lo[0] + rx, y = lo[1] + ry;
        const int64_t sx = bx ? 0 : x - nd.lo[0], sy = by ? 0 : y - nd.lo[1];
        dst[((x - tlo[0]) * TY + (y - tlo[1])) * TZ] = __uint_as_float(src[(sx * s1 + sy) * s2]);
    }
}

// A leaf record as a read of a listed tile uses it: lane l holds leaf l's.
struct LeafRec {
    int lo0, lo1, lo2, sh0, sh1, sh2, st0, st1, st2, offl, offh;
};

// Loads the records of the nl <= 63 leaves of a list tile (lane l: the l-th
// listed record id in `myid`) at once.
__device__ __forceinline__ LeafRec load_leaves(const FieldRef& F, int myid, int nl)
{
    LeafRec q = {0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0};
    if ((int)threadIdx.x < nl) {
        const int32_t* nd = reinterpret_cast<const int32_t*>(F.nodes + myid);
        q.lo0 = gld(nd); q.lo1 = gld(nd + 1); q.lo2 = gld(nd + 2);
        q.sh0 = gld(nd + 3); q.sh1 = gld(nd + 4); q.sh2 = gld(nd + 5);
        q.st0 = gld(nd + 6); q.st1 = gld(nd + 7); q.st2 = gld(nd + 8);
        q.offl = gld(nd + 11); q.offh = gld(nd + 12);
    }
    return q;
}

// Finds the leaf of each of the TY samples (x, y, z) of lane z in the tile
// [tlo, thi) and loads the samples together into v.  Each lane first derives
// its own leaf's terms for the tile (overlap, rows as a bit mask, z range,
// address of row tlo[1] less the z term); the leaves that touch the tile then
// broadcast those terms in turn (few scalar instructions per leaf: the scalar
// unit is shared by a CU's waves) and the rows are selected without branches.
// Returns the rows found (bit k: row k); samples no leaf covers (an operand
// with gaps, whose status already fails the program) are left out.
__device__ __forceinline__ uint32_t gather_leaves(const FieldRef& F, const LeafRec& q, int nl, const int* tlo,
                                                  const int* thi, uint32_t* v)
{
    static_assert(NTH == 64 && TX == 1 && TZ == 64, "one wave per tile, a lane per z");
    const int lane = threadIdx.x;
    const int x = tlo[0];
    bool use = false;
    uint32_t rm = 0;
    int zlo = 0, zhi = 0, rs = 0, zs = 0;
    int64_t base = 0;
    if (lane < nl) {
        const int y0 = max(q.lo1, tlo[1]), y1 = min(q.lo1 + q.sh1, thi[1]);
        use = x >= q.lo0 && x < q.lo0 + q.sh0 && y0 < y1 && q.lo2 < thi[2] && q.lo2 + q.sh2 > tlo[2];
        if (use) rm = ((1u << (y1 - tlo[1])) - 1u) & ~((1u << (y0 - tlo[1])) - 1u);
        rs = q.st1 == 1 ? 0 : q.st2;
        zs = q.st2 == 1 ? 0 : 1;
        zlo = q.lo2;
        zhi = q.lo2 + q.sh2;
        const int64_t off = (int64_t)(((uint64_t)(uint32_t)q.offh << 32) | (uint32_t)q.offl);
        base = off + (int64_t)(q.st0 == 1 ? 0 : x - q.lo0) * q.st1 * q.st2 + (int64_t)(tlo[1] - q.lo1) * rs -
               (int64_t)zs * q.lo2;
    }
    const int z = tlo[2] + lane;
    const bool zin = z < thi[2];
    int64_t addr[TY];
    uint32_t have = 0;
#pragma unroll
    for (int r = 0; r < TY; ++r) addr[r] = 0;
    for (uint64_t todo = __ballot(use); todo; todo &= todo - 1) {
        const int l = __builtin_ctzll(todo);
        const int ZL = __builtin_amdgcn_readlane(zlo, l), ZH = __builtin_amdgcn_readlane(zhi, l);
        const int RS = __builtin_amdgcn_readlane(rs, l), ZS = __builtin_amdgcn_readlane(zs, l);
        const uint32_t RM = (uint32_t)__builtin_amdgcn_readlane((int)rm, l);
        const int64_t B = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(base >> 32), l) << 32) |
                                    (uint32_t)__builtin_amdgcn_readlane((int)base, l));
        const uint32_t lm = zin && z >= ZL && z < ZH ? RM : 0u;
        int64_t a = B + (int64_t)ZS * z;
        have |= lm;
#pragma unroll
        for (int r = 0; r < TY; ++r) {
            addr[r] = (lm >> r & 1u) ? a : addr[r];
            a += RS;
        }
    }
#pragma unroll
    for (int r = 0; r < TY; ++r) v[r] = (have >> r & 1u) ? gld(F.payload + addr[r]) : 0u;
    return have;
}

// Writes gathered rows into a slot.
__device__ __forceinline__ void put_rows(float* slot, const uint32_t* v, uint32_t have)
{
#pragma unroll
    for (int r = 0; r < TY; ++r)
        if (have >> r & 1u) slot[r * TZ + threadIdx.x] = __uint_as_float(v[r]);
}

// Reads field F over the tile [tlo, thi) from the nl <= 63 leaves of its list
// tile: the lanes load the records at once, then each lane loads its TY
// samples together, so a read waits for two loads in a row whatever the
// number of leaves.
__device__ __forceinline__ void read_listed(const FieldRef& F, const int* tlo, const int* thi, float* slot, int myid,
                                           int nl)
{
    const LeafRec q = load_leaves(F, myid, nl);
    uint32_t v[TY];
    const uint32_t have = gather_leaves(F, q, nl, tlo, thi, v);
    put_rows(slot, v, have);
}

// Reads field f over the tile [tlo, thi) into slot (layout x,y,z with TY*TZ, TZ
// strides): the workgroup copies the part of every leaf on the tile's list,
// or, without a list that fits, a breadth-first walk from node `start`
// collects the leaves overlapping the tile first.
__device__ __forceinline__ void read_tile(const FieldRef& F, const int* tlo, const int* thi, float* slot, int* cur, int* nxt, int* leaves,
                          int* counters, int start, const int* __restrict__ list)
{
    const int tid = threadIdx.x;
    if (list) {
        const int myid = tid < LCAP - 1 ? list[1 + tid] : 0;  // loaded with the count
        const int nl = list[0];                                // uniform
        if (nl < LCAP) {
            read_listed(F, tlo, thi, slot, myid, nl);
            __syncthreads();
            return;
        }
    }
    const Node52& sn = F.nodes[start];
    if (sn.children[0] < 0) {  // the tile lies in one leaf: no walk (start is uniform)
        copy_leaf_part(sn, F.payload, tlo, thi, slot);
        __syncthreads();
        return;
    }
    if (tid == 0) {
        counters[0] = 1;  // frontier size
        counters[2] = 0;  // leaves
        counters[3] = 0;  // overflow
        cur[0] = start;
    }
    __syncthreads();
    for (;;) {
        const int n = counters[0];
        if (n == 0 || counters[3]) break;
        if (tid == 0) counters[1] = 0;
        __syncthreads();
        for (int i = tid; i < n; i += NTH) {
            const Node52& nd = F.nodes[cur[i]];
            if (nd.children[0] < 0) {
                const int k = atomicAdd(&counters[2], 1);
                if (k < FRONT) leaves[k] = cur[i];
                else counters[3] = 1;
            } else {
                for (int c = 0; c < 2; ++c) {
                    const int ch = nd.children[c];
                    const Node52& cn = F.nodes[ch];
                    if (box_overlap(cn.lo, cn.shape, tlo, thi)) {
                        const int k = atomicAdd(&counters[1], 1);
                        if (k < FRONT) nxt[k] = ch;
                        else counters[3] = 1;
                    }
                }
            }
        }
        __syncthreads();
        for (int i = tid; i < counters[1] && i < FRONT; i += NTH) cur[i] = nxt[i];
        if (tid == 0) counters[0] = counters[1];
        __syncthreads();
    }
    const int dx = thi[0] - tlo[0], dy = thi[1] - tlo[1], dz = thi[2] - tlo[2];
    if (counters[3]) {  // too many boxes for the lists: per-sample descent
        for (int j = tid; j < dx * dy * dz; j += NTH) {
            const int z = j % dz, y = (j / dz) % dy, x = j / (dz * dy);
            slot[(x * TY + y) * TZ + z] = __uint_as_float(descend(F.nodes, F.payload, tlo[0] + x, tlo[1] + y, tlo[2] + z));
        }
    } else {
        const int nl = counters[2];
        for (int l = 0; l < nl; ++l) copy_leaf_part(F.nodes[leaves[l]], F.payload, tlo, thi, slot);
    }
    __syncthreads();
}

// Reference tile index (x-major over (16,32,16) tiles) of sample (x,y,z).
__device__ __forceinline__ int ref_tile(int x, int y, int z, int nty, int ntz)
{
    return ((x >> 4) * nty + (y >> 5)) * ntz + (z >> 4);
}

// Persistent tiles: workgroup b runs output tiles b, b + G, ...  Slots below
// lds_slots live in LDS; the others (programs with more simultaneously live
// values than LDS holds) in this workgroup's spill_slots * TS floats of
// `arena`.  err_nf receives the first reference tile with a non-finite result.
template <bool SPILL>
__global__ __launch_bounds__(NTH, 5) void k_csg_tiles(const FieldRef* __restrict__ fields, const MOp* __restrict__ prog,
                                                   int nops, int final_slot, int lds_slots, int spill_slots,
                                                   float* __restrict__ arena, int nx, int ny, int nz, int gty, int gtz,
                                                   int ntiles, int nlt, int nreads, const int* __restrict__ reads,
                                                   const int* __restrict__ starts,
                                                   const int* __restrict__ lists, float* __restrict__ out,
                                                   int* __restrict__ err_nf)
{
    extern __shared__ float slots[];
    __shared__ int cur[FRONT], nxt[FRONT], leaves[FRONT];
    __shared__ int counters[4];
    __shared__ int bad_tile;
    const int tid = threadIdx.x;
    float* spill = arena + (size_t)blockIdx.x * spill_slots * TS;
    // without spill slots every slot is in LDS, and the accesses compile to
    // LDS instructions rather than generic ones
    auto slot = [&](int s) -> float* {
        if constexpr (SPILL) return s < lds_slots ? slots + s * TS : spill + (size_t)(s - lds_slots) * TS;
        else return slots + s * TS;
    };
    const int nty = (ny + 31) >> 5, ntz = (nz + 15) >> 4;
    const int lane = tid;  // one wave: lane = z offset, rows = y offsets
    if (tid == 0) bad_tile = INT_MAX;
    __syncthreads();
    // leaf-list words of the first PRE reads, fetched one tile ahead; each
    // tile then gathers those reads' samples together at its start, so that it
    // waits for two loads in a row (leaf records, then samples) for all of them
    constexpr int PRE = 2;
    const int npre = lists ? min(nreads, PRE) : 0;
    FieldRef pref[PRE];
#pragma unroll
    for (int r = 0; r < PRE; ++r)
        if (r < npre) pref[r] = fields[reads[r]];
    int nid[PRE], ncnt[PRE];
#pragma unroll
    for (int r = 0; r < PRE; ++r) {
        nid[r] = 0;
        ncnt[r] = LCAP;
    }
    auto fetch = [&](int t) {
        const int tz = t % gtz, ty = (t / gtz) % gty, tx = t / (gtz * gty);
        const int lt = ((tx / LX) * gty + ty) * gtz + tz;
#pragma unroll
        for (int r = 0; r < PRE; ++r)
            if (r < npre) {
                const int* L = lists + ((int64_t)r * nlt + lt) * LCAP;
                nid[r] = lane < LCAP - 1 ? L[1 + lane] : 0;
                ncnt[r] = L[0];
            }
    };
    if ((int)blockIdx.x < ntiles) fetch(blockIdx.x);
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int tz = t % gtz, ty = (t / gtz) % gty, tx = t / (gtz * gty);
        const int tlo[3] = {tx * TX, ty * TY, tz * TZ};
        const int thi[3] = {min(tlo[0] + TX, nx), min(tlo[1] + TY, ny), min(tlo[2] + TZ, nz)};
        const int lt = ((tx / LX) * gty + ty) * gtz + tz;  // list tile
        const int rows = thi[1] - tlo[1];                   // <= TY
        const bool zin = tlo[2] + lane < thi[2];
        int cid[PRE], ccnt[PRE];
#pragma unroll
        for (int r = 0; r < PRE; ++r) {
            cid[r] = nid[r];
            ccnt[r] = __builtin_amdgcn_readfirstlane(ncnt[r]);
        }
        if (t + (int)gridDim.x < ntiles) fetch(t + gridDim.x);
        LeafRec q[PRE];
#pragma unroll
        for (int r = 0; r < PRE; ++r)
            if (r < npre && ccnt[r] < LCAP) q[r] = load_leaves(pref[r], cid[r], ccnt[r]);
        uint32_t gv[PRE][TY], gm[PRE];
#pragma unroll
        for (int r = 0; r < PRE; ++r)
            gm[r] = r < npre && ccnt[r] < LCAP ? gather_leaves(pref[r], q[r], ccnt[r], tlo, thi, gv[r]) : 0u;
        for (int o = 0; o < nops; ++o) {
            const MOp m = prog[o];
            float* dst = slot(m.dst);
            if (m.kind == 0) {
                if (m.b < npre) {  // uniform
                    int c = ccnt[0];
#pragma unroll
                    for (int r = 1; r < PRE; ++r)
                        if (m.b == r) c = ccnt[r];
                    if (c < LCAP) {  // gathered at the tile's start
#pragma unroll
                        for (int r = 0; r < PRE; ++r)
                            if (m.b == r) put_rows(dst, gv[r], gm[r]);
                        __syncthreads();
                        continue;
                    }
                }
                const int64_t g = (int64_t)m.b * nlt + lt;
                read_tile(fields[m.a], tlo, thi, dst, cur, nxt, leaves, counters, starts ? starts[g] : 0,
                          lists ? lists + g * LCAP : nullptr);
                continue;
            }
            // elementwise over every sample of the slots (dst may be one of the
            // operands' slots); samples outside the grid hold stale values,
            // computed but never flagged nor stored
            const float* A = slot(m.a);
            const float* Bv = slot(m.b);
            uint32_t nf = 0;  // bit k: the sample of row k is not finite
            if (m.kind == 3) {
                const double d = m.off;
#pragma unroll
                for (int k = 0; k < TY; ++k) {
                    const float r = (float)((double)A[k * TZ + lane] - d);
                    dst[k * TZ + lane] = r;
                    nf |= (uint32_t)!isfinite(r) << k;
                }
            } else if (m.kind == 1) {
#pragma unroll
                for (int k = 0; k < TY; ++k) {
                    const float a = A[k * TZ + lane], b = Bv[k * TZ + lane];
                    const float r = a < b ? a : b;
                    dst[k * TZ + lane] = r;
                    nf |= (uint32_t)!isfinite(r) << k;
                }
            } else {
#pragma unroll
                for (int k = 0; k < TY; ++k) {
                    const float a = A[k * TZ + lane], b = -Bv[k * TZ + lane];
                    const float r = a > b ? a : b;
                    dst[k * TZ + lane] = r;
                    nf |= (uint32_t)!isfinite(r) << k;
                }
            }
            nf &= zin ? (1u << rows) - 1u : 0u;
            // the reference tile index grows with y, so a lane's first row decides
            if (nf) atomicMin(&bad_tile, ref_tile(tlo[0], tlo[1] + __builtin_ctz(nf), tlo[2] + lane, nty, ntz));
            __syncthreads();
        }
        const float* F = slot(final_slot);
        if (zin) {
            float* o = out + ((int64_t)tlo[0] * ny + tlo[1]) * nz + tlo[2] + lane;
#pragma unroll
            for (int k = 0; k < TY; ++k)
                if (k < rows) o[(int64_t)k * nz] = F[k * TZ + lane];
        }
        __syncthreads();  // the next tile overwrites the slots
    }
    if (tid == 0 && bad_tile != INT_MAX) atomicMin(err_nf, bad_tile);
}

// Operand check with the reference gather's codes
// (FE/csg_tile_native_v1/gather.cpp:28-47), one workgroup per record:
// malformed leaf records, and internal records that do not point forward to
// records inside the table (the tile reader walks the tree), set `malformed`;
// every leaf marks the grid samples it covers, 32 per word: a bit of `once`
// on its first cover, of `twice` on any later one.
// Marks the samples of a leaf's box clipped to hx x hy x hz: each row (x, y)
// is a run of hz consecutive samples, one atomic per 32-sample word the run
// touches; lanes run over a row's words in groups of the next power of two
// >= the words per row (at most 256), the groups of workgroup slot g (of
// gstep) over the rows, with 32-bit row arithmetic (extents <= 4096).
__device__ __forceinline__ void mark_rows(const Node52& nd, int hx, int hy, int hz, int ny, int nz,
                                          uint32_t* __restrict__ once, uint32_t* __restrict__ twice, uint32_t g,
                                          uint32_t gstep)
{
    const int wpr = ((hz + 31) >> 5) + 1;  // >= 2
    const int lg = wpr >= 256 ? 8 : 32 - __clz(wpr - 1);
    const uint32_t kl = threadIdx.x & ((1u << lg) - 1u), G = 256u >> lg;
    const uint32_t rows = (uint32_t)hx * (uint32_t)hy;
    for (uint32_t r = g * G + (threadIdx.x >> lg); r < rows; r += gstep * G) {
        const uint32_t x = r / (uint32_t)hy, y = r - x * (uint32_t)hy;
        const int64_t s0 = ((int64_t)(nd.lo[0] + (int)x) * ny + nd.lo[1] + (int)y) * nz + nd.lo[2];
        const int64_t s1 = s0 + hz;  // exclusive
        for (uint32_t k = kl; k < (uint32_t)wpr; k += 1u << lg) {
            const int64_t w = (s0 >> 5) + k;
            if (w > ((s1 - 1) >> 5)) break;
            const int lo = (int)(max(s0, w << 5) - (w << 5)), hi = (int)(min(s1, (w << 5) + 32) - (w << 5));
            const uint32_t mask = (hi == 32 ? 0xffffffffu : ((1u << hi) - 1u)) & ~((1u << lo) - 1u);
            const uint32_t dup = atomicOr(&once[w], mask) & mask;
            if (dup) atomicOr(&twice[w], dup);
        }
    }
}

#define CHECK_BIG 65536  // word items (rows x words per row): larger leaves are marked by k_field_mark_big

// Bytes of the check buffers for n grid samples: covered-once and covered-
// twice bit arrays, then a count and up to check_big_cap(n) large-leaf ids.
__host__ __device__ inline int64_t check_big_cap(int64_t n) { return n / 32768 + 2; }
inline size_t check_bits_bytes(int64_t n) { return (size_t)((n + 31) >> 5) * 8 + (size_t)(check_big_cap(n) + 1) * 4; }

__global__ __launch_bounds__(256) void k_field_check(FieldRef F, int nx, int ny, int nz, uint32_t* __restrict__ once,
                                                     uint32_t* __restrict__ twice, int* __restrict__ malformed,
                                                     uint32_t* __restrict__ big)
{
    const int64_t i = blockIdx.x;
    if (i >= F.count) return;
    const Node52& nd = F.nodes[i];
    if (nd.children[0] >= 0) {
        if (threadIdx.x == 0 && (nd.children[0] <= i || nd.children[1] <= i || nd.children[0] >= F.count ||
                                 nd.children[1] >= F.count))
            atomicOr(malformed, 1);
        return;
    }
    int64_t size = 1;
    for (int a = 0; a < 3; ++a) {
        if (nd.lo[a] < 0 || nd.shape[a] < 1 || nd.shape[a] > 4096 || nd.lo[a] > 4096 - nd.shape[a] ||
            (nd.stored[a] != 1 && nd.stored[a] != nd.shape[a])) {
            if (threadIdx.x == 0) atomicOr(malformed, 1);
            return;
        }
        size *= nd.stored[a];
    }
    const int64_t off = noff(nd);
    if (off < 0 || size > F.samples || off > F.samples - size) {
        if (threadIdx.x == 0) atomicOr(malformed, 1);
        return;
    }
    const int hx = min(nd.shape[0], nx - nd.lo[0]), hy = min(nd.shape[1], ny - nd.lo[1]), hz = min(nd.shape[2], nz - nd.lo[2]);
    if (hx <= 0 || hy <= 0 || hz <= 0) return;
    // a large leaf goes to the list (while it has room: overlapping leaves
    // of a bad table may exceed the bound of disjoint ones)
    if ((int64_t)hx * hy * (((hz + 31) >> 5) + 1) > CHECK_BIG) {
        __shared__ int listed;
        if (threadIdx.x == 0) {
            const uint32_t c = atomicAdd(big, 1u);
            listed = c < (uint32_t)check_big_cap((int64_t)nx * ny * nz);
            if (listed) big[1 + c] = (uint32_t)i;
        }
        __syncthreads();
        if (listed) return;
    }
    mark_rows(nd, hx, hy, hz, ny, nz, once, twice, 0, 1);
}

// The listed large leaves, one after another, each spread over every block.
__global__ __launch_bounds__(256) void k_field_mark_big(FieldRef F, int nx, int ny, int nz, uint32_t* __restrict__ once,
                                                        uint32_t* __restrict__ twice, const uint32_t* __restrict__ big)
{
    const uint32_t n = min(big[0], (uint32_t)check_big_cap((int64_t)nx * ny * nz));
    for (uint32_t j = 0; j < n; ++j) {
        const Node52& nd = F.nodes[big[1 + j]];
        const int hx = min(nd.shape[0], nx - nd.lo[0]), hy = min(nd.shape[1], ny - nd.lo[1]),
                  hz = min(nd.shape[2], nz - nd.lo[2]);
        mark_rows(nd, hx, hy, hz, ny, nz, once, twice, blockIdx.x, gridDim.x);
    }
}

// First reference tile holding a sample covered other than exactly once.
__global__ __launch_bounds__(256) void k_cover_scan(const uint32_t* __restrict__ once, const uint32_t* __restrict__ twice,
                                                    int nx, int ny, int nz, int* __restrict__ err_gather)
{
    const int nty = (ny + 31) >> 5, ntz = (nz + 15) >> 4;
    const int64_t n = (int64_t)nx * ny * nz, words = (n + 31) >> 5;
    int bad = INT_MAX;
    for (int64_t w = (int64_t)blockIdx.x * 256 + threadIdx.x; w < words; w += (int64_t)gridDim.x * 256) {
        const int valid = (int)min((int64_t)32, n - (w << 5));
        uint32_t b = ~once[w] | twice[w];
        if (valid < 32) b &= (1u << valid) - 1u;
        while (b) {
            const int64_t s = (w << 5) + __builtin_ctz(b);
            b &= b - 1;
            const int z = (int)(s % nz);
            const int64_t r = s / nz;
            bad = min(bad, ref_tile((int)(r / ny), (int)(r % ny), z, nty, ntz));
        }
    }
    for (int d = 32; d >= 1; d >>= 1) bad = min(bad, __shfl_xor(bad, d));
    if ((threadIdx.x & 63) == 0 && bad != INT_MAX) atomicMin(err_gather, bad);
}

struct Instr24 {
    int32_t op, left, right, pad;
    double offset;
};
static_assert(sizeof(Instr24) == 24, "instruction layout");

// Program checks of FE/csg_batch_native_v1/execute.cpp:14-24; returns 0, 1 or 2
// and the last-use table.  capped: the reference's 2^24-sample limit.
int validate_program(int32_t fields, const int32_t* shape, const Instr24* program, int32_t n, bool capped,
                     std::vector<int>& last)
{
    if (!shape || !program || fields < 1 || fields > 256 || n < 1 || n > 1024) return 1;
    int64_t total = 1;
    for (int a = 0; a < 3; ++a) {
        if (shape[a] < 1 || shape[a] > 4096) return 1;
        total *= shape[a];
    }
    if (capped && total > 16777216) return 1;
    last.assign(fields + n, -1);
    for (int i = 0; i < n; ++i) {
        const Instr24& p = program[i];
        if (p.op < 0 || p.op > 2 || p.left < 0 || p.left >= fields + i) return 2;
        last[p.left] = i;
        if (p.op == 2) {
            if (!std::isfinite(p.offset)) return 2;
        } else {
            if (p.right < 0 || p.right >= fields + i) return 2;
            last[p.right] = i;
        }
    }
    last.back() = n;
    return 0;
}

// Evaluation order of the instructions.  Results are elementwise and every
// instruction still runs on every tile, so any order that puts operands
// first gives the same values and the same per-tile error flags.  This one
// keeps few values live at once: depth first from the results nobody uses
// (in program order, the final result last), and of two operands the one
// needing more slots (its Ershov number, counting shared values as if
// repeated) first.
std::vector<int> eval_order(int32_t fields, const Instr24* program, int32_t n, const std::vector<int>& last)
{
    std::vector<int> need(fields + n, 1);
    for (int i = 0; i < n; ++i) {
        const Instr24& p = program[i];
        const int na = need[p.left];
        if (p.op == 2) need[fields + i] = na;
        else {
            const int nb = need[p.right];
            need[fields + i] = na == nb ? na + 1 : std::max(na, nb);
        }
    }
    std::vector<char> state(n, 0);  // 0 new, 1 operands pending, 2 emitted
    std::vector<int> order, stack;
    order.reserve(n);
    auto visit = [&](int root) {
        stack.push_back(root);
        while (!stack.empty()) {
            const int i = stack.back();
            if (state[i] == 2) { stack.pop_back(); continue; }
            if (state[i] == 1) { state[i] = 2; order.push_back(i); stack.pop_back(); continue; }
            state[i] = 1;
            const Instr24& p = program[i];
            int first = p.left, second = p.op == 2 ? -1 : p.right;
            if (second >= 0 && need[second] > need[first]) std::swap(first, second);
            // pushed in reverse: `first` is evaluated first
            if (second >= fields && state[second - fields] == 0) stack.push_back(second - fields);
            if (first >= fields && state[first - fields] == 0) stack.push_back(first - fields);
        }
    };
    for (int i = 0; i + 1 < n; ++i)
        if (last[fields + i] < 0) visit(i);
    visit(n - 1);
    return order;
}

// Slot allocation over the evaluation order: an input is read into a slot
// right before its first use; a value's slot is free after its last use (the
// lowest free slot is taken first, so short-lived values stay in the low,
// LDS-resident slots); an instruction may write into the slot of an operand
// it is the last user of.  Returns the slot count.
int compile_program(int32_t fields, const Instr24* program, int32_t n, const std::vector<int>& order,
                    std::vector<MOp>& ops, int& final_slot)
{
    std::vector<int> lastpos(fields + n, -1);
    for (int k = 0; k < n; ++k) {
        const Instr24& p = program[order[k]];
        lastpos[p.left] = k;
        if (p.op != 2) lastpos[p.right] = k;
    }
    lastpos[fields + n - 1] = n;
    std::vector<int> slot(fields + n, -1);
    std::vector<int> free_slots;  // kept sorted descending: back() is the lowest
    int nslots = 0;
    auto take = [&]() {
        if (!free_slots.empty()) {
            const int s = free_slots.back();
            free_slots.pop_back();
            return s;
        }
        return nslots++;
    };
    auto release = [&](int s) {
        free_slots.insert(std::upper_bound(free_slots.begin(), free_slots.end(), s, std::greater<int>()), s);
    };
    for (int k = 0; k < n; ++k) {
        const int i = order[k];
        const Instr24& p = program[i];
        const int refs[2] = {p.left, p.op == 2 ? -1 : p.right};
        for (int r : refs)
            if (r >= 0 && r < fields && slot[r] < 0) {
                slot[r] = take();
                ops.push_back({0, slot[r], r, 0, 0.0});
            }
        const int sa = slot[p.left], sb = p.op == 2 ? 0 : slot[p.right];
        if (lastpos[p.left] == k) release(sa);
        if (p.op != 2 && p.right != p.left && lastpos[p.right] == k) release(sb);
        const int d = take();
        ops.push_back({p.op == 0 ? 1 : (p.op == 1 ? 2 : 3), d, sa, sb, p.op == 2 ? p.offset : 0.0});
        slot[fields + i] = d;
        if (lastpos[fields + i] < 0) release(d);
    }
    final_slot = slot[fields + n - 1];
    return nslots;
}

constexpr int LDS_SLOTS_MAX = 16;  // 32 KiB of slots (+ 3 KiB of tree lists) per workgroup
constexpr int LDS_SLOTS_SPILL = 8;  // LDS slots of a program that spills to the arena

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Launch geometry of a compiled program on the current device.
struct Plan {
    std::vector<MOp> ops;
    std::vector<int> reads;  // field of each read micro-op, in read order
    int nslots = 0, final_slot = 0, lds_slots = 0, spill = 0, grid = 0, ntiles = 0, nlt = 0, gty = 0, gtz = 0;
    bool starts = false;     // the start-node map fits (CSG_STARTS_MAX bytes)
    bool lists = false;      // the leaf lists fit (CSG_LISTS_MAX bytes)
    size_t lds = 0, ws_refs = 0, ws_ops = 0, ws_flags = 0, ws_reads = 0, ws_starts = 0, ws_lists = 0, ws_arena = 0,
           ws_total = 0;
};
constexpr size_t CSG_STARTS_MAX = size_t(1) << 28;  // start-node map bytes; beyond, reads start at the root
constexpr size_t CSG_LISTS_MAX = size_t(1) << 28;   // leaf-list bytes (LCAP words per read and list tile); beyond,
                                                    // every read walks the tree

hipError_t make_plan(int32_t fields, const Instr24* program, int32_t n, const std::vector<int>& last,
                     const int32_t* shape, Plan& P)
{
    const std::vector<int> order = eval_order(fields, program, n, last);
    P.nslots = compile_program(fields, program, n, order, P.ops, P.final_slot);
    P.lds_slots = P.nslots <= LDS_SLOTS_MAX ? P.nslots : LDS_SLOTS_SPILL;
    P.spill = P.nslots - P.lds_slots;
    P.lds = (size_t)P.lds_slots * TS * 4;
    const int gtx = (shape[0] + TX - 1) / TX;
    P.gty = (shape[1] + TY - 1) / TY;
    P.gtz = (shape[2] + TZ - 1) / TZ;
    const int64_t nt = (int64_t)gtx * P.gty * P.gtz;
    if (nt > INT_MAX) return hipErrorInvalidValue;
    P.ntiles = (int)nt;
    P.nlt = (int)((int64_t)((shape[0] + LX - 1) / LX) * P.gty * P.gtz);
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    const void* fn = P.spill ? reinterpret_cast<const void*>(&k_csg_tiles<true>)
                             : reinterpret_cast<const void*>(&k_csg_tiles<false>);
    if (P.lds > 65536 && (e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)P.lds)) != hipSuccess)
        return e;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, NTH, P.lds)) != hipSuccess) return e;
    P.grid = (int)std::min<int64_t>(nt, (int64_t)std::max(cus, 1) * std::max(per_cu, 1));
    P.ws_refs = align256(sizeof(FieldRef) * fields);
    P.ws_ops = align256(sizeof(MOp) * P.ops.size());
    P.ws_flags = 256;
    P.ws_arena = (size_t)P.grid * P.spill * TS * 4;
    for (MOp& m : P.ops)
        if (m.kind == 0) {
            m.b = (int)P.reads.size();
            P.reads.push_back(m.a);
        }
    P.ws_reads = align256(sizeof(int) * std::max<size_t>(P.reads.size(), 1));
    const size_t map = (size_t)4 * P.reads.size() * (size_t)P.nlt;
    P.starts = map <= CSG_STARTS_MAX;
    P.ws_starts = P.starts ? align256(map) : 0;
    const size_t lmap = map * LCAP;
    P.lists = !P.reads.empty() && P.reads.size() <= 65535 && lmap <= CSG_LISTS_MAX;
    P.ws_lists = P.lists ? align256(lmap) : 0;
    P.ws_total = P.ws_refs + P.ws_ops + P.ws_flags + P.ws_reads + P.ws_starts + P.ws_lists + P.ws_arena;
    return hipSuccess;
}

// Runs a planned program; first_bad: the first failing reference tile of the
// used operands' reads (INT_MAX: none).  Returns the status.
int run_plan(const Plan& P, const std::vector<FieldRef>& refs, const int32_t* shape, int first_bad, float* output,
             char* ws, hipStream_t st)
{
    FieldRef* dref = reinterpret_cast<FieldRef*>(ws);
    MOp* dops = reinterpret_cast<MOp*>(ws + P.ws_refs);
    int* flag = reinterpret_cast<int*>(ws + P.ws_refs + P.ws_ops);
    int* dreads = reinterpret_cast<int*>(ws + P.ws_refs + P.ws_ops + P.ws_flags);
    int* dstarts = P.starts ? reinterpret_cast<int*>(ws + P.ws_refs + P.ws_ops + P.ws_flags + P.ws_reads) : nullptr;
    int* dlists = P.lists ? reinterpret_cast<int*>(ws + P.ws_refs + P.ws_ops + P.ws_flags + P.ws_reads + P.ws_starts)
                          : nullptr;
    float* arena =
        reinterpret_cast<float*>(ws + P.ws_refs + P.ws_ops + P.ws_flags + P.ws_reads + P.ws_starts + P.ws_lists);
    const int none = INT_MAX;
    int nf = INT_MAX;
    FE_TRY(hipMemcpyAsync(dref, refs.data(), sizeof(FieldRef) * refs.size(), hipMemcpyHostToDevice, st));
    FE_TRY(hipMemcpyAsync(dops, P.ops.data(), sizeof(MOp) * P.ops.size(), hipMemcpyHostToDevice, st));
    FE_TRY(hipMemcpyAsync(flag, &none, sizeof none, hipMemcpyHostToDevice, st));
    const int64_t n = (int64_t)P.reads.size() * P.nlt;
    if ((dstarts || dlists) && n > 0)
        FE_TRY(hipMemcpyAsync(dreads, P.reads.data(), sizeof(int) * P.reads.size(), hipMemcpyHostToDevice, st));
    if (dlists && n > 0) {
        int64_t most = 1;
        for (int f : P.reads) most = std::max(most, refs[f].count);
        hipLaunchKernelGGL(k_csg_list_clear, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dlists, n);
        hipLaunchKernelGGL(k_csg_lists, dim3((unsigned)((most + 3) / 4), (unsigned)P.reads.size()), dim3(256), 0,
                           st, dref, dreads, shape[0], shape[1], shape[2], P.gty, P.gtz, P.nlt, dlists);
        FE_TRY(hipGetLastError());
    }
    if (dstarts && n > 0) {
        hipLaunchKernelGGL(k_csg_starts, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dref, dreads,
                           (int)P.reads.size(), shape[0], shape[1], shape[2], P.gty, P.gtz, P.nlt, dlists, dstarts);
        FE_TRY(hipGetLastError());
    }
    if (P.spill)
        hipLaunchKernelGGL(k_csg_tiles<true>, dim3((unsigned)P.grid), dim3(NTH), P.lds, st, dref, dops,
                           (int)P.ops.size(), P.final_slot, P.lds_slots, P.spill, arena, shape[0], shape[1], shape[2],
                           P.gty, P.gtz, P.ntiles, P.nlt, (int)P.reads.size(), dreads, dstarts, dlists, output,
                           flag);
    else
        hipLaunchKernelGGL(k_csg_tiles<false>, dim3((unsigned)P.grid), dim3(NTH), P.lds, st, dref, dops,
                           (int)P.ops.size(), P.final_slot, P.lds_slots, P.spill, arena, shape[0], shape[1], shape[2],
                           P.gty, P.gtz, P.ntiles, P.nlt, (int)P.reads.size(), dreads, dstarts, dlists, output,
                           flag);
    FE_TRY(hipGetLastError());
    FE_TRY(hipMemcpyAsync(&nf, flag, sizeof nf, hipMemcpyDeviceToHost, st));
    FE_TRY(hipStreamSynchronize(st));
    if (first_bad == INT_MAX && nf == INT_MAX) return 0;
    return first_bad <= nf ? 3 : 4;
}

// Checks the used operands on `stream` (cover bitmaps and the large-leaf list
// in `bits`, check_bits_bytes; results in res[f]); first failing reference
// tile per operand, 0 for a malformed table, INT_MAX when every sample is
// covered exactly once.
hipError_t check_operands(const std::vector<FieldRef>& refs, const std::vector<char>& used, const int32_t* shape,
                          uint32_t* bits, int* dres, std::vector<int>& res, hipStream_t st)
{
    const int nx = shape[0], ny = shape[1], nz = shape[2];
    const int64_t words = ((int64_t)nx * ny * nz + 31) >> 5;
    const int nf = (int)refs.size();
    std::vector<int> init(2 * nf, 0);
    for (int f = 0; f < nf; ++f) init[2 * f + 1] = INT_MAX;
    hipError_t e;
    if ((e = hipMemcpyAsync(dres, init.data(), sizeof(int) * 2 * nf, hipMemcpyHostToDevice, st)) != hipSuccess) return e;
    for (int f = 0; f < nf; ++f) {
        if (!used[f]) continue;
        // both bit arrays and the large-leaf count
        if ((e = hipMemsetAsync(bits, 0, (size_t)words * 8 + 4, st)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_field_check, dim3((unsigned)refs[f].count), dim3(256), 0, st, refs[f], nx, ny, nz, bits,
                           bits + words, dres + 2 * f, bits + 2 * words);
        hipLaunchKernelGGL(k_field_mark_big, dim3(2048), dim3(256), 0, st, refs[f], nx, ny, nz, bits, bits + words,
                           bits + 2 * words);
        hipLaunchKernelGGL(k_cover_scan, dim3((unsigned)std::min<int64_t>((words + 255) / 256, 4096)), dim3(256), 0,
                           st, bits, bits + words, nx, ny, nz, dres + 2 * f + 1);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    std::vector<int> got(2 * nf, 0);
    if ((e = hipMemcpyAsync(got.data(), dres, sizeof(int) * 2 * nf, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    res.assign(nf, INT_MAX);
    for (int f = 0; f < nf; ++f)
        if (used[f]) res[f] = got[2 * f] ? 0 : got[2 * f + 1];
    return hipSuccess;
}

// Device buffers kept between calls of the per-call entry points (grown as
// needed, one set per device; the calls are serialised).
struct Cache {
    char* p = nullptr;
    size_t n = 0;
};
std::mutex cache_lock;
Cache cache_ws[64], cache_bits[64], cache_res[64], cache_host[64];
std::mutex host_lock;  // csg_execute: its operand uploads

hipError_t cached(Cache* table, size_t need, char** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    Cache& c = table[dev];
    if (c.n < need) {
        if (c.p && (e = hipFree(c.p)) != hipSuccess) return e;
        c.p = nullptr;
        c.n = 0;
        if ((e = hipMalloc(&c.p, need)) != hipSuccess) return e;
        c.n = need;
    }
    *out = c.p;
    return hipSuccess;
}

bool operand_ok(const void* nodes, int64_t count, const void* payload, int64_t samples)
{
    return nodes && payload && count >= 1 && count <= INT32_MAX && samples >= 1;
}

std::vector<FieldRef> field_refs(const void* const* nodes, const int64_t* counts, const void* const* payloads,
                                 const int64_t* lengths, int32_t fields)
{
    std::vector<FieldRef> refs(fields);
    for (int f = 0; f < fields; ++f)
        refs[f] = {static_cast<const Node52*>(nodes[f]), static_cast<const uint32_t*>(payloads[f]), counts[f],
                   lengths[f]};
    return refs;
}

// fe_csg_device with or without the reference's size caps.
int csg_device(const void* const* nodes, const int64_t* counts, const void* const* payloads, const int64_t* lengths,
               int32_t fields, const int32_t* shape, const void* program, int32_t n, float* output, bool capped,
               hipStream_t st)
{
    if (!nodes || !counts || !payloads || !lengths || !shape || !program || !output) return 1;
    std::vector<int> last;
    const Instr24* prog = static_cast<const Instr24*>(program);
    const int v = validate_program(fields, shape, prog, n, capped, last);
    if (v) return v;
    std::vector<char> used(fields, 0);
    for (int f = 0; f < fields; ++f) {
        used[f] = last[f] >= 0;
        if (used[f] && (!operand_ok(nodes[f], counts[f], payloads[f], lengths[f]) ||
                        (capped && (counts[f] > 33554432 || lengths[f] > 16777216))))
            return 3;  // the reference gather's code 1 fails the first tile
    }
    const std::vector<FieldRef> refs = field_refs(nodes, counts, payloads, lengths, fields);
    std::lock_guard<std::mutex> hold(cache_lock);
    char *bits = nullptr, *res = nullptr, *ws = nullptr;
    FE_TRY(cached(cache_bits, check_bits_bytes((int64_t)shape[0] * shape[1] * shape[2]), &bits));
    FE_TRY(cached(cache_res, sizeof(int) * 2 * fields, &res));
    std::vector<int> bad;
    FE_TRY(check_operands(refs, used, shape, reinterpret_cast<uint32_t*>(bits), reinterpret_cast<int*>(res), bad, st));
    int first_bad = INT_MAX;
    for (int f = 0; f < fields; ++f) first_bad = std::min(first_bad, bad[f]);
    if (first_bad == 0) return 3;  // tile 0's reads already fail (or a table is malformed)
    Plan P;
    FE_TRY(make_plan(fields, prog, n, last, shape, P));
    FE_TRY(cached(cache_ws, P.ws_total, &ws));
    return run_plan(P, refs, shape, first_bad, output, ws, st);
}

}  // namespace

// Validates one packed operand for programs over a grid of `shape` (once per
// operand; fe_csg_run then reads it without checking): *first_bad = the first
// reference tile, x-major over (16,32,16) tiles, where the reference's
// gather of this operand fails (0 for a malformed table), INT32_MAX when
// every grid sample is covered by exactly one leaf.  Allocates its bitmaps.
extern "C" int fe_csg_check(const void* nodes, int64_t count, const void* payload, int64_t samples,
                            const int32_t* shape, int32_t* first_bad, void* stream) noexcept
{
    try {
        if (!shape || !first_bad || !operand_ok(nodes, count, payload, samples)) return 1;
        for (int a = 0; a < 3; ++a)
            if (shape[a] < 1 || shape[a] > 4096) return 1;
        hipStream_t st = static_cast<hipStream_t>(stream);
        const std::vector<FieldRef> refs = {{static_cast<const Node52*>(nodes), static_cast<const uint32_t*>(payload),
                                             count, samples}};
        std::lock_guard<std::mutex> hold(cache_lock);
        char *bits = nullptr, *res = nullptr;
        FE_TRY(cached(cache_bits, check_bits_bytes((int64_t)shape[0] * shape[1] * shape[2]), &bits));
        FE_TRY(cached(cache_res, sizeof(int) * 2, &res));
        std::vector<int> bad;
        FE_TRY(check_operands(refs, std::vector<char>(1, 1), shape, reinterpret_cast<uint32_t*>(bits),
                              reinterpret_cast<int*>(res), bad, st));
        *first_bad = bad[0];
        return 0;
    } catch (...) {
        return 5;
    }
}

// Device bytes of the workspace fe_csg_run needs for a program over `shape`
// on the current device (-1: an invalid program, as fe_csg_run would refuse).
extern "C" int64_t fe_csg_workspace_bytes(int32_t fields, const int32_t* shape, const void* program,
                                          int32_t n) noexcept
{
    try {
        std::vector<int> last;
        const Instr24* prog = static_cast<const Instr24*>(program);
        if (validate_program(fields, shape, prog, n, false, last)) return -1;
        Plan P;
        if (make_plan(fields, prog, n, last, shape, P) != hipSuccess) return -1;
        return (int64_t)P.ws_total;
    } catch (...) {
        return -1;
    }
}

// The program over operands checked by fe_csg_check (first_bad: their
// results, host int32 per field; unused fields are ignored), with a caller
// workspace of fe_csg_workspace_bytes: no allocation, no operand scan, one
// host synchronisation for the status.  Axes <= 4096, no sample cap.
extern "C" int fe_csg_run(const void* const* nodes, const int64_t* counts, const void* const* payloads,
                          const int64_t* lengths, int32_t fields, const int32_t* shape, const void* program,
                          int32_t n, const int32_t* first_bad, float* output, void* workspace, int64_t workspace_bytes,
                          void* stream) noexcept
{
    try {
        if (!nodes || !counts || !payloads || !lengths || !shape || !program || !first_bad || !output || !workspace)
            return 1;
        std::vector<int> last;
        const Instr24* prog = static_cast<const Instr24*>(program);
        const int v = validate_program(fields, shape, prog, n, false, last);
        if (v) return v;
        int bad = INT_MAX;
        for (int f = 0; f < fields; ++f) {
            if (last[f] < 0) continue;
            if (!operand_ok(nodes[f], counts[f], payloads[f], lengths[f])) return 3;
            bad = std::min(bad, first_bad[f] < 0 ? 0 : first_bad[f]);
        }
        if (bad == 0) return 3;
        Plan P;
        FE_TRY(make_plan(fields, prog, n, last, shape, P));
        if ((int64_t)P.ws_total > workspace_bytes) return 1;
        return run_plan(P, field_refs(nodes, counts, payloads, lengths, fields), shape, bad, output,
                        static_cast<char*>(workspace), static_cast<hipStream_t>(stream));
    } catch (...) {
        return 5;
    }
}

// Device CSG with a check of every used operand per call: fields[f] =
// {device nodes, count, device payload, samples}; output: device float32
// nx*ny*nz.  Status as FE/csg_batch_native_v1/execute.cpp:10-54; axes <= 4096,
// no sample cap.  Buffers are kept between calls.
extern "C" int fe_csg_device(const void* const* nodes, const int64_t* counts, const void* const* payloads,
                             const int64_t* lengths, int32_t fields, const int32_t* shape, const void* program,
                             int32_t n, float* output, void* stream) noexcept
{
    try {
        return csg_device(nodes, counts, payloads, lengths, fields, shape, program, n, output, false,
                          static_cast<hipStream_t>(stream));
    } catch (...) {
        return 5;
    }
}

// FE/csg_batch_native_v1/execute.cpp:10 (csg_execute) over host arrays, with
// the reference's caps: upload the used operands, run, download the output.
extern "C" int csg_execute(const void* const* nodes, const int64_t* counts, const void* const* payload,
                           const int64_t* lengths, int32_t fields, const int32_t* shape, const void* program,
                           int32_t instructions, float* output) noexcept
{
    try {
        if (!nodes || !counts || !payload || !lengths || !shape || !program || !output) return 1;
        std::vector<int> last;
        int v = validate_program(fields, shape, static_cast<const Instr24*>(program), instructions, true, last);
        if (v) return v;
        const int64_t N = (int64_t)shape[0] * shape[1] * shape[2];
        std::vector<size_t> noff(fields, 0), poff(fields, 0);
        size_t total = align256((size_t)N * 4);
        for (int f = 0; f < fields; ++f) {
            if (last[f] < 0) continue;
            if (!operand_ok(nodes[f], counts[f], payload[f], lengths[f]) || counts[f] > 33554432 ||
                lengths[f] > 16777216)
                return 3;
            noff[f] = total;
            total += align256((size_t)counts[f] * 52);
            poff[f] = total;
            total += align256((size_t)lengths[f] * 4);
        }
        std::lock_guard<std::mutex> hold_host(host_lock);
        char* base = nullptr;
        {
            std::lock_guard<std::mutex> hold(cache_lock);
            FE_TRY(cached(cache_host, total, &base));
        }
        std::vector<const void*> dn(fields, nullptr), dp(fields, nullptr);
        for (int f = 0; f < fields; ++f) {
            if (last[f] < 0) continue;
            dn[f] = base + noff[f];
            dp[f] = base + poff[f];
            FE_TRY(hipMemcpy(base + noff[f], nodes[f], (size_t)counts[f] * 52, hipMemcpyHostToDevice));
            FE_TRY(hipMemcpy(base + poff[f], payload[f], (size_t)lengths[f] * 4, hipMemcpyHostToDevice));
        }
        float* dout = reinterpret_cast<float*>(base);
        const int status = csg_device(dn.data(), counts, dp.data(), lengths, fields, shape, program, instructions,
                                      dout, true, nullptr);
        if (status == 0) FE_TRY(hipMemcpy(output, dout, (size_t)N * 4, hipMemcpyDeviceToHost));
        return status;
    } catch (...) {
        return 5;
    }
}

// ---------------------------------------------------------------- tile reader
namespace {
// Per node, in record order: status 3 for a malformed leaf; for a leaf
// overlapping the tile, every covered target records the first covering node.
// A workgroup per node record (grid-stride), its threads over the overlap.
__global__ __launch_bounds__(256) void k_gather_first(const Node52* __restrict__ nodes, int64_t count, int64_t samples,
                                                      const int* __restrict__ lower, const int* __restrict__ upper,
                                                      int* __restrict__ first, int* __restrict__ mal)
{
    for (int64_t i = blockIdx.x; i < count; i += gridDim.x) {
        const Node52 nd = nodes[i];
        if (nd.children[0] >= 0) continue;
        int lo[3], hi[3];
        bool ov = true, bad = false;
        int64_t size = 1;
        for (int a = 0; a < 3; ++a) {
            if (nd.lo[a] < 0 || nd.shape[a] < 1 || nd.shape[a] > 4096 || nd.lo[a] > 4096 - nd.shape[a] ||
                (nd.stored[a] != 1 && nd.stored[a] != nd.shape[a]))
                bad = true;
            size *= nd.stored[a];
            lo[a] = max(lower[a], nd.lo[a]);
            hi[a] = min(upper[a], nd.lo[a] + nd.shape[a]);
            ov = ov && hi[a] > lo[a];
        }
        const int64_t off = noff(nd);
        if (bad || off < 0 || size > samples || off > samples - size) {
            if (threadIdx.x == 0) atomicMin(mal, (int)i);
            continue;
        }
        if (!ov) continue;
        const int dy = upper[1] - lower[1], dz = upper[2] - lower[2];
        const int ey = hi[1] - lo[1], ez = hi[2] - lo[2];
        const int n = (hi[0] - lo[0]) * ey * ez;
        for (int t = threadIdx.x; t < n; t += 256) {
            const int x = lo[0] + t / (ey * ez), y = lo[1] + (t / ez) % ey, z = lo[2] + t % ez;
            atomicMin(&first[((x - lower[0]) * dy + (y - lower[1])) * dz + (z - lower[2])], (int)i);
        }
    }
}

// Second visit: a covered target whose first node is another one marks an
// overlap at this node; the first node copies its bits.  A workgroup per node
// record, as above.
__global__ __launch_bounds__(256) void k_gather_copy(const Node52* __restrict__ nodes, int64_t count,
                                                     const uint32_t* __restrict__ payload, const int* __restrict__ lower,
                                                     const int* __restrict__ upper, const int* __restrict__ first,
                                                     int* __restrict__ ovl, uint32_t* __restrict__ out,
                                                     const int* __restrict__ mal)
{
    const int64_t stop = min(count, (int64_t)*mal);  // records after a malformed one are never read
    for (int64_t i = blockIdx.x; i < stop; i += gridDim.x) {
        const Node52 nd = nodes[i];
        if (nd.children[0] >= 0) continue;
        int lo[3], hi[3];
        bool ov = true;
        for (int a = 0; a < 3; ++a) {
            lo[a] = max(lower[a], nd.lo[a]);
            hi[a] = min(upper[a], nd.lo[a] + nd.shape[a]);
            ov = ov && hi[a] > lo[a];
        }
        if (!ov) continue;
        const int dy = upper[1] - lower[1], dz = upper[2] - lower[2];
        const int ey = hi[1] - lo[1], ez = hi[2] - lo[2];
        const int n = (hi[0] - lo[0]) * ey * ez;
        const int64_t off = noff(nd);
        for (int t = threadIdx.x; t < n; t += 256) {
            const int x = lo[0] + t / (ey * ez), y = lo[1] + (t / ez) % ey, z = lo[2] + t % ez;
            const int g = ((x - lower[0]) * dy + (y - lower[1])) * dz + (z - lower[2]);
            if (first[g] != (int)i) {
                atomicMin(ovl, (int)i);
            } else {
                int64_t s2 = off;
                s2 += (int64_t)(nd.stored[0] == 1 ? 0 : x - nd.lo[0]) * nd.stored[1] * nd.stored[2];
                s2 += (int64_t)(nd.stored[1] == 1 ? 0 : y - nd.lo[1]) * nd.stored[2];
                s2 += nd.stored[2] == 1 ? 0 : z - nd.lo[2];
                out[g] = payload[s2];
            }
        }
    }
}
}  // namespace

// FE/csg_tile_native_v1/gather.cpp:11-50 (csg_gather): exact bits of every leaf
// intersecting [lower, upper) copied into the tile, with its status codes
// (1 arguments, 2 bounds, 3 malformed node, 4 overlap or gap, 5 exception);
// host arrays in and out, the scan itself on the device.
extern "C" int csg_gather(const void* raw, int64_t count, const void* payload, int64_t samples, const int32_t* lower,
                          const int32_t* upper, void* output) noexcept
{
    try {
        if (!raw || !payload || !lower || !upper || !output || count < 1 || count > 33554432 || samples < 1 ||
            samples > 16777216)
            return 1;
        int64_t total = 1;
        for (int a = 0; a < 3; ++a) {
            if (lower[a] < 0 || upper[a] <= lower[a] || upper[a] > 4096) return 2;
            total *= upper[a] - lower[a];
        }
        if (total > 8192) return 2;
        Node52* dn = nullptr;
        uint32_t* dp = nullptr;
        int* di = nullptr;
        uint32_t* dout = nullptr;
        int status = 5;
        do {
            if (hipMalloc(&dn, (size_t)count * 52) != hipSuccess) break;
            if (hipMalloc(&dp, (size_t)samples * 4) != hipSuccess) break;
            if (hipMalloc(&di, (size_t)(8 + total) * 4) != hipSuccess) break;
            if (hipMalloc(&dout, (size_t)total * 4) != hipSuccess) break;
            std::vector<int> init(8 + total, INT_MAX);
            for (int a = 0; a < 3; ++a) {
                init[a] = lower[a];
                init[3 + a] = upper[a];
            }
            if (hipMemcpy(dn, raw, (size_t)count * 52, hipMemcpyHostToDevice) != hipSuccess) break;
            if (hipMemcpy(dp, payload, (size_t)samples * 4, hipMemcpyHostToDevice) != hipSuccess) break;
            if (hipMemcpy(di, init.data(), init.size() * 4, hipMemcpyHostToDevice) != hipSuccess) break;
            const unsigned g = (unsigned)std::min<int64_t>(count, 65536);  // workgroups over the node records
            hipLaunchKernelGGL(k_gather_first, dim3(g), dim3(256), 0, nullptr, dn, count, samples, di, di + 3, di + 8, di + 6);
            hipLaunchKernelGGL(k_gather_copy, dim3(g), dim3(256), 0, nullptr, dn, count, dp, di, di + 3, di + 8, di + 7,
                               dout, di + 6);
            std::vector<int> res(8 + total);
            if (hipMemcpy(res.data(), di, res.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) break;
            const int mal = res[6], ovl = res[7];
            // the reference stops at the first bad record in order: a malformed
            // leaf or the first leaf that writes an already covered target
            if (mal != INT_MAX && (ovl == INT_MAX || mal <= ovl)) { status = 3; break; }
            if (ovl != INT_MAX) { status = 4; break; }
            bool gap = false;
            for (int64_t t = 0; t < total; ++t) gap = gap || res[8 + t] == INT_MAX;
            if (gap) { status = 4; break; }
            if (hipMemcpy(output, dout, (size_t)total * 4, hipMemcpyDeviceToHost) != hipSuccess) break;
            status = 0;
        } while (0);
        if (dn) (void)hipFree(dn);
        if (dp) (void)hipFree(dp);
        if (di) (void)hipFree(di);
        if (dout) (void)hipFree(dout);
        return status;
    } catch (...) {
        return 5;
    }
}
