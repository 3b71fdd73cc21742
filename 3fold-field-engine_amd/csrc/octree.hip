// 2-bit adaptive block tree (FE/field_adaptive_blocks_v1.py:20-138) on gfx950.
//
// The tree stores one 2-bit code per node in breadth-first order: 0 exterior,
// 1 interior, 2 branch, 3 value leaf.  The children of the b-th branch start
// at root_count + 8*b, and value leaves hold their samples in node order, so
// a node's rank among the branches (or value leaves) before it replaces child
// pointers.  Ranks come from checkpoints every 64 packed bytes (256 nodes).
//
// Build: the dense-source kinds at block sides 8, 4, 2 and 1 (computed by
// fe_dense_classify) decide every node.  Breadth-first order is level order,
// so the build runs one level at a time: codes of the level's nodes, an
// exclusive scan of the branch (or value) flags, then the next level's nodes
// (8 per branch, octant order i,j,k) or the value samples.  Packing and the
// checkpoints are two more passes.  Lookups descend at most four levels; a
// rank is the checkpoint plus a popcount over at most 16 words.
#include "fe_common.hpp"
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

namespace {

constexpr int SCAN_ITEMS = 1024;  // per block of the device scan (256 threads x 4)

// Exclusive scan of n int32 flags into out, block totals into sums.
__global__ __launch_bounds__(256) void k_scan_blocks(const int32_t* __restrict__ in, int64_t n, int32_t* __restrict__ out,
                                                     int32_t* __restrict__ sums)
{
    __shared__ int32_t wsum[4];
    const int64_t base = (int64_t)blockIdx.x * SCAN_ITEMS + threadIdx.x * 4;
    int32_t v[4], t = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        v[i] = base + i < n ? in[base + i] : 0;
        t += v[i];
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int32_t incl = t;  // inclusive wave scan of the thread totals
    for (int d = 1; d < 64; d <<= 1) {
        const int32_t u = __shfl_up(incl, d, 64);
        if (lane >= d) incl += u;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int32_t off = 0;
    for (int i = 0; i < w; ++i) off += wsum[i];
    int32_t run = off + incl - t;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (threadIdx.x == 255) sums[blockIdx.x] = off + incl;
}

// Exclusive scan of the block totals in place (one block, chunks of 256); the
// grand total goes to sums[nblocks].
__global__ __launch_bounds__(256) void k_scan_sums(int32_t* __restrict__ sums, int nblocks)
{
    __shared__ int32_t wsum[4];
    __shared__ int32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int c0 = 0; c0 < nblocks; c0 += 256) {
        const int i = c0 + threadIdx.x;
        const int32_t t = i < nblocks ? sums[i] : 0;
        int32_t incl = t;
        for (int d = 1; d < 64; d <<= 1) {
            const int32_t u = __shfl_up(incl, d, 64);
            if (lane >= d) incl += u;
        }
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        int32_t off = carry;
        for (int k = 0; k < w; ++k) off += wsum[k];
        if (i < nblocks) sums[i] = off + incl - t;
        __syncthreads();
        if (threadIdx.x == 255) carry = off + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[nblocks] = carry;
}

__global__ __launch_bounds__(256) void k_scan_add(int32_t* __restrict__ out, int64_t n, const int32_t* __restrict__ sums)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] += sums[i / SCAN_ITEMS];
}

// Codes of one level.  Level 0 nodes are the root blocks in C order (implicit);
// deeper levels read their (x,y,z) from xyz.  flag: branch (levels 0-2) or
// value leaf (level 3).  stats[level*4 + code]: code counts; first[level*4 +
// code]: first node index with that code in the level.
__global__ __launch_bounds__(256) void k_oct_level(const int32_t* __restrict__ xyz, int64_t n, int level, int side,
                                                   int rx, int ry, int rz, int nx, int ny, int nz,
                                                   const int32_t* __restrict__ kind, int nby, int nbz,
                                                   uint8_t* __restrict__ codes, int32_t* __restrict__ flag,
                                                   unsigned long long* __restrict__ stats,
                                                   unsigned long long* __restrict__ first)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    int x, y, z;
    if (level == 0) {
        z = 8 * (int)(t % rz);
        y = 8 * (int)((t / rz) % ry);
        x = 8 * (int)(t / ((int64_t)rz * ry));
    } else {
        x = xyz[3 * t];
        y = xyz[3 * t + 1];
        z = xyz[3 * t + 2];
    }
    int code;
    if (x >= nx || y >= ny || z >= nz) code = 0;
    else {
        const int32_t k = kind[((int64_t)(x / side) * nby + y / side) * nbz + z / side];
        code = k ? (k > 0 ? 0 : 1) : (side == 1 ? 3 : 2);
    }
    codes[t] = (uint8_t)code;
    flag[t] = code == (level == 3 ? 3 : 2);
    // one count and one first index per code and wave (every thread's own
    // atomics on these few words serialised the pass): the lowest lane with
    // a code holds the wave's first node with it (t grows with the lane)
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const uint64_t m = __ballot(code == c);
        if (m && lane == __builtin_ctzll(m)) {
            atomicAdd(&stats[level * 4 + c], (unsigned long long)__popcll(m));
            atomicMin(&first[level * 4 + c], (unsigned long long)t);
        }
    }
}

// Children of the branches of a level: branch b (exclusive rank pos[t]) writes
// its eight children at 8*pos .. 8*pos+7.
__global__ __launch_bounds__(256) void k_oct_children(const int32_t* __restrict__ xyz, int64_t n, int level, int rx,
                                                      int ry, int rz, int half, const uint8_t* __restrict__ codes,
                                                      const int32_t* __restrict__ pos, int32_t* __restrict__ next)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n || codes[t] != 2) return;
    int x, y, z;
    if (level == 0) {
        z = 8 * (int)(t % rz);
        y = 8 * (int)((t / rz) % ry);
        x = 8 * (int)(t / ((int64_t)rz * ry));
    } else {
        x = xyz[3 * t];
        y = xyz[3 * t + 1];
        z = xyz[3 * t + 2];
    }
    int32_t* o = next + (int64_t)24 * pos[t];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        o[3 * c] = x + ((c >> 2) & 1) * half;
        o[3 * c + 1] = y + ((c >> 1) & 1) * half;
        o[3 * c + 2] = z + (c & 1) * half;
    }
}

// Samples of the value leaves of the last level, in node order.
__global__ __launch_bounds__(256) void k_oct_values(const int32_t* __restrict__ xyz, int64_t n,
                                                    const uint8_t* __restrict__ codes, const int32_t* __restrict__ pos,
                                                    const float* __restrict__ dense, int ny, int nz,
                                                    float* __restrict__ values)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n || codes[t] != 3) return;
    values[pos[t]] = dense[((int64_t)xyz[3 * t] * ny + xyz[3 * t + 1]) * nz + xyz[3 * t + 2]];
}

// Four codes per byte (code i at bits 2i); bytes past the end stay zero.
__global__ __launch_bounds__(256) void k_oct_pack(const uint8_t* __restrict__ codes, int64_t n, int64_t nbytes_padded,
                                                  uint8_t* __restrict__ packed)
{
    const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (b >= nbytes_padded) return;
    uint32_t v = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (4 * b + i < n) v |= (uint32_t)codes[4 * b + i] << (2 * i);
    packed[b] = (uint8_t)v;
}

// Number of 2-bit fields equal to `code` among the first `fields` of word w.
__device__ __forceinline__ int count_code(uint32_t w, int code, int fields)
{
    const uint32_t rep = code == 2 ? 0xAAAAAAAAu : 0xFFFFFFFFu;
    const uint32_t x = w ^ rep;                          // matching fields become 00
    const uint32_t nz = (x | (x >> 1)) & 0x55555555u;    // 1 in each non-matching field
    const uint32_t keep = fields >= 16 ? 0x55555555u : (0x55555555u & ((1u << (2 * fields)) - 1u));
    return fields - __popc(nz & keep);
}

// Per group of 64 bytes: the counts of branch and value codes (into sums2 and
// sums3; an exclusive scan of them gives the checkpoints).
__global__ __launch_bounds__(256) void k_oct_group_counts(const uint32_t* __restrict__ words, int64_t groups,
                                                          int32_t* __restrict__ c2, int32_t* __restrict__ c3)
{
    const int64_t g = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (g >= groups) return;
    int a = 0, b = 0;
    if (lane < 16) {
        const uint32_t w = words[g * 16 + lane];
        a = count_code(w, 2, 16);
        b = count_code(w, 3, 16);
    }
    for (int d = 8; d > 0; d >>= 1) {
        a += __shfl_down(a, d, 64);
        b += __shfl_down(b, d, 64);
    }
    if (lane == 0) {
        c2[g] = a;
        c3[g] = b;
    }
}

// Codes equal to `code` before `node` (checkpoint + popcounts over the group).
__device__ __forceinline__ int64_t oct_rank(const uint32_t* __restrict__ words, const int32_t* __restrict__ prefix,
                                            int64_t node, int code)
{
    const int64_t group = node >> 8;
    int64_t rank = prefix[group];
    const int64_t first = group << 8;  // first node of the group
    const int before = (int)(node - first);
    for (int i = 0; i * 16 < before; ++i) rank += count_code(words[(group << 4) + i], code, min(16, before - i * 16));
    return rank;
}

// FE/field_adaptive_blocks_v1.py:52-82 per query; status 1 (kept in *bad) if a
// branch would go below side 1.
__global__ __launch_bounds__(256) void k_oct_sample(const uint32_t* __restrict__ words,
                                                    const int32_t* __restrict__ bpre, const int32_t* __restrict__ vpre,
                                                    const float* __restrict__ values, int ry, int rz,
                                                    int64_t root_count, double pitch, const int64_t* __restrict__ ijk,
                                                    int64_t n, float* __restrict__ out, int* __restrict__ bad)
{
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    const int64_t i = ijk[3 * q], j = ijk[3 * q + 1], k = ijk[3 * q + 2];
    int64_t node = ((i >> 3) * ry + (j >> 3)) * rz + (k >> 3), size = 8;
    for (int depth = 0; depth < 5; ++depth) {
        const int code = (words[node >> 4] >> (2 * (node & 15))) & 3;
        if (code < 2) {
            out[q] = (float)(((code == 0 ? 1.0 : -1.0) * 1.7320508075688772) * (double)size * pitch);
            return;
        }
        if (code == 3) {
            out[q] = values[oct_rank(words, vpre, node, 3)];
            return;
        }
        const int64_t half = size >> 1;
        if (half < 1) break;
        const int64_t child = 4 * ((i / half) & 1) + 2 * ((j / half) & 1) + ((k / half) & 1);
        node = root_count + 8 * oct_rank(words, bpre, node, 2) + child;
        size = half;
    }
    out[q] = 0.0f;
    *bad = 1;
}

struct Octree {
    int shape[3] = {0, 0, 0}, root[3] = {0, 0, 0};
    int64_t nodes = 0, nvalues = 0, nbytes = 0, groups = 0;
    uint8_t* packed = nullptr;    // groups*64 bytes (zero padded)
    int32_t* bpre = nullptr;      // groups + 1
    int32_t* vpre = nullptr;      // groups + 1
    float* values = nullptr;      // max(nvalues, 1)
    int64_t stats[16] = {}, first[16] = {};
};

void oct_free(Octree* t)
{
    if (!t) return;
    (void)hipFree(t->packed);
    (void)hipFree(t->bpre);
    (void)hipFree(t->vpre);
    (void)hipFree(t->values);
    delete t;
}

// Exclusive scan of flags[0..n) into pos; returns the total in *total (host).
hipError_t scan(const int32_t* flags, int64_t n, int32_t* pos, int32_t* sums, int32_t* total, hipStream_t st)
{
    const int nb = (int)((n + SCAN_ITEMS - 1) / SCAN_ITEMS);
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(256), 0, st, flags, n, pos, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, st, sums, nb);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, pos, n, sums);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if ((e = hipMemcpyAsync(total, sums + nb, sizeof(int32_t), hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    return hipStreamSynchronize(st);
}

template <class T>
hipError_t dalloc(T** p, int64_t count)
{
    return hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (size_t)(count > 0 ? count : 1));
}

// Temporary device buffers of fe_octree_build, kept between builds (grown
// as needed, one set per device; builds are serialised by oct_lock), so a
// build allocates only the arrays its tree keeps.
enum { S_CODES0, S_CODES1, S_STATS, S_FLAG, S_POS, S_SUMS, S_XYZ0, S_XYZ1, S_C2, S_C3, S_COUNT };
std::mutex oct_lock;
void* oct_buf[64][S_COUNT];
size_t oct_len[64][S_COUNT];

template <class T>
hipError_t scratch(int slot, T** p, int64_t count)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    const size_t need = sizeof(T) * (size_t)(count > 0 ? count : 1);
    if (oct_len[dev][slot] < need) {
        if (oct_buf[dev][slot] && (e = hipFree(oct_buf[dev][slot])) != hipSuccess) return e;
        oct_buf[dev][slot] = nullptr;
        oct_len[dev][slot] = 0;
        const size_t grown = need + need / 4;
        if ((e = hipMalloc(&oct_buf[dev][slot], grown)) != hipSuccess) return e;
        oct_len[dev][slot] = grown;
    }
    *p = static_cast<T*>(oct_buf[dev][slot]);
    return hipSuccess;
}

}  // namespace

// Builds the tree of a device dense field (nx,ny,nz) from its dense-source
// kinds at block sides 8, 4, 2, 1 (kinds[l]: device int32 of that side's
// blocks, C order).  Returns an opaque handle (fe_octree_* accessors), NULL
// with *status 1 (invalid) or 2 (runtime).
extern "C" void* fe_octree_build(const float* dense, int nx, int ny, int nz, const int32_t* const* kinds,
                                 int* status, void* stream) noexcept
{
    int dummy = 0;
    int& st_out = status ? *status : dummy;
    st_out = FE_INVALID;
    if (!dense || !kinds || nx < 1 || ny < 1 || nz < 1 || nx > 4096 || ny > 4096 || nz > 4096) return nullptr;
    for (int l = 0; l < 4; ++l)
        if (!kinds[l]) return nullptr;
    st_out = FE_RUNTIME;
    hipStream_t st = static_cast<hipStream_t>(stream);
    Octree* t = nullptr;
    uint8_t* codes = nullptr;
    int32_t *flag = nullptr, *pos = nullptr, *sums = nullptr, *xyz = nullptr, *next = nullptr;
    unsigned long long* stats = nullptr;
    hipError_t e = hipSuccess;
    std::lock_guard<std::mutex> hold(oct_lock);
    int cs = S_CODES0, xs = S_XYZ0;  // the slots holding codes and xyz now
    try {
        t = new Octree();
        t->shape[0] = nx; t->shape[1] = ny; t->shape[2] = nz;
        const int rx = (nx + 7) / 8, ry = (ny + 7) / 8, rz = (nz + 7) / 8;
        t->root[0] = rx; t->root[1] = ry; t->root[2] = rz;
        const int64_t roots = (int64_t)rx * ry * rz;
        // level sizes are known one level ahead; node codes of all levels go
        // into one array grown as needed
        int64_t cap_codes = roots * 2 + 1024, off = 0;
        if ((e = scratch(cs, &codes, cap_codes)) != hipSuccess) throw e;
        if ((e = scratch(S_STATS, &stats, 32)) != hipSuccess) throw e;
        std::vector<unsigned long long> init(32, 0ull);
        for (int i = 16; i < 32; ++i) init[i] = ~0ull;
        if ((e = hipMemcpyAsync(stats, init.data(), 32 * sizeof(unsigned long long), hipMemcpyHostToDevice, st)) !=
            hipSuccess)
            throw e;
        int64_t n = roots;
        for (int level = 0; level < 4; ++level) {
            const int side = 8 >> level;
            if (n == 0) break;  // no branches above: the tree ends here
            const int nby = (ny + side - 1) / side, nbz = (nz + side - 1) / side;
            if (off + n > cap_codes) {  // grow the code array
                uint8_t* bigger = nullptr;
                const int64_t cap2 = (off + n) * 2;
                const int other = cs == S_CODES0 ? S_CODES1 : S_CODES0;
                if ((e = scratch(other, &bigger, cap2)) != hipSuccess) throw e;
                if ((e = hipMemcpyAsync(bigger, codes, (size_t)off, hipMemcpyDeviceToDevice, st)) != hipSuccess) throw e;
                codes = bigger;
                cs = other;
                cap_codes = cap2;
            }
            if ((e = scratch(S_FLAG, &flag, n)) != hipSuccess) throw e;
            if ((e = scratch(S_POS, &pos, n)) != hipSuccess) throw e;
            if ((e = scratch(S_SUMS, &sums, (n + SCAN_ITEMS - 1) / SCAN_ITEMS + 1)) != hipSuccess) throw e;
            const unsigned grid = (unsigned)((n + 255) / 256);
            hipLaunchKernelGGL(k_oct_level, dim3(grid), dim3(256), 0, st, xyz, n, level, side, rx, ry, rz, nx, ny, nz,
                               kinds[level], nby, nbz, codes + off, flag, stats, stats + 16);
            if ((e = hipGetLastError()) != hipSuccess) throw e;
            int32_t total = 0;
            if ((e = scan(flag, n, pos, sums, &total, st)) != hipSuccess) throw e;
            if (level < 3) {
                if ((e = scratch(xs == S_XYZ0 ? S_XYZ1 : S_XYZ0, &next, (int64_t)total * 24)) != hipSuccess) throw e;
                hipLaunchKernelGGL(k_oct_children, dim3(grid), dim3(256), 0, st, xyz, n, level, rx, ry, rz, side / 2,
                                   codes + off, pos, next);
            } else {
                if ((e = dalloc(&t->values, total)) != hipSuccess) throw e;
                hipLaunchKernelGGL(k_oct_values, dim3(grid), dim3(256), 0, st, xyz, n, codes + off, pos, dense, ny, nz,
                                   t->values);
                t->nvalues = total;
            }
            if ((e = hipGetLastError()) != hipSuccess) throw e;
            if ((e = hipStreamSynchronize(st)) != hipSuccess) throw e;
            if (next) {
                xyz = next;
                xs = xs == S_XYZ0 ? S_XYZ1 : S_XYZ0;
                next = nullptr;
            }
            off += n;
            n = level < 3 ? (int64_t)total * 8 : 0;
        }
        if (!t->values && (e = dalloc(&t->values, 0)) != hipSuccess) throw e;
        t->nodes = off;
        t->nbytes = (off + 3) / 4;
        t->groups = t->nbytes / 64 + 1;  // checkpoints at bytes 0, 64, ..., <= nbytes
        if ((e = dalloc(&t->packed, t->groups * 64)) != hipSuccess) throw e;
        hipLaunchKernelGGL(k_oct_pack, dim3((unsigned)((t->groups * 64 + 255) / 256)), dim3(256), 0, st, codes, off,
                           t->groups * 64, t->packed);
        if ((e = dalloc(&t->bpre, t->groups + 1)) != hipSuccess) throw e;
        if ((e = dalloc(&t->vpre, t->groups + 1)) != hipSuccess) throw e;
        int32_t *c2 = nullptr, *c3 = nullptr;
        if ((e = scratch(S_C2, &c2, t->groups)) != hipSuccess) throw e;
        if ((e = scratch(S_C3, &c3, t->groups)) != hipSuccess) throw e;
        hipLaunchKernelGGL(k_oct_group_counts, dim3((unsigned)((t->groups + 3) / 4)), dim3(256), 0, st,
                           reinterpret_cast<const uint32_t*>(t->packed), t->groups, c2, c3);
        int32_t tot = 0;
        if ((e = scratch(S_SUMS, &sums, (t->groups + SCAN_ITEMS - 1) / SCAN_ITEMS + 1)) == hipSuccess &&
            (e = scan(c2, t->groups, t->bpre, sums, &tot, st)) == hipSuccess)
            e = scan(c3, t->groups, t->vpre, sums, &tot, st);
        if (e != hipSuccess) throw e;
        unsigned long long host[32];
        if ((e = hipMemcpyAsync(host, stats, sizeof host, hipMemcpyDeviceToHost, st)) != hipSuccess) throw e;
        if ((e = hipStreamSynchronize(st)) != hipSuccess) throw e;
        for (int i = 0; i < 16; ++i) {
            t->stats[i] = (int64_t)host[i];
            t->first[i] = host[16 + i] == ~0ull ? -1 : (int64_t)host[16 + i];
        }
        st_out = FE_OK;
        return t;
    } catch (hipError_t err) {
        fe_note_error(err, "fe_octree_build");
    } catch (...) {
    }
    // the temporaries stay cached; let queued work on them finish first
    (void)hipStreamSynchronize(st);
    oct_free(t);
    return nullptr;
}

// Sizes of a built tree: [0] node count, [1] value count, [2] packed bytes,
// [3] checkpoints (= packed bytes / 64 + 1); counts[16] per level (side 8, 4,
// 2, 1) and code; first[16] the first node of that level with that code
// (index within the level, -1 if none).
extern "C" int fe_octree_info(const void* handle, int64_t* sizes, int64_t* counts, int64_t* first) noexcept
{
    const auto* t = static_cast<const Octree*>(handle);
    if (!t || !sizes) return FE_INVALID;
    sizes[0] = t->nodes;
    sizes[1] = t->nvalues;
    sizes[2] = t->nbytes;
    sizes[3] = t->groups;
    if (counts) std::memcpy(counts, t->stats, sizeof t->stats);
    if (first) std::memcpy(first, t->first, sizeof t->first);
    return FE_OK;
}

// Device copies of the packed codes (nbytes), checkpoints (groups each) and
// values (nvalues); any destination may be NULL.
extern "C" int fe_octree_copy(const void* handle, uint8_t* packed, int32_t* branch_prefix, int32_t* value_prefix,
                              float* values, void* stream) noexcept
{
    const auto* t = static_cast<const Octree*>(handle);
    if (!t) return FE_INVALID;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (packed) FE_TRY(hipMemcpyAsync(packed, t->packed, (size_t)t->nbytes, hipMemcpyDeviceToDevice, st));
    if (branch_prefix)
        FE_TRY(hipMemcpyAsync(branch_prefix, t->bpre, sizeof(int32_t) * (size_t)t->groups, hipMemcpyDeviceToDevice, st));
    if (value_prefix)
        FE_TRY(hipMemcpyAsync(value_prefix, t->vpre, sizeof(int32_t) * (size_t)t->groups, hipMemcpyDeviceToDevice, st));
    if (values && t->nvalues)
        FE_TRY(hipMemcpyAsync(values, t->values, sizeof(float) * (size_t)t->nvalues, hipMemcpyDeviceToDevice, st));
    return FE_OK;
}

extern "C" void fe_octree_free(void* handle) noexcept { oct_free(static_cast<Octree*>(handle)); }

// Lookups of FE/field_adaptive_blocks_v1.py:52-82.  packed: device codes,
// zero-padded to a multiple of 64 bytes; prefixes: device checkpoints
// (packed bytes / 64 + 1 each); root_shape: host int32[3]; ijk: device int64 [n][3] inside the domain
// (checked by the caller).  Returns 1 if the tree is malformed (a branch below
// side 1), 0 otherwise; `flag` is a device int used for that report.
extern "C" int fe_octree_sample(const uint8_t* packed, const int32_t* branch_prefix, const int32_t* value_prefix,
                                const float* values, const int32_t* root_shape, int64_t root_count, double pitch,
                                const int64_t* ijk, int64_t n, float* out, int* flag, void* stream) noexcept
{
    if (!packed || !branch_prefix || !value_prefix || !values || !root_shape || !ijk || !out || !flag || n < 0 ||
        !std::isfinite(pitch))
        return FE_INVALID;
    if (n == 0) return FE_OK;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_oct_sample, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       reinterpret_cast<const uint32_t*>(packed), branch_prefix, value_prefix, values, root_shape[1],
                       root_shape[2], root_count, pitch, ijk, n, out, flag);
    FE_TRY(hipGetLastError());
    return FE_OK;
}
