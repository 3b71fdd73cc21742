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
// MI355X design: one workgroup per output tile of 4x8x64 samples.  The host
// compiles the program into micro-ops over LDS slots (linear-scan liveness,
// every input read once right before its first use).  An input read walks the
// field's box tree breadth-first with the whole workgroup to list the leaves
// overlapping the tile, then copies their intersections into the slot with
// coalesced reads along z.  Each instruction then runs over the slot arrays
// in LDS; only the final value is written to HBM.
#include "fe_common.hpp"
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
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
    int32_t kind;  // 0 read input a into slot dst, 1 union, 2 difference, 3 offset
    int32_t dst, a, b;
    double off;
};
static_assert(sizeof(MOp) == 24, "micro-op");

constexpr int TX = 4, TY = 8, TZ = 64, TS = TX * TY * TZ;  // 2048 samples per tile
constexpr int FRONT = 1024;                                  // BFS frontier / leaf list capacity

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

// Reads field f over the tile [tlo, thi) into slot (layout x,y,z with TY*TZ, TZ strides).
__device__ void read_tile(const FieldRef& F, const int* tlo, const int* thi, float* slot, int* cur, int* nxt, int* leaves,
                          int* counters)
{
    const int tid = threadIdx.x;
    if (tid == 0) {
        counters[0] = 1;  // frontier size
        counters[2] = 0;  // leaves
        counters[3] = 0;  // overflow
        cur[0] = 0;
    }
    __syncthreads();
    for (;;) {
        const int n = counters[0];
        if (n == 0 || counters[3]) break;
        if (tid == 0) counters[1] = 0;
        __syncthreads();
        for (int i = tid; i < n; i += 256) {
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
        for (int i = tid; i < counters[1] && i < FRONT; i += 256) cur[i] = nxt[i];
        if (tid == 0) counters[0] = counters[1];
        __syncthreads();
    }
    const int dx = thi[0] - tlo[0], dy = thi[1] - tlo[1], dz = thi[2] - tlo[2];
    if (counters[3]) {  // too many boxes for the lists: per-sample descent
        for (int j = tid; j < dx * dy * dz; j += 256) {
            const int z = j % dz, y = (j / dz) % dy, x = j / (dz * dy);
            slot[(x * TY + y) * TZ + z] = __uint_as_float(descend(F.nodes, F.payload, tlo[0] + x, tlo[1] + y, tlo[2] + z));
        }
    } else {
        const int nl = counters[2];
        for (int l = 0; l < nl; ++l) {
            const Node52& nd = F.nodes[leaves[l]];
            int lo[3], hi[3];
            for (int a = 0; a < 3; ++a) {
                lo[a] = max(tlo[a], nd.lo[a]);
                hi[a] = min(thi[a], nd.lo[a] + nd.shape[a]);
            }
            const int ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
            if (ex <= 0 || ey <= 0 || ez <= 0) continue;
            const int64_t off = noff(nd);
            for (int j = tid; j < ex * ey * ez; j += 256) {
                const int z = lo[2] + j % ez, y = lo[1] + (j / ez) % ey, x = lo[0] + j / (ez * ey);
                const int64_t sx = nd.stored[0] == 1 ? 0 : x - nd.lo[0], sy = nd.stored[1] == 1 ? 0 : y - nd.lo[1],
                              sz = nd.stored[2] == 1 ? 0 : z - nd.lo[2];
                slot[((x - tlo[0]) * TY + (y - tlo[1])) * TZ + (z - tlo[2])] =
                    __uint_as_float(F.payload[off + (sx * nd.stored[1] + sy) * nd.stored[2] + sz]);
            }
        }
    }
    __syncthreads();
}

// Reference tile index (x-major over (16,32,16) tiles) of sample (x,y,z).
__device__ __forceinline__ int ref_tile(int x, int y, int z, int nty, int ntz)
{
    return ((x >> 4) * nty + (y >> 5)) * ntz + (z >> 4);
}

__global__ __launch_bounds__(256) void k_csg_tiles(const FieldRef* __restrict__ fields, const MOp* __restrict__ prog,
                                                   int nops, int final_slot, int nx, int ny, int nz, int gty, int gtz,
                                                   int ntiles, float* __restrict__ out, int* __restrict__ err_nf)
{
    extern __shared__ float slots[];
    __shared__ int cur[FRONT], nxt[FRONT], leaves[FRONT];
    __shared__ int counters[4];
    __shared__ int bad_tile;
    const int t = blockIdx.x;
    if (t >= ntiles) return;
    const int tz = t % gtz, ty = (t / gtz) % gty, tx = t / (gtz * gty);
    const int tlo[3] = {tx * TX, ty * TY, tz * TZ};
    const int thi[3] = {min(tlo[0] + TX, nx), min(tlo[1] + TY, ny), min(tlo[2] + TZ, nz)};
    const int nty = (ny + 31) >> 5, ntz = (nz + 15) >> 4;
    const int tid = threadIdx.x;
    if (tid == 0) bad_tile = INT_MAX;
    __syncthreads();
    for (int o = 0; o < nops; ++o) {
        const MOp m = prog[o];
        float* dst = slots + m.dst * TS;
        if (m.kind == 0) {
            read_tile(fields[m.a], tlo, thi, dst, cur, nxt, leaves, counters);
            continue;
        }
        const float* A = slots + m.a * TS;
        const float* Bv = slots + m.b * TS;
        int bad = INT_MAX;
        for (int j = tid; j < TS; j += 256) {
            const int z = j % TZ, y = (j / TZ) % TY, x = j / (TZ * TY);
            const int gx = tlo[0] + x, gy = tlo[1] + y, gz = tlo[2] + z;
            if (gx >= thi[0] || gy >= thi[1] || gz >= thi[2]) continue;
            const float a = A[j];
            float r;
            if (m.kind == 3) r = (float)((double)a - m.off);
            else {
                const float b = m.kind == 1 ? Bv[j] : -Bv[j];
                r = m.kind == 1 ? (a < b ? a : b) : (a > b ? a : b);
            }
            if (!isfinite(r)) bad = min(bad, ref_tile(gx, gy, gz, nty, ntz));
            dst[j] = r;
        }
        if (bad != INT_MAX) atomicMin(&bad_tile, bad);
        __syncthreads();
    }
    const float* F = slots + final_slot * TS;
    for (int j = tid; j < TS; j += 256) {
        const int z = j % TZ, y = (j / TZ) % TY, x = j / (TZ * TY);
        const int gx = tlo[0] + x, gy = tlo[1] + y, gz = tlo[2] + z;
        if (gx < thi[0] && gy < thi[1] && gz < thi[2]) out[((int64_t)gx * ny + gy) * nz + gz] = F[j];
    }
    if (tid == 0 && bad_tile != INT_MAX) atomicMin(err_nf, bad_tile);
}

// Field checks with the reference's gather codes: malformed leaf records
// (FE/csg_tile_native_v1/gather.cpp:34-41) raise `malformed`; internal records
// must point forward to records inside the table (pre-order tables do), else
// `malformed`; every leaf adds one to each domain sample it covers.
__global__ __launch_bounds__(256) void k_field_check(FieldRef F, int nx, int ny, int nz, uint32_t* __restrict__ cover,
                                                     int* __restrict__ malformed)
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
    const int64_t cnt = (int64_t)hx * hy * hz;
    for (int64_t j = threadIdx.x; j < cnt; j += 256) {
        const int z = (int)(j % hz);
        const int64_t r = j / hz;
        const int y = (int)(r % hy), x = (int)(r / hy);
        const int64_t s = ((int64_t)(nd.lo[0] + x) * ny + nd.lo[1] + y) * nz + nd.lo[2] + z;
        if (cover[s] < 2) atomicAdd(&cover[s], 1u);  // 1 = exactly once, >= 2 = more than once
    }
}

// First reference tile containing a sample covered other than exactly once.
__global__ __launch_bounds__(256) void k_cover_scan(const uint32_t* __restrict__ cover, int nx, int ny, int nz,
                                                    int* __restrict__ err_gather)
{
    const int nty = (ny + 31) >> 5, ntz = (nz + 15) >> 4;
    const int64_t n = (int64_t)nx * ny * nz;
    int bad = INT_MAX;
    for (int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x; s < n; s += (int64_t)gridDim.x * 256) {
        if (cover[s] != 1) {
            const int z = (int)(s % nz);
            const int64_t r = s / nz;
            const int y = (int)(r % ny), x = (int)(r / ny);
            bad = min(bad, ref_tile(x, y, z, nty, ntz));
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

// Host validation of FE/csg_batch_native_v1/execute.cpp:14-24; returns 0, 1 or 2
// and the last-use table.
int validate_program(int32_t fields, const int32_t* shape, const Instr24* program, int32_t n, std::vector<int>& last)
{
    if (!shape || !program || fields < 1 || fields > 256 || n < 1 || n > 1024) return 1;
    int64_t total = 1;
    for (int a = 0; a < 3; ++a) {
        if (shape[a] < 1 || shape[a] > 4096) return 1;
        total *= shape[a];
    }
    if (total > 16777216) return 1;
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

// Linear-scan slot allocation: inputs are read right before their first use.
int compile_program(int32_t fields, const Instr24* program, int32_t n, const std::vector<int>& last, std::vector<MOp>& ops,
                    int& final_slot)
{
    std::vector<int> slot(fields + n, -1), free_slots;
    int nslots = 0;
    auto take = [&]() {
        if (!free_slots.empty()) {
            int s = free_slots.back();
            free_slots.pop_back();
            return s;
        }
        return nslots++;
    };
    for (int i = 0; i < n; ++i) {
        const Instr24& p = program[i];
        const int refs[2] = {p.left, p.op == 2 ? -1 : p.right};
        for (int r : refs) {
            if (r >= 0 && r < fields && slot[r] < 0) {
                slot[r] = take();
                ops.push_back({0, slot[r], r, 0, 0.0});
            }
        }
        // operands whose last use is this instruction free their slots after it
        const int d = take();
        ops.push_back({p.op == 0 ? 1 : (p.op == 1 ? 2 : 3), d, slot[p.left], p.op == 2 ? 0 : slot[p.right],
                       p.op == 2 ? p.offset : 0.0});
        slot[fields + i] = d;
        if (last[p.left] == i) free_slots.push_back(slot[p.left]);
        if (p.op != 2 && p.right != p.left && last[p.right] == i) free_slots.push_back(slot[p.right]);
        if (last[fields + i] < 0) free_slots.push_back(d);
    }
    final_slot = slot[fields + n - 1];
    return nslots;
}

}  // namespace

// Device CSG: fields[f] = {device nodes, count, device payload, samples};
// output: device float32 nx*ny*nz.  Status as FE/csg_batch_native_v1/execute.cpp.
extern "C" int fe_csg_device(const void* const* nodes, const int64_t* counts, const void* const* payloads,
                             const int64_t* lengths, int32_t fields, const int32_t* shape, const void* program,
                             int32_t n, float* output, void* stream) noexcept
{
    try {
        if (!nodes || !counts || !payloads || !lengths || !shape || !program || !output) return 1;
        std::vector<int> last;
        const Instr24* prog = static_cast<const Instr24*>(program);
        int v = validate_program(fields, shape, prog, n, last);
        if (v) return v;
        hipStream_t st = static_cast<hipStream_t>(stream);
        const int nx = shape[0], ny = shape[1], nz = shape[2];
        const int64_t N = (int64_t)nx * ny * nz;
        std::vector<FieldRef> refs(fields);
        for (int f = 0; f < fields; ++f) {
            refs[f] = {static_cast<const Node52*>(nodes[f]), static_cast<const uint32_t*>(payloads[f]), counts[f],
                       lengths[f]};
            if (last[f] >= 0 && (!nodes[f] || !payloads[f] || counts[f] < 1 || counts[f] > 33554432 || lengths[f] < 1 ||
                                 lengths[f] > 16777216))
                return 3;
        }
        std::vector<MOp> ops;
        int final_slot = 0;
        const int nslots = compile_program(fields, prog, n, last, ops, final_slot);
        const size_t lds = (size_t)nslots * TS * 4;
        if (lds > 128 * 1024) return 5;  // more simultaneously live values than the tile slots hold
        FieldRef* dref = nullptr;
        MOp* dops = nullptr;
        int* flags = nullptr;  // [0] malformed, [1] first gather-error tile, [2] first non-finite tile
        uint32_t* cover = nullptr;
        int status = 5;
        int hflags[3] = {0, INT_MAX, INT_MAX};
        do {
            if (hipMalloc(&dref, sizeof(FieldRef) * fields) != hipSuccess) break;
            if (hipMalloc(&dops, sizeof(MOp) * ops.size()) != hipSuccess) break;
            if (hipMalloc(&flags, 16) != hipSuccess) break;
            if (hipMalloc(&cover, (size_t)N * 4) != hipSuccess) break;
            if (hipMemcpyAsync(dref, refs.data(), sizeof(FieldRef) * fields, hipMemcpyHostToDevice, st) != hipSuccess) break;
            if (hipMemcpyAsync(dops, ops.data(), sizeof(MOp) * ops.size(), hipMemcpyHostToDevice, st) != hipSuccess) break;
            if (hipMemcpyAsync(flags, hflags, 12, hipMemcpyHostToDevice, st) != hipSuccess) break;
            bool fail = false;
            for (int f = 0; f < fields && !fail; ++f) {
                if (last[f] < 0) continue;
                if (hipMemsetAsync(cover, 0, (size_t)N * 4, st) != hipSuccess) { fail = true; break; }
                hipLaunchKernelGGL(k_field_check, dim3((unsigned)refs[f].count), dim3(256), 0, st, refs[f], nx, ny, nz,
                                   cover, flags);
                hipLaunchKernelGGL(k_cover_scan, dim3((unsigned)std::min<int64_t>((N + 255) / 256, 8192)), dim3(256), 0,
                                   st, cover, nx, ny, nz, flags + 1);
            }
            if (fail) break;
            int pre[2];
            if (hipMemcpyAsync(pre, flags, 8, hipMemcpyDeviceToHost, st) != hipSuccess) break;
            if (hipStreamSynchronize(st) != hipSuccess) break;
            if (pre[0]) { status = 3; break; }  // tile 0's reads already fail
            const int gtx = (nx + TX - 1) / TX, gty = (ny + TY - 1) / TY, gtz = (nz + TZ - 1) / TZ;
            const int ntiles = gtx * gty * gtz;
            if (lds > 65536 &&
                hipFuncSetAttribute(reinterpret_cast<const void*>(&k_csg_tiles), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds) != hipSuccess)
                break;
            hipLaunchKernelGGL(k_csg_tiles, dim3((unsigned)ntiles), dim3(256), lds, st, dref, dops, (int)ops.size(),
                               final_slot, nx, ny, nz, gty, gtz, ntiles, output, flags + 2);
            if (hipGetLastError() != hipSuccess) break;
            if (hipMemcpyAsync(hflags, flags, 12, hipMemcpyDeviceToHost, st) != hipSuccess) break;
            if (hipStreamSynchronize(st) != hipSuccess) break;
            const int g = hflags[1], nf = hflags[2];
            if (g == INT_MAX && nf == INT_MAX) status = 0;
            else status = (g <= nf) ? 3 : 4;
        } while (0);
        if (dref) (void)hipFree(dref);
        if (dops) (void)hipFree(dops);
        if (flags) (void)hipFree(flags);
        if (cover) (void)hipFree(cover);
        return status;
    } catch (...) {
        return 5;
    }
}

// FE/csg_batch_native_v1/execute.cpp:10 (csg_execute) over host arrays:
// upload the used operands, run fe_csg_device, download the dense output.
extern "C" int csg_execute(const void* const* nodes, const int64_t* counts, const void* const* payload,
                           const int64_t* lengths, int32_t fields, const int32_t* shape, const void* program,
                           int32_t instructions, float* output) noexcept
{
    try {
        if (!nodes || !counts || !payload || !lengths || !shape || !program || !output) return 1;
        std::vector<int> last;
        int v = validate_program(fields, shape, static_cast<const Instr24*>(program), instructions, last);
        if (v) return v;
        const int64_t N = (int64_t)shape[0] * shape[1] * shape[2];
        std::vector<void*> dn(fields, nullptr), dp(fields, nullptr);
        float* dout = nullptr;
        int status = 5;
        do {
            bool fail = false;
            for (int f = 0; f < fields; ++f) {
                if (last[f] < 0) continue;
                if (!nodes[f] || !payload[f] || counts[f] < 1 || counts[f] > 33554432 || lengths[f] < 1 ||
                    lengths[f] > 16777216) {
                    status = 3;
                    fail = true;
                    break;
                }
                if (hipMalloc(&dn[f], (size_t)counts[f] * 52) != hipSuccess ||
                    hipMalloc(&dp[f], (size_t)lengths[f] * 4) != hipSuccess ||
                    hipMemcpy(dn[f], nodes[f], (size_t)counts[f] * 52, hipMemcpyHostToDevice) != hipSuccess ||
                    hipMemcpy(dp[f], payload[f], (size_t)lengths[f] * 4, hipMemcpyHostToDevice) != hipSuccess) {
                    fail = true;
                    break;
                }
            }
            if (fail) break;
            if (hipMalloc(&dout, (size_t)N * 4) != hipSuccess) break;
            status = fe_csg_device(const_cast<const void* const*>(dn.data()), counts,
                                   const_cast<const void* const*>(dp.data()), lengths, fields, shape, program,
                                   instructions, dout, nullptr);
            if (status == 0 && hipMemcpy(output, dout, (size_t)N * 4, hipMemcpyDeviceToHost) != hipSuccess) status = 5;
        } while (0);
        for (int f = 0; f < fields; ++f) {
            if (dn[f]) (void)hipFree(dn[f]);
            if (dp[f]) (void)hipFree(dp[f]);
        }
        if (dout) (void)hipFree(dout);
        return status;
    } catch (...) {
        return 5;
    }
}

// ---------------------------------------------------------------- tile reader
namespace {
// Per node, in record order: status 3 for a malformed leaf; for a leaf
// overlapping the tile, every covered target records the first covering node.
__global__ __launch_bounds__(256) void k_gather_first(const Node52* __restrict__ nodes, int64_t count, int64_t samples,
                                                      const int* __restrict__ lower, const int* __restrict__ upper,
                                                      int* __restrict__ first, int* __restrict__ mal)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    const Node52 nd = nodes[i];
    if (nd.children[0] >= 0) return;
    int lo[3], hi[3];
    bool ov = true;
    int64_t size = 1;
    for (int a = 0; a < 3; ++a) {
        if (nd.lo[a] < 0 || nd.shape[a] < 1 || nd.shape[a] > 4096 || nd.lo[a] > 4096 - nd.shape[a] ||
            (nd.stored[a] != 1 && nd.stored[a] != nd.shape[a])) {
            atomicMin(mal, (int)i);
            return;
        }
        size *= nd.stored[a];
        lo[a] = max(lower[a], nd.lo[a]);
        hi[a] = min(upper[a], nd.lo[a] + nd.shape[a]);
        ov = ov && hi[a] > lo[a];
    }
    const int64_t off = noff(nd);
    if (off < 0 || size > samples || off > samples - size) {
        atomicMin(mal, (int)i);
        return;
    }
    if (!ov) return;
    const int dy = upper[1] - lower[1], dz = upper[2] - lower[2];
    for (int x = lo[0]; x < hi[0]; ++x)
        for (int y = lo[1]; y < hi[1]; ++y)
            for (int z = lo[2]; z < hi[2]; ++z)
                atomicMin(&first[((x - lower[0]) * dy + (y - lower[1])) * dz + (z - lower[2])], (int)i);
}

// Second visit: a covered target whose first node is another one marks an
// overlap at this node; the first node copies its bits.
__global__ __launch_bounds__(256) void k_gather_copy(const Node52* __restrict__ nodes, int64_t count,
                                                     const uint32_t* __restrict__ payload, const int* __restrict__ lower,
                                                     const int* __restrict__ upper, const int* __restrict__ first,
                                                     int* __restrict__ ovl, uint32_t* __restrict__ out,
                                                     const int* __restrict__ mal)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= count || (int64_t)*mal <= i) return;  // records after a malformed one are never read
    const Node52 nd = nodes[i];
    if (nd.children[0] >= 0) return;
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = max(lower[a], nd.lo[a]);
        hi[a] = min(upper[a], nd.lo[a] + nd.shape[a]);
        if (hi[a] <= lo[a]) return;
    }
    const int dy = upper[1] - lower[1], dz = upper[2] - lower[2];
    const int64_t off = noff(nd);
    for (int x = lo[0]; x < hi[0]; ++x)
        for (int y = lo[1]; y < hi[1]; ++y)
            for (int z = lo[2]; z < hi[2]; ++z) {
                const int t = ((x - lower[0]) * dy + (y - lower[1])) * dz + (z - lower[2]);
                if (first[t] != (int)i) {
                    atomicMin(ovl, (int)i);
                } else {
                    int64_t s = off;
                    s += (int64_t)(nd.stored[0] == 1 ? 0 : x - nd.lo[0]) * nd.stored[1] * nd.stored[2];
                    s += (int64_t)(nd.stored[1] == 1 ? 0 : y - nd.lo[1]) * nd.stored[2];
                    s += nd.stored[2] == 1 ? 0 : z - nd.lo[2];
                    out[t] = payload[s];
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
            const unsigned g = (unsigned)((count + 255) / 256);
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
