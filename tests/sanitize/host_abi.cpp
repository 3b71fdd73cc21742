// AddressSanitizer / UndefinedBehaviorSanitizer run of the library's host
// code with no GPU: the argument and program checks of the C ABI and the CSG
// program compiler (validation, evaluation order, slot allocation), driven
// with seeded random and malformed inputs.  Built host-only (no device code)
// by tests/test_sanitizers_cpu.py; every entry point called here returns
// before any kernel launch, or at the first HIP call, which fails without a
// device.  Prints one JSON line on success; exits non-zero on a wrong status.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

extern "C" {
const char* fe_last_error();
int fe_ikarus_check(const void* ins_host, int n_ins, int64_t n_params, int n_grids);
int64_t fe_csg_workspace_bytes(int32_t fields, const int32_t* shape, const void* program, int32_t instructions);
int csg_execute(const void* const* nodes, const int64_t* counts, const void* const* payloads, const int64_t* lengths,
                int32_t fields, const int32_t* shape, const void* program, int32_t instructions, float* output);
}

struct Instr24 {
    int32_t op, left, right, pad;
    double offset;
};

static int fail(const char* what, long a, long b)
{
    std::printf("{\"ok\": false, \"case\": \"%s\", \"got\": %ld, \"want\": %ld}\n", what, a, b);
    return 1;
}

int main()
{
    std::mt19937_64 rng(7);
    long ik_ok = 0, ik_bad = 0, csg_ok = 0, csg_bad = 0;

    // IKARUS programs: random instruction streams, most malformed.
    const int ops[] = {0, 1, 2, 3, 10, 11, 12, 13, 14, 15, 16, 20, 21, 22, 23, 24, 4, 99, -1};
    for (int t = 0; t < 20000; ++t) {
        const int n = 1 + (int)(rng() % 48);
        std::vector<int32_t> ins(4 * n);
        for (int i = 0; i < n; ++i) {
            ins[4 * i] = ops[rng() % (sizeof ops / sizeof ops[0])];
            ins[4 * i + 1] = (int)(rng() % 5) - 1;
            ins[4 * i + 2] = (int)(rng() % 40) - 2;
            ins[4 * i + 3] = 0;
        }
        const int s = fe_ikarus_check(ins.data(), n, 32, 2);
        if (s != 0 && s != 1) return fail("ikarus status", s, 1);
        if (s == 1 && !std::strlen(fe_last_error())) return fail("ikarus message", 0, 1);
        (s ? ik_bad : ik_ok)++;
    }
    // well-formed: a balanced tree at the stack bound, and one past it
    for (int depth : {17, 18}) {
        std::vector<int32_t> ins;
        for (int i = 0; i < depth; ++i) ins.insert(ins.end(), {11, 0, 0, 0});
        for (int i = 0; i + 1 < depth; ++i) ins.insert(ins.end(), {22, 0, 4, 0});
        const int s = fe_ikarus_check(ins.data(), (int)ins.size() / 4, 8, 0);
        if (s != (depth > 17)) return fail("ikarus depth", s, depth > 17);
    }
    if (fe_ikarus_check(nullptr, 1, 1, 0) != 1 || fe_ikarus_check(ops, 0, 1, 0) != 1) return fail("ikarus args", 0, 1);

    // CSG programs: random programs over up to 256 fields and 1024
    // instructions, valid ones through the slot compiler (the workspace
    // query stops at the first HIP call: -1 without a device).
    for (int t = 0; t < 4000; ++t) {
        const int fields = 1 + (int)(rng() % (t % 8 == 0 ? 256 : 12));
        const int n = 1 + (int)(rng() % (t % 8 == 0 ? 1024 : 40));
        std::vector<Instr24> prog(n);
        const bool corrupt = rng() % 4 == 0;
        for (int i = 0; i < n; ++i) {
            Instr24& p = prog[i];
            p.op = (int)(rng() % 3);
            p.left = (int)(rng() % (fields + i));
            p.right = p.op == 2 ? -1 : (int)(rng() % (fields + i));
            p.pad = 0;
            p.offset = p.op == 2 ? (double)(rng() % 1000) / 37.0 - 13.0 : 0.0;
        }
        if (corrupt) {
            Instr24& p = prog[rng() % n];
            switch (rng() % 4) {
                case 0: p.op = 3 + (int)(rng() % 5); break;
                case 1: p.left = fields + n + (int)(rng() % 9); break;
                case 2: p.op = 0; p.right = -1 - (int)(rng() % 3); break;
                default: p.op = 2; p.offset = NAN; break;
            }
        }
        const int32_t shape[3] = {1 + (int)(rng() % 200), 1 + (int)(rng() % 200), 1 + (int)(rng() % 200)};  // <= 2^24 samples
        const int64_t ws = fe_csg_workspace_bytes(fields, shape, prog.data(), n);
        if (ws != -1) return fail("csg workspace without a device", (long)ws, -1);
        // csg_execute: program errors (2) come before operand and device work
        std::vector<const void*> nodes(fields, nullptr), pay(fields, nullptr);
        std::vector<int64_t> counts(fields, 0), lens(fields, 0);
        float out = 0.0f;
        const int s = csg_execute(nodes.data(), counts.data(), pay.data(), lens.data(), fields, shape, prog.data(), n,
                                  &out);
        if (corrupt && s != 2) return fail("csg corrupt program", s, 2);
        if (!corrupt && s != 3) return fail("csg null operands", s, 3);
        (corrupt ? csg_bad : csg_ok)++;
    }
    // argument checks
    const int32_t shape[3] = {4, 4, 4}, big[3] = {4097, 1, 1};
    Instr24 one{0, 0, 0, 0, 0.0};
    const void* np[1] = {nullptr};
    int64_t c0[1] = {0};
    float out = 0.0f;
    if (csg_execute(nullptr, c0, np, c0, 1, shape, &one, 1, &out) != 1) return fail("csg null nodes", 0, 1);
    if (csg_execute(np, c0, np, c0, 0, shape, &one, 1, &out) != 1) return fail("csg zero fields", 0, 1);
    if (csg_execute(np, c0, np, c0, 1, big, &one, 1, &out) != 1) return fail("csg axis cap", 0, 1);
    if (csg_execute(np, c0, np, c0, 1, shape, &one, 0, &out) != 1) return fail("csg no instructions", 0, 1);
    if (fe_csg_workspace_bytes(1, big, &one, 1) != -1) return fail("csg workspace axis cap", 0, -1);

    std::printf("{\"ok\": true, \"ikarus_valid\": %ld, \"ikarus_rejected\": %ld, \"csg_valid\": %ld, "
                "\"csg_rejected\": %ld, \"sanitizers\": \"address,undefined\"}\n",
                ik_ok, ik_bad, csg_ok, csg_bad);
    return 0;
}
