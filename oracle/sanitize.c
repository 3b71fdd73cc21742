/*
 * AddressSanitizer / UndefinedBehaviorSanitizer run of the C oracle.  TEST
 * INFRASTRUCTURE: built and run by `make -C oracle sanitize` (which
 * tests/test_sanitizers_cpu.py calls); one translation unit with fe_oracle.c.
 *
 * Seeded inputs, each with an invariant that must hold exactly:
 *   pack -> gather of every 7^3 tile and samples_at of random indices give
 *   the input bits (the modes and thresholds of the reference's own
 *   sanitizer exercise, tests/native_csg_sanitizer.cpp:13-35);
 *   CSG over packed operands: valid programs return 0 and equal a direct
 *   dense fold, corrupted ones return 2;
 *   classify -> compact -> tiles -> reconstruct on ragged grids: every
 *   active voxel equals the recipe fold at its sample;
 *   the exhaustive and lower-envelope EDTs agree on random masks.
 * Prints one JSON line on success; exits non-zero at the first mismatch.
 */
#include "fe_oracle.c"

#include <stdio.h>

static uint64_t rng_state = 0x9e3779b97f4a7c15ull;
static uint64_t rnd(void)
{
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return rng_state;
}
static double unif(void) { return (double)(rnd() >> 11) / 9007199254740992.0; }

static int fail(const char *what, long a)
{
    printf("{\"ok\": false, \"case\": \"%s\", \"at\": %ld}\n", what, a);
    return 1;
}

static float *field_of(int mode, const int32_t shape[3])
{
    int64_t total = (int64_t)shape[0] * shape[1] * shape[2];
    float *v = (float *)malloc(sizeof(float) * (size_t)total);
    for (int64_t i = 0; i < total; ++i) {
        v[i] = mode == 3 ? 0.25f : (float)((i * 37) % 101 - 50) * 0.03125f;
        if (mode == 4 && i % 2 == 0) v[i] = -0.0f;
        if (mode == 5) v[i] = (float)(i % shape[2]);
        if (mode == 6) v[i] = (float)(unif() * 8.0 - 4.0);
    }
    return v;
}

static int pack_cases(int *count)
{
    static const int thresholds[3] = {8, 512, 4096};
    for (int mode = 0; mode < 8; ++mode)
        for (int t = 0; t < 3; ++t) {
            int32_t shape[3] = {19, 21, 17};
            if (mode < 3) shape[mode] = 1;
            float *in = field_of(mode, shape);
            fe_node *nodes = NULL;
            uint32_t *pay = NULL;
            int64_t nn = 0, ns = 0;
            if (or_pack(in, shape, thresholds[t], &nodes, &nn, &pay, &ns)) return fail("pack", mode);
            for (int x = 0; x < shape[0]; x += 7)
                for (int y = 0; y < shape[1]; y += 7)
                    for (int z = 0; z < shape[2]; z += 7) {
                        int32_t lo[3] = {x, y, z};
                        int32_t hi[3] = {x + 7 < shape[0] ? x + 7 : shape[0], y + 7 < shape[1] ? y + 7 : shape[1],
                                         z + 7 < shape[2] ? z + 7 : shape[2]};
                        int64_t n = (int64_t)(hi[0] - x) * (hi[1] - y) * (hi[2] - z), q = 0;
                        uint32_t *tile = (uint32_t *)malloc(4 * (size_t)n);
                        if (or_gather(nodes, nn, pay, ns, lo, hi, tile)) return fail("gather status", mode);
                        for (int a = x; a < hi[0]; ++a)
                            for (int b = y; b < hi[1]; ++b)
                                for (int c = z; c < hi[2]; ++c) {
                                    int64_t i = ((int64_t)a * shape[1] + b) * shape[2] + c;
                                    if (memcmp(&tile[q++], &in[i], 4)) return fail("gather bits", (long)i);
                                }
                        free(tile);
                    }
            int64_t idx[3 * 64];
            uint32_t got[64];
            for (int k = 0; k < 64; ++k)
                for (int a = 0; a < 3; ++a) idx[3 * k + a] = (int64_t)(rnd() % (uint64_t)shape[a]);
            or_samples_at(nodes, pay, idx, 64, got);
            for (int k = 0; k < 64; ++k) {
                int64_t i = (idx[3 * k] * shape[1] + idx[3 * k + 1]) * shape[2] + idx[3 * k + 2];
                if (memcmp(&got[k], &in[i], 4)) return fail("samples_at", (long)i);
            }
            or_free(nodes);
            or_free(pay);
            free(in);
            ++*count;
        }
    return 0;
}

static int csg_cases(int *count)
{
    const int32_t shape[3] = {19, 21, 17};
    const int64_t total = 19 * 21 * 17;
    enum { F = 3 };
    float *in[F];
    fe_node *nodes[F];
    uint32_t *pay[F];
    int64_t nn[F], ns[F];
    for (int f = 0; f < F; ++f) {
        in[f] = field_of(6, shape);
        if (or_pack(in[f], shape, 64, &nodes[f], &nn[f], &pay[f], &ns[f])) return fail("csg pack", f);
    }
    float *out = (float *)malloc(sizeof(float) * (size_t)total);
    float *ref = (float *)malloc(sizeof(float) * (size_t)total * (F + 40));
    for (int t = 0; t < 200; ++t) {
        int n = 1 + (int)(rnd() % 40);
        fe_instr prog[40];
        for (int i = 0; i < n; ++i) {
            prog[i].op = (int)(rnd() % 3);
            prog[i].left = (int)(rnd() % (uint64_t)(F + i));
            prog[i].right = prog[i].op == 2 ? 0 : (int)(rnd() % (uint64_t)(F + i));
            prog[i].pad = 0;
            prog[i].offset = (double)(int)(rnd() % 17) / 8.0 - 1.0;
        }
        int corrupt = t % 5 == 4;
        if (corrupt) prog[rnd() % (uint64_t)n].left = F + n + 3;
        const fe_node *np[F] = {nodes[0], nodes[1], nodes[2]};
        const uint32_t *pp[F] = {pay[0], pay[1], pay[2]};
        int s = or_csg_execute(np, nn, pp, ns, F, shape, prog, n, out);
        if (s != (corrupt ? 2 : 0)) return fail("csg status", s);
        if (corrupt) continue;
        for (int f = 0; f < F; ++f) memcpy(ref + (int64_t)f * total, in[f], sizeof(float) * (size_t)total);
        for (int i = 0; i < n; ++i) {
            float *r = ref + (int64_t)(F + i) * total;
            const float *a = ref + (int64_t)prog[i].left * total, *b = ref + (int64_t)prog[i].right * total;
            for (int64_t k = 0; k < total; ++k) {
                if (prog[i].op == 0) r[k] = sel_min(a[k], b[k]);
                else if (prog[i].op == 1) r[k] = sel_max(a[k], -b[k]);
                else r[k] = (float)((double)a[k] - prog[i].offset);
            }
        }
        if (memcmp(out, ref + (int64_t)(F + n - 1) * total, sizeof(float) * (size_t)total))
            return fail("csg values", t);
        ++*count;
    }
    free(out);
    free(ref);
    for (int f = 0; f < F; ++f) {
        or_free(nodes[f]);
        or_free(pay[f]);
        free(in[f]);
    }
    return 0;
}

static int build_cases(int *count)
{
    for (int t = 0; t < 24; ++t) {
        int np = 1 + (int)(rnd() % 12), B = t % 2 ? 8 : 4;
        fe_grid g = {(float)(unif() * 4 - 2), (float)(unif() * 4 - 2), (float)(unif() * 4 - 2), 0.5f, 0.0f,
                     9 + (int)(rnd() % 30), 9 + (int)(rnd() % 30), 9 + (int)(rnd() % 30), B};
        g.margin = (float)(sqrt(3.0) * B * g.pitch);
        fe_prim *p = (fe_prim *)malloc(sizeof(fe_prim) * (size_t)np);
        for (int i = 0; i < np; ++i) {
            p[i].kind = (int)(rnd() % 7);
            p[i].op = i ? (int)(rnd() % 3) : 0;
            p[i].cx = (float)(unif() * 16);
            p[i].cy = (float)(unif() * 16);
            p[i].cz = (float)(unif() * 16);
            p[i].a = (float)(1 + unif() * 6);
            p[i].b = (float)(0.5 + unif() * 3);
            p[i].c = (float)(1 + unif() * 6);
        }
        int nbx = nblk(g.nx, B), nby = nblk(g.ny, B), nbz = nblk(g.nz, B);
        int64_t nb = (int64_t)nbx * nby * nbz;
        int32_t *kind = (int32_t *)malloc(4 * (size_t)nb), *ids = (int32_t *)malloc(4 * (size_t)nb);
        float *cma = (float *)malloc(4 * (size_t)nb);
        or_classify(p, np, &g, kind, cma);
        int64_t na = or_compact(kind, nb, ids);
        float *tiles = (float *)malloc(4 * (size_t)(na ? na : 1) * B * B * B);
        or_eval_tiles(p, np, &g, ids, na, 0, tiles);
        int64_t N = (int64_t)g.nx * g.ny * g.nz;
        float *dense = (float *)malloc(4 * (size_t)N);
        or_reconstruct(kind, ids, na, tiles, g.nx, g.ny, g.nz, B, dense);
        for (int64_t a = 0; a < na; ++a) {
            int64_t b = ids[a];
            int bi = (int)(b / ((int64_t)nby * nbz)), bj = (int)((b / nbz) % nby), bk = (int)(b % nbz);
            int i = bi * B + (int)(rnd() % (uint64_t)B), j = bj * B + (int)(rnd() % (uint64_t)B),
                k = bk * B + (int)(rnd() % (uint64_t)B);
            if (i >= g.nx || j >= g.ny || k >= g.nz) continue;
            float x = g.x0 + (float)i * g.pitch, y = g.y0 + (float)j * g.pitch, z = g.z0 + (float)k * g.pitch;
            float want = or_recipe_eval(p, np, x, y, z);
            if (memcmp(&dense[((int64_t)i * g.ny + j) * g.nz + k], &want, 4)) return fail("build voxel", (long)t);
        }
        free(p);
        free(kind);
        free(ids);
        free(cma);
        free(tiles);
        free(dense);
        ++*count;
    }
    return 0;
}

static int edt_cases(int *count)
{
    for (int t = 0; t < 40; ++t) {
        int nx = 1 + (int)(rnd() % 12), ny = 1 + (int)(rnd() % 12), nz = 1 + (int)(rnd() % 12);
        int64_t n = (int64_t)nx * ny * nz;
        double fill = unif();
        uint8_t *m = (uint8_t *)malloc((size_t)n);
        for (int64_t i = 0; i < n; ++i) m[i] = unif() < fill;
        int64_t *ex = (int64_t *)malloc(8 * (size_t)n);
        int32_t *env = (int32_t *)malloc(4 * (size_t)n), *sq = (int32_t *)malloc(8 * (size_t)n);
        float *sd = (float *)malloc(4 * (size_t)n);
        or_edt_exhaustive(m, nx, ny, nz, ex);
        or_edt_envelope(m, nx, ny, nz, env);
        for (int64_t i = 0; i < n; ++i)
            if (ex[i] != (int64_t)env[i]) return fail("edt envelope", (long)t);
        or_edt_pair(m, nx, ny, nz, 0.5, sd, sq);
        free(m);
        free(ex);
        free(env);
        free(sq);
        free(sd);
        ++*count;
    }
    return 0;
}

int main(void)
{
    int packs = 0, csgs = 0, builds = 0, edts = 0;
    if (pack_cases(&packs) || csg_cases(&csgs) || build_cases(&builds) || edt_cases(&edts)) return 1;
    printf("{\"ok\": true, \"pack\": %d, \"csg\": %d, \"build\": %d, \"edt\": %d, "
           "\"sanitizers\": \"address,undefined\"}\n", packs, csgs, builds, edts);
    return 0;
}
