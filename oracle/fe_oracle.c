/*
 * CPU oracle for the block-sparse SDF build path.  TEST INFRASTRUCTURE ONLY:
 * the parity tests, the smoke check and the CPU-baseline leg of the benchmark
 * load this file's shared library as the checker.  The product path never
 * calls it.
 *
 * Every function restates one reference algorithm in plain C99 and cites the
 * reference file:line it follows.  Floating point follows the reference's
 * native build policy: no contraction (-ffp-contract=off), no fast-math,
 * correctly rounded sqrt and division.  Selection min/max use the NumPy/Warp
 * convention "a<b?a:b" / "a>b?a:b" (a tie returns the second operand), which
 * is what FE/csg_batch_native_v1/execute.cpp:37-38 reproduces for NumPy.
 *
 * Parity pinning (which check pins which function) is listed in the tests
 * that exercise each one: tests/test_oracle_golden.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define FE_BIG 1.0e6f

static inline float sel_min(float a, float b) { return a < b ? a : b; }
static inline float sel_max(float a, float b) { return a > b ? a : b; }

/* ------------------------------------------------------------------ EDT -- */

/* Sentinel of FE/field_edt_integer_v1.py:17 and FE/edt_native_v1/edt.cu:73. */
int64_t or_edt_sentinel(int nx, int ny, int nz)
{
    return (int64_t)(nx - 1) * (nx - 1) + (int64_t)(ny - 1) * (ny - 1) +
           (int64_t)(nz - 1) * (nz - 1) + 1;
}

/* Exhaustive separable min-plus, FE/field_edt_integer_v1.py:14-26:
 * f = binary ? sentinel : 0, then for axes 0,1,2: g[p] = min_s f[s]+(p-s)^2.
 * Result: squared distance from each nonzero voxel to the nearest zero voxel. */
void or_edt_exhaustive(const uint8_t *binary, int nx, int ny, int nz, int64_t *out)
{
    int64_t n = (int64_t)nx * ny * nz, sentinel = or_edt_sentinel(nx, ny, nz);
    int64_t *tmp = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) out[i] = binary[i] ? sentinel : 0;
    int len[3] = {nx, ny, nz};
    int64_t stride[3] = {(int64_t)ny * nz, nz, 1};
    for (int axis = 0; axis < 3; ++axis) {
        int L = len[axis];
        int64_t s = stride[axis];
        for (int64_t i = 0; i < n; ++i) {
            int p = (int)((i / s) % L);
            int64_t base = i - (int64_t)p * s, best = INT64_MAX;
            for (int q = 0; q < L; ++q) {
                int64_t d = p - q, v = out[base + q * s] + d * d;
                if (v < best) best = v;
            }
            tmp[i] = best;
        }
        memcpy(out, tmp, sizeof(int64_t) * (size_t)n);
    }
    free(tmp);
}

/* First integer position where the parabola of site q is strictly below the
 * one of site v, with an exact floor for negative numerators:
 * FE/edt_native_linear_v1/edt.cu:8-12. */
static inline int64_t separation(int64_t q, int64_t v, int64_t fq, int64_t fv)
{
    int64_t num = fq + q * q - fv - v * v, den = 2 * (q - v);
    int64_t fl = num >= 0 ? num / den : -((-num + den - 1) / den);
    return fl + 1;
}

/* One lower-envelope pass over every line of one axis, in place through a
 * scratch buffer: FE/edt_native_linear_v1/edt.cu:13-29.  src/dst are int32
 * planes of n voxels; lines are (n/len) lines of len samples at stride s. */
static void envelope_pass(const int32_t *src, int32_t *dst, int64_t n, int64_t s, int L)
{
    int64_t lines = n / L;
#pragma omp parallel
    {
        int *sites = (int *)malloc(sizeof(int) * (size_t)L);
        int64_t *starts = (int64_t *)malloc(sizeof(int64_t) * (size_t)(L + 1));
#pragma omp for schedule(static)
        for (int64_t line = 0; line < lines; ++line) {
            int64_t base = (line / s) * L * s + line % s;
            int k = 0;
            sites[0] = 0;
            starts[0] = INT64_MIN / 2;
            starts[1] = INT64_MAX / 2;
            for (int q = 1; q < L; ++q) {
                int64_t st = separation(q, sites[k], src[base + q * s], src[base + (int64_t)sites[k] * s]);
                while (st <= starts[k]) {
                    --k;
                    st = separation(q, sites[k], src[base + q * s], src[base + (int64_t)sites[k] * s]);
                }
                ++k;
                sites[k] = q;
                starts[k] = st;
                starts[k + 1] = INT64_MAX / 2;
            }
            k = 0;
            for (int q = 0; q < L; ++q) {
                while (starts[k + 1] <= q) ++k;
                int64_t d = q - sites[k];
                dst[base + q * s] = (int32_t)(src[base + (int64_t)sites[k] * s] + d * d);
            }
        }
        free(sites);
        free(starts);
    }
}

/* Envelope EDT of one polarity, FE/edt_native_linear_v1/edt.cu:33-61 without
 * the size caps: squared distance from each nonzero voxel to the nearest zero. */
void or_edt_envelope(const uint8_t *binary, int nx, int ny, int nz, int32_t *out)
{
    int64_t n = (int64_t)nx * ny * nz;
    int32_t sentinel = (int32_t)or_edt_sentinel(nx, ny, nz);
    int32_t *a = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) a[i] = binary[i] ? sentinel : 0;
    int len[3] = {nx, ny, nz};
    int64_t stride[3] = {(int64_t)ny * nz, nz, 1};
    int32_t *src = a, *dst = out;
    for (int axis = 0; axis < 3; ++axis) {
        envelope_pass(src, dst, n, stride[axis], len[axis]);
        int32_t *t = src; src = dst; dst = t;
    }
    if (src != out) memcpy(out, src, sizeof(int32_t) * (size_t)n);
    free(a);
}

/* Signed output of FE/edt_persistent_pair_v1/edt.cu:36-42 (equal to the host
 * form FE/field_edt_integer_v1.py:40-43): correctly rounded f64 sqrt and mul,
 * f32 subtraction and half-pitch correction. */
void or_edt_signed(const int32_t *outside_sq, const int32_t *inside_sq, int64_t n, double pitch, float *out)
{
    float half = (float)(0.5 * pitch);
    for (int64_t i = 0; i < n; ++i) {
        float o = (float)(sqrt((double)outside_sq[i]) * pitch);
        float in = (float)(sqrt((double)inside_sq[i]) * pitch);
        float d = o - in;
        float sg = d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f);
        float c = sg * half;
        out[i] = d - c;
    }
}

/* Paired EDT, FE/edt_persistent_pair_v1/edt.cu:32-35 and :59-78: plane 0 is
 * the squared distance of every non-mask voxel to the mask, plane 1 of every
 * mask voxel to the complement.  squared (2n int32) may be NULL. */
void or_edt_pair(const uint8_t *mask, int nx, int ny, int nz, double pitch, float *out, int32_t *squared)
{
    int64_t n = (int64_t)nx * ny * nz;
    uint8_t *inv = (uint8_t *)malloc((size_t)n);
    int32_t *p = (int32_t *)malloc(sizeof(int32_t) * (size_t)(2 * n));
    for (int64_t i = 0; i < n; ++i) inv[i] = mask[i] ? 0 : 1;
    or_edt_envelope(inv, nx, ny, nz, p);
    or_edt_envelope(mask, nx, ny, nz, p + n);
    if (out) or_edt_signed(p, p + n, n, pitch, out);
    if (squared) memcpy(squared, p, sizeof(int32_t) * (size_t)(2 * n));
    free(p);
    free(inv);
}

/* -------------------------------------------------------- primitives ----- */

/* Recipe record shared with the device library (32 bytes). */
typedef struct {
    int32_t kind; /* 0 sphere,1 box,2/3/4 capped cylinder x/y/z,5 torus z,6 infinite z-cylinder */
    int32_t op;   /* 0 union(min), 1 subtract(max(f,-p)), 2 intersect(max) */
    float cx, cy, cz;
    float a, b, c;
} fe_prim;

/* fp32 primitive formulas.  Sphere, box and capped cylinder follow the
 * emitted IKARUS source FE/ikarus_v1/eval_warp.py:182-224 (operation order
 * as written there); the box of FE/faltkarna_v1.py:91-99 is the same
 * arithmetic; the infinite z-cylinder is FE/faltkarna_v1.py:101-104.  The
 * torus has no reference; it is length(length(x,y)-R, z)-r in the same style. */
float or_prim_eval(const fe_prim *p, float x, float y, float z)
{
    float lx = x - p->cx, ly = y - p->cy, lz = z - p->cz;
    switch (p->kind) {
    case 0: {
        float s = lx * lx + ly * ly;
        s = s + lz * lz;
        return sqrtf(s) - p->a;
    }
    case 1: {
        float qx = fabsf(lx) - p->a, qy = fabsf(ly) - p->b, qz = fabsf(lz) - p->c;
        float ox = sel_max(qx, 0.0f), oy = sel_max(qy, 0.0f), oz = sel_max(qz, 0.0f);
        float s = ox * ox + oy * oy;
        s = s + oz * oz;
        float outside = sqrtf(s);
        float inside = sel_min(sel_max(qx, sel_max(qy, qz)), 0.0f);
        return outside + inside;
    }
    case 2: case 3: case 4: {
        float r2, ax;
        if (p->kind == 4) { r2 = lx * lx + ly * ly; ax = lz; }
        else if (p->kind == 3) { r2 = lx * lx + lz * lz; ax = ly; }
        else { r2 = ly * ly + lz * lz; ax = lx; }
        float qx = sqrtf(r2) - p->a, qy = fabsf(ax) - p->b;
        float ox = sel_max(qx, 0.0f), oy = sel_max(qy, 0.0f);
        float outside = sqrtf(ox * ox + oy * oy);
        float inside = sel_min(sel_max(qx, qy), 0.0f);
        return outside + inside;
    }
    case 5: {
        float qx = sqrtf(lx * lx + ly * ly) - p->a;
        return sqrtf(qx * qx + lz * lz) - p->b;
    }
    case 6:
        return sqrtf(lx * lx + ly * ly) - p->a;
    default:
        return NAN;
    }
}

/* Ordered CSG fold: f = p0, then f = min(f,p) / max(f,-p) / max(f,p)
 * (FE/faltkarna_v1.py:106-126, FE/ikarus_v1/eval_warp.py:295-320 hard ops). */
float or_recipe_eval(const fe_prim *prims, int n, float x, float y, float z)
{
    float f = or_prim_eval(&prims[0], x, y, z);
    for (int i = 1; i < n; ++i) {
        float v = or_prim_eval(&prims[i], x, y, z);
        if (prims[i].op == 0) f = sel_min(f, v);
        else if (prims[i].op == 1) f = sel_max(f, -v);
        else f = sel_max(f, v);
    }
    return f;
}

void or_recipe_eval_points(const fe_prim *prims, int n, const float *pts, int64_t count, float *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < count; ++i)
        out[i] = or_recipe_eval(prims, n, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
}

/* Grid metadata, FE/faltkarna_v1.py:50-64 (margin = fp32(sqrt(3)*B*h)). */
typedef struct {
    float x0, y0, z0, pitch, margin;
    int32_t nx, ny, nz, block;
} fe_grid;

static inline int nblk(int n, int b) { return (n + b - 1) / b; }

/* Eight-corner classification, FE/faltkarna_v1.py:128-164. */
void or_classify(const fe_prim *prims, int n, const fe_grid *g, int32_t *kind, float *cma)
{
    int B = g->block, nbx = nblk(g->nx, B), nby = nblk(g->ny, B), nbz = nblk(g->nz, B);
    int64_t nb = (int64_t)nbx * nby * nbz;
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < nb; ++t) {
        int bk = (int)(t % nbz), bj = (int)((t / nbz) % nby), bi = (int)(t / ((int64_t)nbz * nby));
        float ox = g->x0 + (float)(bi * B) * g->pitch;
        float oy = g->y0 + (float)(bj * B) * g->pitch;
        float oz = g->z0 + (float)(bk * B) * g->pitch;
        float side = (float)B * g->pitch;
        float mn = 1.0e30f;
        int pos = 0, neg = 0;
        for (int c = 0; c < 8; ++c) {
            float dx = (float)(c & 1) * side, dy = (float)((c >> 1) & 1) * side, dz = (float)((c >> 2) & 1) * side;
            float v = or_recipe_eval(prims, n, ox + dx, oy + dy, oz + dz);
            mn = sel_min(mn, fabsf(v));
            if (v >= 0.0f) pos = 1; else neg = 1;
        }
        if (cma) cma[t] = mn;
        kind[t] = (pos && neg) ? 0 : (mn <= g->margin ? 0 : (neg ? -1 : 1));
    }
}

/* Ascending compaction of surface blocks, FE/faltkarna_v1.py:277. */
int64_t or_compact(const int32_t *kind, int64_t nb, int32_t *ids)
{
    int64_t m = 0;
    for (int64_t t = 0; t < nb; ++t)
        if (kind[t] == 0) ids[m++] = (int32_t)t;
    return m;
}

/* Tile evaluation of active blocks, FE/faltkarna_v1.py:166-190 (no clamp;
 * corrupt != 0 plants the +1 block-index fault of :177-180). */
void or_eval_tiles(const fe_prim *prims, int n, const fe_grid *g, const int32_t *ids, int64_t n_active,
                   int corrupt, float *tiles)
{
    int B = g->block, nby = nblk(g->ny, B), nbz = nblk(g->nz, B);
    int64_t b3 = (int64_t)B * B * B;
#pragma omp parallel for schedule(static)
    for (int64_t a = 0; a < n_active; ++a) {
        int64_t lin = ids[a] + (corrupt ? 1 : 0);
        int bk = (int)(lin % nbz), bi = (int)(lin / ((int64_t)nbz * nby)), bj = (int)((lin / nbz) % nby);
        for (int64_t l = 0; l < b3; ++l) {
            int li = (int)(l / (B * B)), lj = (int)((l / B) % B), lk = (int)(l % B);
            float wx = g->x0 + (float)(bi * B + li) * g->pitch;
            float wy = g->y0 + (float)(bj * B + lj) * g->pitch;
            float wz = g->z0 + (float)(bk * B + lk) * g->pitch;
            tiles[a * b3 + l] = or_recipe_eval(prims, n, wx, wy, wz);
        }
    }
}

/* Dense reconstruction cropped to (nx,ny,nz), FE/faltkarna_v1.py:302-325:
 * background -BIG for interior blocks, +BIG otherwise, active tiles scattered. */
void or_reconstruct(const int32_t *kind, const int32_t *ids, int64_t n_active, const float *tiles,
                    int nx, int ny, int nz, int B, float *dense)
{
    int nby = nblk(ny, B), nbz = nblk(nz, B);
    int64_t b3 = (int64_t)B * B * B;
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                int64_t blk = ((int64_t)(x / B) * nby + y / B) * nbz + z / B;
                dense[((int64_t)x * ny + y) * nz + z] = kind[blk] < 0 ? -FE_BIG : FE_BIG;
            }
    for (int64_t a = 0; a < n_active; ++a) {
        int64_t lin = ids[a];
        int bk = (int)(lin % nbz), bi = (int)(lin / ((int64_t)nbz * nby)), bj = (int)((lin / nbz) % nby);
        for (int64_t l = 0; l < b3; ++l) {
            int x = bi * B + (int)(l / (B * B)), y = bj * B + (int)((l / B) % B), z = bk * B + (int)(l % B);
            if (x < nx && y < ny && z < nz) dense[((int64_t)x * ny + y) * nz + z] = tiles[a * b3 + l];
        }
    }
}

/* Deterministic int64 fixed-point checksum, FE/faltkarna_v1.py:249-255:
 * sum over v<0 of int64(f64(v)*scale) (truncation toward zero). */
int64_t or_checksum_i64(const float *vals, int64_t n, double scale)
{
    int64_t s = 0;
    for (int64_t i = 0; i < n; ++i)
        if (vals[i] < 0.0f) s += (int64_t)((double)vals[i] * scale);
    return s;
}

/* Dense-source classification, FE/faltkarna_v1_mesh_to_sdf.py:602-632
 * (corner indices clamped to n-1, margin from f64 rounded to fp32). */
void or_classify_dense(const float *dense, int nx, int ny, int nz, int B, float margin, int32_t *kind)
{
    int nbx = nblk(nx, B), nby = nblk(ny, B), nbz = nblk(nz, B);
    int64_t nb = (int64_t)nbx * nby * nbz;
    for (int64_t t = 0; t < nb; ++t) {
        int bk = (int)(t % nbz), bj = (int)((t / nbz) % nby), bi = (int)(t / ((int64_t)nbz * nby));
        float mn = 1.0e30f;
        int pos = 0, neg = 0;
        for (int c = 0; c < 8; ++c) {
            int cx = bi * B + (c & 1) * B, cy = bj * B + ((c >> 1) & 1) * B, cz = bk * B + ((c >> 2) & 1) * B;
            if (cx > nx - 1) cx = nx - 1;
            if (cy > ny - 1) cy = ny - 1;
            if (cz > nz - 1) cz = nz - 1;
            float v = dense[((int64_t)cx * ny + cy) * nz + cz];
            mn = sel_min(mn, fabsf(v));
            if (v >= 0.0f) pos = 1; else neg = 1;
        }
        kind[t] = (pos && neg) ? 0 : (mn <= margin ? 0 : (neg ? -1 : 1));
    }
}

/* Dense-source tile extraction, FE/faltkarna_v1_mesh_to_sdf.py:635-654 (clamped). */
void or_eval_tiles_dense(const float *dense, int nx, int ny, int nz, int B, const int32_t *ids, int64_t n_active,
                         float *tiles)
{
    int nby = nblk(ny, B), nbz = nblk(nz, B);
    int64_t b3 = (int64_t)B * B * B;
    for (int64_t a = 0; a < n_active; ++a) {
        int64_t lin = ids[a];
        int bk = (int)(lin % nbz), bi = (int)(lin / ((int64_t)nbz * nby)), bj = (int)((lin / nbz) % nby);
        for (int64_t l = 0; l < b3; ++l) {
            int x = bi * B + (int)(l / (B * B)), y = bj * B + (int)((l / B) % B), z = bk * B + (int)(l % B);
            if (x > nx - 1) x = nx - 1;
            if (y > ny - 1) y = ny - 1;
            if (z > nz - 1) z = nz - 1;
            tiles[a * b3 + l] = dense[((int64_t)x * ny + y) * nz + z];
        }
    }
}

/* ------------------------------------------------- leaf-box packing ----- */

#pragma pack(push, 1)
typedef struct {
    int32_t lo[3], shape[3], stored[3], children[2];
    int64_t offset;
} fe_node;
#pragma pack(pop)

typedef struct tnode {
    int lo[3], shape[3], stored[3];
    int64_t cost;
    struct tnode *left, *right;
} tnode;

typedef struct {
    const uint32_t *bits;
    int64_t sx, sy;
    int leaf;
} packer;

static inline uint32_t pk_at(const packer *p, int x, int y, int z)
{
    return p->bits[(int64_t)x * p->sx + (int64_t)y * p->sy + z];
}

/* Recursive bisection with bitwise axis invariance and byte-cost split test,
 * FE/adaptive_sample_blocks_v1.py:19-43. */
static tnode *pk_build(const packer *p, const int lo[3], const int shape[3])
{
    tnode *t = (tnode *)calloc(1, sizeof(tnode));
    int64_t count = 1, reduced = 1;
    for (int a = 0; a < 3; ++a) { t->lo[a] = lo[a]; t->shape[a] = shape[a]; count *= shape[a]; }
    for (int axis = 0; axis < 3; ++axis) {
        int same = 1;
        for (int x = 0; x < shape[0] && same; ++x)
            for (int y = 0; y < shape[1] && same; ++y)
                for (int z = 0; z < shape[2] && same; ++z) {
                    int b[3] = {x, y, z};
                    b[axis] = 0;
                    if (pk_at(p, lo[0] + x, lo[1] + y, lo[2] + z) != pk_at(p, lo[0] + b[0], lo[1] + b[1], lo[2] + b[2]))
                        same = 0;
                }
        t->stored[axis] = same ? 1 : shape[axis];
        reduced *= t->stored[axis];
    }
    t->cost = 52 + 4 * reduced;
    if (reduced <= 64 || count <= p->leaf) return t;
    int axis = -1;
    for (int a = 0; a < 3; ++a)
        if (t->stored[a] > 1 && (axis < 0 || shape[a] > shape[axis])) axis = a;
    int ls[3], rs[3], rl[3];
    for (int a = 0; a < 3; ++a) { ls[a] = rs[a] = shape[a]; rl[a] = lo[a]; }
    ls[axis] = shape[axis] / 2;
    rs[axis] -= ls[axis];
    rl[axis] += ls[axis];
    tnode *l = pk_build(p, lo, ls), *r = pk_build(p, rl, rs);
    int64_t split = 52 + l->cost + r->cost;
    if (split < t->cost) {
        t->cost = split;
        t->left = l;
        t->right = r;
    } else {
        /* discard the candidate subtrees */
        tnode *stack[128];
        int top = 0;
        stack[top++] = l;
        stack[top++] = r;
        while (top) {
            tnode *u = stack[--top];
            if (u->left) { stack[top++] = u->left; stack[top++] = u->right; }
            free(u);
        }
    }
    return t;
}

typedef struct {
    fe_node *nodes;
    int64_t n_nodes, cap_nodes;
    uint32_t *payload;
    int64_t n_payload;
} packed;

/* Pre-order flatten, left first; payload C-order per leaf,
 * FE/adaptive_sample_blocks_v1.py:46-58. */
static int64_t pk_flatten(const packer *p, const tnode *t, packed *out)
{
    int64_t index = out->n_nodes++;
    fe_node node;
    memset(&node, 0, sizeof node);
    for (int a = 0; a < 3; ++a) { node.lo[a] = t->lo[a]; node.shape[a] = t->shape[a]; }
    if (t->left) {
        node.offset = -1;
        node.children[0] = (int32_t)pk_flatten(p, t->left, out);
        node.children[1] = (int32_t)pk_flatten(p, t->right, out);
    } else {
        node.offset = out->n_payload;
        node.children[0] = node.children[1] = -1;
        for (int a = 0; a < 3; ++a) node.stored[a] = t->stored[a];
        for (int x = 0; x < t->stored[0]; ++x)
            for (int y = 0; y < t->stored[1]; ++y)
                for (int z = 0; z < t->stored[2]; ++z)
                    out->payload[out->n_payload++] = pk_at(p, t->lo[0] + x, t->lo[1] + y, t->lo[2] + z);
    }
    out->nodes[index] = node;
    return index;
}

static int64_t pk_count(const tnode *t) { return t->left ? 1 + pk_count(t->left) + pk_count(t->right) : 1; }
static void pk_free(tnode *t)
{
    if (t->left) { pk_free(t->left); pk_free(t->right); }
    free(t);
}

/* Packs a dense C-order float32 (nx,ny,nz) array.  On success returns 0 and
 * fills *nodes (count n_nodes) and *payload (count n_payload), both freed
 * with or_free.  Returns 1 for a non-finite sample (FE/adaptive_pack_native_v1/pack.cpp:63-64). */
int or_pack(const float *data, const int32_t shape[3], int leaf, fe_node **nodes, int64_t *n_nodes,
            uint32_t **payload, int64_t *n_payload)
{
    int64_t total = (int64_t)shape[0] * shape[1] * shape[2];
    const uint32_t *bits = (const uint32_t *)data;
    for (int64_t i = 0; i < total; ++i)
        if ((bits[i] & 0x7f800000u) == 0x7f800000u) return 1;
    packer p = {bits, (int64_t)shape[1] * shape[2], shape[2], leaf};
    int lo[3] = {0, 0, 0}, sh[3] = {shape[0], shape[1], shape[2]};
    tnode *root = pk_build(&p, lo, sh);
    packed out;
    out.n_nodes = 0;
    out.cap_nodes = pk_count(root);
    out.nodes = (fe_node *)malloc(sizeof(fe_node) * (size_t)out.cap_nodes);
    out.payload = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)total);
    out.n_payload = 0;
    pk_flatten(&p, root, &out);
    pk_free(root);
    *nodes = out.nodes;
    *n_nodes = out.n_nodes;
    *payload = out.payload;
    *n_payload = out.n_payload;
    return 0;
}

void or_free(void *p) { free(p); }

/* ------------------------------------------------- tile gather / CSG ----- */

/* Exact bit copies from intersecting leaves into one tile with the status
 * codes of FE/csg_tile_native_v1/gather.cpp:11-50 (1 arguments, 2 bounds,
 * 3 malformed node, 4 overlap or gap). */
int or_gather(const fe_node *nodes, int64_t count, const uint32_t *payload, int64_t samples,
              const int32_t lower[3], const int32_t upper[3], uint32_t *out)
{
    if (!nodes || !payload || !out || count < 1 || count > 33554432 || samples < 1 || samples > 16777216) return 1;
    int64_t total = 1, dims[3];
    for (int a = 0; a < 3; ++a) {
        if (lower[a] < 0 || upper[a] <= lower[a] || upper[a] > 4096) return 2;
        dims[a] = upper[a] - lower[a];
        total *= dims[a];
    }
    if (total > 8192) return 2;
    uint8_t cover[8192];
    memset(cover, 0, sizeof cover);
    for (int64_t n = 0; n < count; ++n) {
        const fe_node *nd = &nodes[n];
        if (nd->children[0] >= 0) continue;
        int lo[3], hi[3], overlap = 1;
        int64_t size = 1;
        for (int a = 0; a < 3; ++a) {
            if (nd->lo[a] < 0 || nd->shape[a] < 1 || nd->shape[a] > 4096 || nd->lo[a] > 4096 - nd->shape[a] ||
                (nd->stored[a] != 1 && nd->stored[a] != nd->shape[a]))
                return 3;
            size *= nd->stored[a];
            lo[a] = lower[a] > nd->lo[a] ? lower[a] : nd->lo[a];
            hi[a] = upper[a] < nd->lo[a] + nd->shape[a] ? upper[a] : nd->lo[a] + nd->shape[a];
            overlap &= hi[a] > lo[a];
        }
        if (nd->offset < 0 || size > samples || nd->offset > samples - size) return 3;
        if (!overlap) continue;
        for (int x = lo[0]; x < hi[0]; ++x)
            for (int y = lo[1]; y < hi[1]; ++y)
                for (int z = lo[2]; z < hi[2]; ++z) {
                    int64_t src = nd->offset;
                    src += (int64_t)(nd->stored[0] == 1 ? 0 : x - nd->lo[0]) * nd->stored[1] * nd->stored[2];
                    src += (int64_t)(nd->stored[1] == 1 ? 0 : y - nd->lo[1]) * nd->stored[2];
                    src += nd->stored[2] == 1 ? 0 : z - nd->lo[2];
                    int64_t dst = ((int64_t)(x - lower[0]) * dims[1] + (y - lower[1])) * dims[2] + (z - lower[2]);
                    if (cover[dst]) return 4;
                    cover[dst] = 1;
                    out[dst] = payload[src];
                }
    }
    for (int64_t i = 0; i < total; ++i)
        if (!cover[i]) return 4;
    return 0;
}

typedef struct {
    int32_t op, left, right, pad;
    double offset;
} fe_instr;

/* Ordered fp32 CSG over (16,32,16) tiles with the status codes of
 * FE/csg_batch_native_v1/execute.cpp:10-54: union a<b?a:b, difference
 * a>(-b)?a:(-b), offset f32(f64(a)-d); 4 on a non-finite result. */
int or_csg_execute(const fe_node *const *nodes, const int64_t *counts, const uint32_t *const *payload,
                   const int64_t *lengths, int32_t fields, const int32_t shape[3], const fe_instr *prog,
                   int32_t ninstr, float *output)
{
    if (!nodes || !counts || !payload || !lengths || !shape || !prog || !output || fields < 1 || fields > 256 ||
        ninstr < 1 || ninstr > 1024)
        return 1;
    int64_t total = 1;
    for (int a = 0; a < 3; ++a) {
        if (shape[a] < 1 || shape[a] > 4096) return 1;
        total *= shape[a];
    }
    if (total > 16777216) return 1;
    int nv = fields + ninstr;
    int *last = (int *)malloc(sizeof(int) * (size_t)nv);
    for (int i = 0; i < nv; ++i) last[i] = -1;
    for (int n = 0; n < ninstr; ++n) {
        const fe_instr *p = &prog[n];
        if (p->op < 0 || p->op > 2 || p->left < 0 || p->left >= fields + n) { free(last); return 2; }
        last[p->left] = n;
        if (p->op == 2) {
            if (!isfinite(p->offset)) { free(last); return 2; }
        } else {
            if (p->right < 0 || p->right >= fields + n) { free(last); return 2; }
            last[p->right] = n;
        }
    }
    last[nv - 1] = ninstr;
    float **vals = (float **)calloc((size_t)nv, sizeof(float *));
    int status = 0;
    for (int x = 0; x < shape[0] && !status; x += 16)
        for (int y = 0; y < shape[1] && !status; y += 32)
            for (int z = 0; z < shape[2] && !status; z += 16) {
                int32_t lo[3] = {x, y, z};
                int32_t hi[3] = {x + 16 < shape[0] ? x + 16 : shape[0], y + 32 < shape[1] ? y + 32 : shape[1],
                                 z + 16 < shape[2] ? z + 16 : shape[2]};
                int size = (hi[0] - x) * (hi[1] - y) * (hi[2] - z);
                for (int f = 0; f < fields && !status; ++f)
                    if (last[f] >= 0) {
                        vals[f] = (float *)malloc(sizeof(float) * (size_t)size);
                        if (or_gather(nodes[f], counts[f], payload[f], lengths[f], lo, hi, (uint32_t *)vals[f]))
                            status = 3;
                    }
                for (int n = 0; n < ninstr && !status; ++n) {
                    const fe_instr *p = &prog[n];
                    float *res = (float *)malloc(sizeof(float) * (size_t)size);
                    const float *a = vals[p->left];
                    for (int i = 0; i < size; ++i) {
                        float r;
                        if (p->op == 2) r = (float)((double)a[i] - p->offset);
                        else {
                            float b = p->op == 0 ? vals[p->right][i] : -vals[p->right][i];
                            r = p->op == 0 ? (a[i] < b ? a[i] : b) : (a[i] > b ? a[i] : b);
                        }
                        if (!isfinite(r)) { status = 4; break; }
                        res[i] = r;
                    }
                    vals[fields + n] = res;
                    if (last[p->left] == n) { free(vals[p->left]); vals[p->left] = NULL; }
                    if (p->op != 2 && p->right != p->left && last[p->right] == n) {
                        free(vals[p->right]);
                        vals[p->right] = NULL;
                    }
                    if (last[fields + n] < 0) { free(vals[fields + n]); vals[fields + n] = NULL; }
                }
                if (!status) {
                    const float *r = vals[nv - 1];
                    int local = 0;
                    for (int a = x; a < hi[0]; ++a)
                        for (int b = y; b < hi[1]; ++b) {
                            int len = hi[2] - z;
                            memcpy(output + ((int64_t)a * shape[1] + b) * shape[2] + z, r + local, sizeof(float) * (size_t)len);
                            local += len;
                        }
                }
                for (int i = 0; i < nv; ++i) { free(vals[i]); vals[i] = NULL; }
            }
    free(vals);
    free(last);
    return status;
}

/* -------------------------------------------------------- point query ---- */

/* Leaf lookup by tree descent, FE/adaptive_sample_blocks_v1.py:103-128. */
static uint32_t sample_at(const fe_node *nodes, const uint32_t *payload, int64_t x, int64_t y, int64_t z)
{
    int64_t id = 0;
    for (;;) {
        const fe_node *nd = &nodes[id];
        if (nd->children[0] < 0) {
            int64_t lx = nd->stored[0] == 1 ? 0 : x - nd->lo[0];
            int64_t ly = nd->stored[1] == 1 ? 0 : y - nd->lo[1];
            int64_t lz = nd->stored[2] == 1 ? 0 : z - nd->lo[2];
            return payload[nd->offset + (lx * nd->stored[1] + ly) * nd->stored[2] + lz];
        }
        const fe_node *l = &nodes[nd->children[0]];
        int in = x >= l->lo[0] && x < l->lo[0] + l->shape[0] && y >= l->lo[1] && y < l->lo[1] + l->shape[1] &&
                 z >= l->lo[2] && z < l->lo[2] + l->shape[2];
        id = in ? nd->children[0] : nd->children[1];
    }
}

void or_samples_at(const fe_node *nodes, const uint32_t *payload, const int64_t *idx, int64_t n, uint32_t *out)
{
    for (int64_t i = 0; i < n; ++i) out[i] = sample_at(nodes, payload, idx[3 * i], idx[3 * i + 1], idx[3 * i + 2]);
}

/* Trilinear query, FE/adaptive_sample_blocks_v1.py:130-151: f64 grid
 * coordinates clipped to [0,n-1], eight clamped neighbours, then order-1
 * spline interpolation as scipy.ndimage.map_coordinates computes it
 * (weights w0 = 1-t, w1 = 1-w0; per term value*w_x*w_y*w_z, terms summed
 * with the last axis fastest). */
void or_query(const fe_node *nodes, const uint32_t *payload, const int32_t shape[3], const double origin[3],
              double pitch, const double *pts, int64_t n, float *out)
{
    for (int64_t i = 0; i < n; ++i) {
        int64_t lower[3];
        double w[3][2];
        for (int a = 0; a < 3; ++a) {
            double c = (pts[3 * i + a] - origin[a]) / pitch;
            double hi = (double)(shape[a] - 1);
            if (c < 0.0) c = 0.0;
            if (c > hi) c = hi;
            double fl = floor(c);
            lower[a] = (int64_t)fl;
            double t = c - fl;
            t -= floor(t);
            w[a][0] = 1.0 - t;
            w[a][1] = 1.0 - w[a][0];
        }
        double acc = 0.0;
        for (int h = 0; h < 8; ++h) {
            int ox = (h >> 2) & 1, oy = (h >> 1) & 1, oz = h & 1;
            int64_t q[3] = {lower[0] + ox, lower[1] + oy, lower[2] + oz};
            for (int a = 0; a < 3; ++a)
                if (q[a] > shape[a] - 1) q[a] = shape[a] - 1;
            uint32_t bits = sample_at(nodes, payload, q[0], q[1], q[2]);
            float v;
            memcpy(&v, &bits, 4);
            double coeff = (double)v;
            coeff *= 1.0;
            coeff *= w[0][ox];
            coeff *= w[1][oy];
            coeff *= w[2][oz];
            acc += coeff;
        }
        out[i] = (float)acc;
    }
}

/* ---------------------------------------------------------------- 2-bit adaptive block tree */
/* Node codes in breadth-first order, FE/field_adaptive_blocks_v1.py:97-138:
 * the roots are the 8-voxel blocks in C order; every node (x, y, z, side) is
 * 0 when it lies outside the domain, else its dense-source kind at that side
 * (FE/faltkarna_v1_mesh_to_sdf.py:602-632 with margin f32(sqrt(3)*side*pitch))
 * decides: exterior 0, interior 1, unresolved side-1 node 3 (a value leaf
 * holding dense[x,y,z]), unresolved larger node 2 (a branch whose children
 * (x+i*h, y+j*h, z+k*h, h), h = side/2, are appended in (i,j,k) order).
 * Returns the node count (codes[0..count)), -1 if cap or vcap is exceeded.
 * hist[4][4]: per level (side 8,4,2,1) the count of each code. */
int64_t or_octree_build(const float *dense, int nx, int ny, int nz, double pitch, uint8_t *codes, int64_t cap,
                        float *values, int64_t vcap, int64_t *nvalues, int64_t *hist)
{
    static const int sides[4] = {8, 4, 2, 1};
    int32_t *lv[4];
    int nb[4][3];
    for (int l = 0; l < 4; ++l) {
        int B = sides[l];
        nb[l][0] = nblk(nx, B);
        nb[l][1] = nblk(ny, B);
        nb[l][2] = nblk(nz, B);
        lv[l] = (int32_t *)malloc(sizeof(int32_t) * (size_t)nb[l][0] * nb[l][1] * nb[l][2]);
        or_classify_dense(dense, nx, ny, nz, B, (float)(sqrt(3.0) * B * pitch), lv[l]);
    }
    for (int i = 0; i < 16; ++i) hist[i] = 0;
    int32_t *qx = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap), *qy = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap),
            *qz = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
    int8_t *ql = (int8_t *)malloc((size_t)cap);
    int64_t tail = 0, nv = 0, result = -1;
    for (int i = 0; i < nb[0][0] && tail <= cap; ++i)
        for (int j = 0; j < nb[0][1]; ++j)
            for (int k = 0; k < nb[0][2]; ++k) {
                if (tail >= cap) goto done;
                qx[tail] = 8 * i; qy[tail] = 8 * j; qz[tail] = 8 * k; ql[tail] = 0; ++tail;
            }
    for (int64_t head = 0; head < tail; ++head) {
        int x = qx[head], y = qy[head], z = qz[head], l = ql[head], s = sides[l];
        uint8_t code;
        if (x >= nx || y >= ny || z >= nz) code = 0;
        else {
            int32_t kind = lv[l][((int64_t)(x / s) * nb[l][1] + y / s) * nb[l][2] + z / s];
            if (kind) code = kind > 0 ? 0 : 1;
            else if (s == 1) {
                if (nv >= vcap) goto done;
                code = 3;
                values[nv++] = dense[((int64_t)x * ny + y) * nz + z];
            } else {
                code = 2;
                int h = s / 2;
                for (int c = 0; c < 8; ++c) {
                    if (tail >= cap) goto done;
                    qx[tail] = x + ((c >> 2) & 1) * h; qy[tail] = y + ((c >> 1) & 1) * h; qz[tail] = z + (c & 1) * h;
                    ql[tail] = (int8_t)(l + 1); ++tail;
                }
            }
        }
        codes[head] = code;
        hist[l * 4 + code] += 1;
    }
    result = tail;
done:
    *nvalues = nv;
    free(qx); free(qy); free(qz); free(ql);
    for (int l = 0; l < 4; ++l) free(lv[l]);
    return result;
}

/* Number of nodes before `node` whose code is `code`, from the checkpoint of
 * its 64-byte group plus the codes in between: FE/field_adaptive_blocks_v1.py:35-50. */
static int64_t octree_rank(const uint8_t *packed, const int32_t *prefix, int64_t node, int code)
{
    int64_t byte = node / 4, group = byte / 64, rank = prefix[group];
    for (int64_t b = group * 64; b < byte; ++b)
        for (int f = 0; f < 4; ++f) rank += ((packed[b] >> (2 * f)) & 3) == code;
    for (int f = 0; f < (int)(node % 4); ++f) rank += ((packed[byte] >> (2 * f)) & 3) == code;
    return rank;
}

/* Tree lookup of FE/field_adaptive_blocks_v1.py:52-82: descend from the root
 * block; exterior/interior leaves give +/- f32(((1.0 * sqrt(3)) * size) * pitch)
 * (f64 product), value leaves their stored sample.  Returns 0, or 1 when a
 * branch would go below side 1. */
int or_octree_sample(const uint8_t *packed, const int32_t *branch_prefix, const int32_t *value_prefix,
                     const float *values, const int32_t root_shape[3], int64_t root_count, double pitch,
                     const int64_t *ijk, int64_t n, float *out)
{
    for (int64_t q = 0; q < n; ++q) {
        int64_t i = ijk[3 * q], j = ijk[3 * q + 1], k = ijk[3 * q + 2];
        int64_t node = ((i / 8) * root_shape[1] + j / 8) * root_shape[2] + k / 8, size = 8;
        for (;;) {
            int code = (packed[node / 4] >> (2 * (node % 4))) & 3;
            if (code < 2) {
                out[q] = (float)(((code == 0 ? 1.0 : -1.0) * sqrt(3.0)) * (double)size * pitch);
                break;
            }
            if (code == 3) {
                out[q] = values[octree_rank(packed, value_prefix, node, 3)];
                break;
            }
            int64_t half = size / 2;
            if (half < 1) return 1;
            int64_t child = 4 * ((i / half) % 2) + 2 * ((j / half) % 2) + (k / half) % 2;
            node = root_count + 8 * octree_rank(packed, branch_prefix, node, 2) + child;
            size = half;
        }
    }
    return 0;
}

/* ----------------------------------------------------------- mesh stage -- */

/* Projected-column scan, FE/cuda_columns_v1/columns.cu:28-48 with the area
 * floor of :61-67 (1e-12 * scale^2, scale = max(1, |x|, |y| over all
 * corners)).  corners: ntri x 9 doubles (a, b, c).  For each query (x, y) the
 * triangles are visited in face order; a triangle is hit when the query lies
 * in its closed xy bounding box, its projected doubled area d clears the
 * floor and the three edge functions l0, l1, l2 have the sign of d (zero
 * allowed).  Depth z = (l0*az + l1*bz + l2*cz)/d, sign = sign(d).  The first
 * `cap` hits are stored at [q*cap + k]; counts hold the full number, which
 * may exceed cap (the reference's status-4 condition). */
void or_columns_scan(const double *corners, int64_t ntri, const double *xy, int64_t nq, int cap,
                     int32_t *counts, double *depths, int32_t *signs)
{
    double scale = 1.0;
    for (int64_t t = 0; t < ntri; ++t)
        for (int j = 0; j < 3; ++j) {
            double ax = fabs(corners[9 * t + 3 * j]), ay = fabs(corners[9 * t + 3 * j + 1]);
            if (ax > scale) scale = ax;
            if (ay > scale) scale = ay;
        }
    const double floor_area = 1e-12 * scale * scale;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t q = 0; q < nq; ++q) {
        const double qx = xy[2 * q], qy = xy[2 * q + 1];
        int32_t found = 0;
        for (int64_t t = 0; t < ntri; ++t) {
            const double *p = corners + 9 * t;
            double ax = p[0], ay = p[1], az = p[2], bx = p[3], by = p[4], bz = p[5], cx = p[6], cy = p[7], cz = p[8];
            double xmin = fmin(ax, fmin(bx, cx)), xmax = fmax(ax, fmax(bx, cx));
            double ymin = fmin(ay, fmin(by, cy)), ymax = fmax(ay, fmax(by, cy));
            if (qx < xmin || qx > xmax || qy < ymin || qy > ymax) continue;
            double d = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
            if (!(fabs(d) > floor_area)) continue;
            double l0 = (bx - qx) * (cy - qy) - (by - qy) * (cx - qx);
            double l1 = (cx - qx) * (ay - qy) - (cy - qy) * (ax - qx);
            double l2 = (ax - qx) * (by - qy) - (ay - qy) * (bx - qx);
            int s = d > 0 ? 1 : -1;
            if (l0 * s < 0 || l1 * s < 0 || l2 * s < 0) continue;
            double z = (l0 * az + l1 * bz + l2 * cz) / d;
            if (found < cap) {
                depths[q * cap + found] = z;
                signs[q * cap + found] = s;
            }
            ++found;
        }
        counts[q] = found;
    }
}

/* Column mask and surface from per-column hits, the device chain of
 * FE/mesh_fused_native_v1/field.cu:5-37 (hit_range, prepare_range,
 * prepare_keys, classify_hits, fused_surface), which equals the host form of
 * FE/column_mask_native_v1.py:27-33 with FE/column_mask_native_v1/mask.cu:5-16.
 *   low/high = min/max stored depth (low 0, span 1 when there is none),
 *   span = max(high - low, 1e-12),
 *   key(hit) = (col + .25) + (.5*(z - low))/span,
 *   fraction(z) = clamp(.25 + (.5*(zs[z] - low))/span, 0, .999),
 *   voxel (col, z) solid iff the signs of hits with key > col + fraction sum
 *   to a nonzero value;
 *   surface = solid voxel on the grid border or with a 6-neighbour outside.
 * Returns the reference's flag bits: 1 some voxel outside, 2 some voxel
 * inside, 4 a count outside [0, cap], 8 a non-finite depth. */
int or_column_mask(const int32_t *counts, const double *depths, const int32_t *signs, int cap, int nx, int ny, int nz,
                   const double *zs, uint8_t *mask, uint8_t *surface)
{
    const int64_t columns = (int64_t)nx * ny, size = columns * nz;
    int flags = 0;
    double lo = INFINITY, hi = -INFINITY;
    for (int64_t c = 0; c < columns; ++c) {
        int n = counts[c];
        if (n < 0 || n > cap) flags |= 4;
        n = n < 0 ? 0 : (n > cap ? cap : n);
        for (int j = 0; j < n; ++j) {
            double z = depths[c * cap + j];
            if (!isfinite(z)) flags |= 8;
            else {
                lo = fmin(lo, z);
                hi = fmax(hi, z);
            }
        }
    }
    double span = 1.0;
    if (!isfinite(lo)) lo = 0.0;
    else span = fmax(hi - lo, 1e-12);
    double *fr = (double *)malloc(sizeof(double) * (size_t)nz);
    for (int z = 0; z < nz; ++z) {
        double f = .25 + (.5 * (zs[z] - lo)) / span;
        fr[z] = fmax(0.0, fmin(.999, f));
    }
    for (int64_t c = 0; c < columns; ++c) {
        int n = counts[c];
        n = n < 0 ? 0 : (n > cap ? cap : n);
        const double base = (double)c + .25;
        for (int z = 0; z < nz; ++z) {
            const double q = (double)c + fr[z];
            int sum = 0;
            for (int j = 0; j < n; ++j) {
                double key = base + (.5 * (depths[c * cap + j] - lo)) / span;
                if (key > q) sum += signs[c * cap + j];
            }
            mask[c * nz + z] = sum != 0;
            flags |= sum != 0 ? 2 : 1;
        }
    }
    free(fr);
    for (int64_t i = 0; i < size; ++i) {
        int64_t z = i % nz, y = (i / nz) % ny, x = i / ((int64_t)ny * nz);
        surface[i] = mask[i] && (x == 0 || x == nx - 1 || y == 0 || y == ny - 1 || z == 0 || z == nz - 1 ||
                                 !mask[i - (int64_t)ny * nz] || !mask[i + (int64_t)ny * nz] || !mask[i - nz] ||
                                 !mask[i + nz] || !mask[i - 1] || !mask[i + 1]);
    }
    return flags;
}

/* Ordered generalized winding number, FE/generalized_winding_ordered_v1.py:11-32:
 * for each point, the sum over triangles in face order of
 * atan2(a.(b x c), |a||b||c| + (a.b)|c| + (b.c)|a| + (c.a)|b|) with a, b, c
 * the corners minus the point, divided by 2*pi.  Dot products add their
 * three terms left to right.  Returns 1 when a point coincides with a corner
 * (the reference's ValueError), else 0. */
int or_gwn(const double *corners, int64_t ntri, const double *points, int64_t n, double *w)
{
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t i = 0; i < n; ++i) {
        const double px = points[3 * i], py = points[3 * i + 1], pz = points[3 * i + 2];
        double total = 0.0;
        for (int64_t t = 0; t < ntri; ++t) {
            const double *p = corners + 9 * t;
            double a0 = p[0] - px, a1 = p[1] - py, a2 = p[2] - pz;
            double b0 = p[3] - px, b1 = p[4] - py, b2 = p[5] - pz;
            double c0 = p[6] - px, c1 = p[7] - py, c2 = p[8] - pz;
            double la = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
            double lb = sqrt(b0 * b0 + b1 * b1 + b2 * b2);
            double lc = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
            if (la == 0 || lb == 0 || lc == 0) bad = 1;
            double x0 = b1 * c2 - b2 * c1, x1 = b2 * c0 - b0 * c2, x2 = b0 * c1 - b1 * c0;
            double num = a0 * x0 + a1 * x1 + a2 * x2;
            double den = la * lb * lc + (a0 * b0 + a1 * b1 + a2 * b2) * lc + (b0 * c0 + b1 * c1 + b2 * c2) * la +
                         (c0 * a0 + c1 * a1 + c2 * a2) * lb;
            total += atan2(num, den);
        }
        w[i] = total / (2.0 * 3.141592653589793);
    }
    return bad;
}
