"""Recipes (ordered primitive folds) and the seeded synthetic workloads.

A recipe is a numpy structured array of PRIM records evaluated as
f = p0; f = min(f, p) (union), max(f, -p) (subtract) or max(f, p) (intersect)
in array order -- the hard CSG ops of FE/ikarus_v1/eval_warp.py:295-320 and
FE/faltkarna_v1.py:106-126.  Primitive formulas follow the emitted fp32
IKARUS source (FE/ikarus_v1/eval_warp.py:182-224); parameters are stored as
fp32, as the reference bakes them into fp32 literals.

Kinds: 0 sphere(r) 1 box(hx,hy,hz) 2/3/4 capped cylinder along x/y/z
(radius, half-height) 5 torus about z (R, r) 6 infinite z-cylinder (r).
"""
import math

import numpy as np

PRIM = np.dtype([("kind", "<i4"), ("op", "<i4"), ("cx", "<f4"), ("cy", "<f4"), ("cz", "<f4"),
                 ("a", "<f4"), ("b", "<f4"), ("c", "<f4")])
UNION, SUBTRACT, INTERSECT = 0, 1, 2
SPHERE, BOX, CYL_X, CYL_Y, CYL_Z, TORUS, ZCYL = 0, 1, 2, 3, 4, 5, 6
MAX_PRIMS = 4096


def make(records):
    """records: iterable of (kind, op, (cx,cy,cz), (a,b,c))."""
    out = np.zeros(len(records), PRIM)
    for i, (kind, op, c, p) in enumerate(records):
        out[i] = (kind, op, *c, *(tuple(p) + (0.0,) * (3 - len(p))))
    return out


def validate(recipe):
    r = np.asarray(recipe)
    if r.dtype != PRIM or r.ndim != 1 or not 1 <= len(r) <= MAX_PRIMS:
        raise ValueError(f"Recipe of 1..{MAX_PRIMS} PRIM records required")
    if np.any((r["kind"] < 0) | (r["kind"] > 6)) or np.any((r["op"] < 0) | (r["op"] > 2)):
        raise ValueError("Unknown primitive kind or operation")
    vals = np.stack([r[k] for k in ("cx", "cy", "cz", "a", "b", "c")])
    if not np.isfinite(vals).all():
        raise ValueError("Finite primitive parameters required")
    return np.ascontiguousarray(r)


def testfalt(hx, hy, hz, r):
    """The reference test field, box minus a through z-cylinder (FE/faltkarna_v1.py:121-126)."""
    return make([(BOX, UNION, (0, 0, 0), (hx, hy, hz)), (ZCYL, SUBTRACT, (0, 0, 0), (r,))])


def extent(recipe):
    """Largest coordinate/size magnitude of a recipe (for the culling slack)."""
    r = np.asarray(recipe)
    return float(np.max(np.abs(np.stack([r["cx"], r["cy"], r["cz"]])))) + \
        float(np.max(np.abs(np.stack([r["a"], r["b"], r["c"]]))))


# ---------------------------------------------------------------- seeded workloads
def plain_part_128(seed=1):
    """128^3 part at 1 mm: 40 spheres r~U[2,12], 24 boxes half-extents~U[2,15]^3,
    centres~U[0,128)^3, fold order a seeded permutation whose last 8 are subtractions."""
    rng = np.random.default_rng(seed)
    recs = []
    for _ in range(40):
        recs.append((SPHERE, UNION, tuple(rng.uniform(0, 128, 3)), (rng.uniform(2, 12),)))
    for _ in range(24):
        recs.append((BOX, UNION, tuple(rng.uniform(0, 128, 3)), tuple(rng.uniform(2, 15, 3))))
    order = rng.permutation(len(recs))
    recs = [recs[i] for i in order]
    for i in range(len(recs) - 8, len(recs)):
        recs[i] = (recs[i][0], SUBTRACT, recs[i][2], recs[i][3])
    return make(recs)


def mixed_recipe(rng, n_prims, n, spread=0.3):
    """n_prims primitives, a quarter each of sphere, box, capped cylinder (random
    axis) and z-torus (r/R~U[0.2,0.5]); sizes log-uniform in [1,40] voxels
    (scaled by n/512); centres uniform in a central cube of side spread*n; fold
    order a seeded permutation whose last eighth are subtractions."""
    scale = n / 512.0
    recs = []
    kinds = np.arange(n_prims) % 4
    lo, hi = 0.5 * n * (1 - spread), 0.5 * n * (1 + spread)
    for k in kinds:
        c = tuple(rng.uniform(lo, hi, 3))
        if k == 0:
            recs.append((SPHERE, UNION, c, (math.exp(rng.uniform(0, math.log(40))) * scale,)))
        elif k == 1:
            recs.append((BOX, UNION, c, tuple(np.exp(rng.uniform(0, math.log(40), 3)) * scale / 2)))
        elif k == 2:
            axis = int(rng.integers(0, 3))
            recs.append((CYL_X + axis, UNION, c, (math.exp(rng.uniform(0, math.log(40))) * scale / 2,
                                                  math.exp(rng.uniform(0, math.log(40))) * scale / 2)))
        else:
            big = math.exp(rng.uniform(0, math.log(40))) * scale / 2
            recs.append((TORUS, UNION, c, (big, big * rng.uniform(0.2, 0.5))))
    order = rng.permutation(len(recs))
    recs = [recs[i] for i in order]
    n_sub = len(recs) // 8
    for i in range(len(recs) - n_sub, len(recs)):
        if i > 0:
            recs[i] = (recs[i][0], SUBTRACT, recs[i][2], recs[i][3])
    return make(recs)


def large_field_512(seed=2):
    """1024 mixed primitives in a 512^3 domain at 1 mm."""
    return mixed_recipe(np.random.default_rng(seed), 1024, 512)


def served_requests(seed=5, count=48):
    """`count` recipes of the large-field distribution (1024 mixed primitives,
    512^3), drawn one after another from one seeded generator: the request
    stream of the field server."""
    rng = np.random.default_rng(seed)
    return [mixed_recipe(rng, 1024, 512) for _ in range(count)]


def batch_recipes(seed=3, count=256):
    """count recipes: primitive counts log-uniform in [4,2048], cubic domains
    uniform in {64,96,...,384}; returns [(recipe, n)]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        p = int(round(math.exp(rng.uniform(math.log(4), math.log(2048)))))
        n = int(rng.choice(np.arange(64, 385, 32)))
        out.append((mixed_recipe(rng, p, n), n))
    return out


# ---------------------------------------------------------------- obstacle volumes (EDT workloads)
def obstacle_volume(shape, rng, count, scale=100.0):
    """uint8 occupancy of `count` seeded axis-aligned obstacles, in voxel units.

    Kinds (probabilities): solid box 0.35 (edges log-uniform in [2, scale]);
    1-voxel-thick wall 0.25 (the other two edges log-uniform in [8, 2*scale]);
    capped cylinder along a random axis 0.25 (radius log-uniform in
    [1, scale/3], length log-uniform in [4, 2*scale]; voxel (i,j) of the
    cross-section is inside when (i-ci)^2 + (j-cj)^2 <= r^2); split box 0.15,
    a solid box with a 1-voxel gap cut through its middle along a random axis.
    Obstacles are painted in order (later gaps clear earlier obstacles).  At
    scale 100 the occupancy is about 12% for any volume when `count` is
    proportional to the voxel count (4000 at 768^3)."""
    n = np.asarray(shape, dtype=np.int64)
    vol = np.zeros(tuple(int(x) for x in shape), np.uint8)
    kinds = rng.choice(4, size=count, p=[0.35, 0.25, 0.25, 0.15])
    log = np.log
    for k in kinds:
        if k in (0, 3):
            e = np.rint(np.exp(rng.uniform(log(2), log(scale), 3))).astype(np.int64)
            lo = (rng.uniform(0, 1, 3) * n).astype(np.int64) - e // 2
            hi = np.minimum(lo + e, n)
            lo = np.maximum(lo, 0)
            vol[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = 1
            if k == 3:
                ax = int(rng.integers(3))
                if hi[ax] - lo[ax] >= 3:
                    g = (lo[ax] + hi[ax]) // 2
                    sl = [slice(lo[0], hi[0]), slice(lo[1], hi[1]), slice(lo[2], hi[2])]
                    sl[ax] = slice(g, g + 1)
                    vol[tuple(sl)] = 0
        elif k == 1:
            ax = int(rng.integers(3))
            e = np.rint(np.exp(rng.uniform(log(8), log(2 * scale), 3))).astype(np.int64)
            e[ax] = 1
            lo = (rng.uniform(0, 1, 3) * n).astype(np.int64) - e // 2
            hi = np.minimum(lo + e, n)
            lo = np.maximum(lo, 0)
            vol[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = 1
        else:
            ax = int(rng.integers(3))
            r = float(np.exp(rng.uniform(0.0, log(scale / 3))))
            length = int(np.rint(np.exp(rng.uniform(log(4), log(2 * scale)))))
            c = (rng.uniform(0, 1, 3) * n).astype(np.int64)
            ri = int(np.ceil(r))
            lo, hi = c - ri, c + ri + 1
            lo[ax] = c[ax] - length // 2
            hi[ax] = lo[ax] + length
            lo, hi = np.maximum(lo, 0), np.minimum(hi, n)
            if np.any(hi <= lo):
                continue
            idx = [np.arange(lo[a], hi[a]) - c[a] for a in range(3)]
            idx[ax] = np.zeros(1, np.int64)
            d2 = idx[0][:, None, None] ** 2 + idx[1][None, :, None] ** 2 + idx[2][None, None, :] ** 2
            disk = (d2 <= r * r).astype(np.uint8)
            sl = tuple(slice(lo[a], hi[a]) for a in range(3))
            vol[sl] |= np.broadcast_to(disk, tuple(int(x) for x in hi - lo))
    return vol


def obstacle_count(shape, per_768=4000):
    """Obstacle count proportional to the voxel count (4000 at 768^3)."""
    return max(1, int(round(per_768 * float(np.prod(shape)) / 768.0 ** 3)))


def thin_feature_volume_768(seed=4):
    """768^3 from 4000 seeded obstacles (about 12% occupied)."""
    return obstacle_volume((768, 768, 768), np.random.default_rng(seed), 4000)


def thin_feature_shapes(seed=4, count=64, lo=96, hi=256):
    """Edges of `count` uneven volumes, each axis uniform in [lo, hi]."""
    rng = np.random.default_rng([seed, 1])
    return [tuple(int(x) for x in rng.integers(lo, hi + 1, 3)) for _ in range(count)]


def thin_feature_batch(seed=4, count=64, lo=96, hi=256):
    """`count` volumes of the 768^3 construction with edges in [lo, hi]."""
    out = []
    for i, shape in enumerate(thin_feature_shapes(seed, count, lo, hi)):
        out.append(obstacle_volume(shape, np.random.default_rng([seed, 2, i]), obstacle_count(shape)))
    return out


# ------------------------------------------------------------------ meshes
def torus_mesh(R, r, nu, nv, center, axis_angles=(0.0, 0.0)):
    """Closed, outward-wound torus of nu x nv quads (two triangles each),
    tilted by axis_angles (about x, then z) and moved to `center`."""
    u = 2 * np.pi * np.arange(nu) / nu
    w = 2 * np.pi * np.arange(nv) / nv
    U, W = np.meshgrid(u, w, indexing="ij")
    v = np.stack([(R + r * np.cos(W)) * np.cos(U), (R + r * np.cos(W)) * np.sin(U), r * np.sin(W)], -1).reshape(-1, 3)
    a, b = axis_angles
    rx = np.array([[1, 0, 0], [0, math.cos(a), -math.sin(a)], [0, math.sin(a), math.cos(a)]])
    rz = np.array([[math.cos(b), -math.sin(b), 0], [math.sin(b), math.cos(b), 0], [0, 0, 1]])
    v = v @ (rz @ rx).T + np.asarray(center, np.float64)
    i, j = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    p00 = (i * nv + j).ravel()
    p10 = (((i + 1) % nu) * nv + j).ravel()
    p01 = (i * nv + (j + 1) % nv).ravel()
    p11 = (((i + 1) % nu) * nv + (j + 1) % nv).ravel()
    f = np.concatenate([np.stack([p00, p10, p11], 1), np.stack([p00, p11, p01], 1)])
    return v, f.astype(np.int64)


def mesh_scene(seed=7, count=64, extent=500.0, nu=128, nv=32):
    """Closed triangle soup of `count` seeded tori (R ~ U[15, 45] mm, r/R ~
    U[0.15, 0.35], random tilt) inside [0, extent)^3: count*nu*nv*2
    triangles (524,288 by default)."""
    rng = np.random.default_rng(seed)
    vs, fs, off = [], [], 0
    for _ in range(count):
        R = rng.uniform(15, 45)
        r = R * rng.uniform(0.15, 0.35)
        c = rng.uniform(R + r, extent - R - r, 3)
        v, f = torus_mesh(R, r, nu, nv, c, tuple(rng.uniform(0, math.pi, 2)))
        vs.append(v)
        fs.append(f + off)
        off += len(v)
    return np.concatenate(vs), np.concatenate(fs)
