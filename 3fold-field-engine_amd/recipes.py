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
