"""Pins the oracle's recipe primitives (oracle/fe_oracle.c or_prim_eval)
independently of their own formulas.

* Sphere, box and the capped cylinders along x, y and z equal, bit for bit,
  the fp32 arithmetic the reference emits for the IKARUS sphere, box and
  cylinder nodes (FE/ikarus_v1/eval_warp.py:182-224, restated by
  oracle/ikarus_ref.py from that text): a recipe cylinder of radius a and
  half-height b is the node cylinder(radius=a, height=2b, axis).  A wrong
  axis mapping or a full height taken for a half height fails here.
* Capped cylinders and the torus (which the reference lacks) against brute
  force: the distance to a dense sampling of the surface, with the sign from
  the solid's own membership test.  A swapped R and r, a wrong axis or a
  wrong height convention fails here.
* A sphere of radius 20 at pitch 1 and 0.5 (centre on a cell centre): the
  voxel volume (count of samples with sdf < 0) within 2e-3 of 4/3 pi r^3.
* The fold order: union = min, subtraction = max(f, -p), intersection =
  max, ties to the second operand (+0/-0).
"""
import math

import numpy as np
import pytest

import ikarus_ref as IK
import oracle as O

AXES = {2: "x", 3: "y", 4: "z"}


def prims(records):
    out = np.zeros(len(records), O.PRIM)
    for i, (kind, op, c, p) in enumerate(records):
        out[i] = (kind, op, *c, *(tuple(p) + (0.0,) * (3 - len(p))))
    return out


def one(kind, c, p):
    return prims([(kind, 0, c, p)])


def cloud(rng, c, spread, n=20000):
    pts = rng.normal(0.0, spread, (n, 3)) + np.asarray(c, np.float64)
    return pts.astype(np.float32)


def test_sphere_box_cylinder_equal_reference_emitted_f32():
    rng = np.random.default_rng(11)
    for trial in range(40):
        c = tuple(float(v) for v in np.float32(rng.uniform(-50, 50, 3)))
        pts = cloud(rng, c, float(rng.uniform(1, 30)))
        r = float(np.float32(rng.uniform(0.5, 20)))
        node = {"op": "sphere", "radius": r, "center": list(c)}
        assert O.recipe_eval_points(one(0, c, (r,)), pts).tobytes() == IK.eval_f32(node, pts).tobytes(), trial
        he = [float(v) for v in np.float32(rng.uniform(0.5, 20, 3))]
        node = {"op": "box", "half_extents": he, "center": list(c)}
        assert O.recipe_eval_points(one(1, c, he), pts).tobytes() == IK.eval_f32(node, pts).tobytes(), trial
        for kind, axis in AXES.items():
            a, b = (float(v) for v in np.float32(rng.uniform(0.5, 20, 2)))
            node = {"op": "cylinder", "radius": a, "height": 2.0 * b, "axis": axis, "center": list(c)}
            got = O.recipe_eval_points(one(kind, c, (a, b)), pts)
            assert got.tobytes() == IK.eval_f32(node, pts).tobytes(), (trial, axis)


def test_wrong_cylinder_conventions_are_caught():
    """Controls: the full height as the node's half height, or another axis,
    differ from the recipe cylinder."""
    rng = np.random.default_rng(12)
    c, a, b = (1.0, 2.0, 3.0), 4.0, 6.0
    pts = cloud(rng, c, 8.0)
    got = O.recipe_eval_points(one(4, c, (a, b)), pts)
    assert got.tobytes() != IK.eval_f32({"op": "cylinder", "radius": a, "height": b, "axis": "z",
                                         "center": list(c)}, pts).tobytes()
    assert got.tobytes() != IK.eval_f32({"op": "cylinder", "radius": a, "height": 2 * b, "axis": "y",
                                         "center": list(c)}, pts).tobytes()


def _surface_cylinder(a, b, axis, step):
    """Dense samples of a capped cylinder's surface in local coordinates,
    (radial u, radial v, axial) mapped to (x, y, z) by axis."""
    nt = int(math.ceil(2 * math.pi * a / step))
    t = 2 * math.pi * np.arange(nt) / nt
    h = np.linspace(-b, b, int(math.ceil(2 * b / step)) + 1)
    T, H = np.meshgrid(t, h, indexing="ij")
    side = np.stack([a * np.cos(T).ravel(), a * np.sin(T).ravel(), H.ravel()], -1)
    g = np.arange(-a, a + step / 2, step)
    U, V = np.meshgrid(g, g, indexing="ij")
    disk = np.stack([U.ravel(), V.ravel()], -1)
    disk = disk[(disk ** 2).sum(-1) <= a * a]
    rim = np.stack([a * np.cos(t), a * np.sin(t)], -1)
    disk = np.concatenate([disk, rim])
    caps = np.concatenate([np.column_stack([disk, np.full(len(disk), s)]) for s in (-b, b)])
    loc = np.concatenate([side, caps])
    order = {"z": (0, 1, 2), "y": (0, 2, 1), "x": (2, 0, 1)}[axis]
    out = np.empty_like(loc)
    for dst, src in enumerate(order):
        out[:, dst] = loc[:, src]
    return out


def _surface_torus(R, r, step):
    nu = int(math.ceil(2 * math.pi * (R + r) / step))
    nv = int(math.ceil(2 * math.pi * r / step))
    u = 2 * math.pi * np.arange(nu) / nu
    v = 2 * math.pi * np.arange(nv) / nv
    U, V = np.meshgrid(u, v, indexing="ij")
    return np.stack([((R + r * np.cos(V)) * np.cos(U)).ravel(), ((R + r * np.cos(V)) * np.sin(U)).ravel(),
                     (r * np.sin(V)).ravel()], -1)


def _brute(surface, inside, pts):
    from scipy.spatial import cKDTree
    d, _ = cKDTree(surface).query(pts)
    return np.where(inside, -d, d)


@pytest.mark.parametrize("kind", [2, 3, 4])
def test_capped_cylinders_against_brute_force(kind):
    rng = np.random.default_rng(20 + kind)
    axis = AXES[kind]
    a, b = 5.0, 8.0
    step = 0.02
    c = (3.0, -2.0, 1.0)
    surf = _surface_cylinder(a, b, axis, step) + np.asarray(c)
    pts = cloud(rng, c, 7.0, 4000).astype(np.float64)
    loc = pts - np.asarray(c)
    ax = {"x": 0, "y": 1, "z": 2}[axis]
    radial = np.sqrt((loc ** 2).sum(-1) - loc[:, ax] ** 2)
    inside = (radial < a) & (np.abs(loc[:, ax]) < b)
    want = _brute(surf, inside, pts)
    got = O.recipe_eval_points(one(kind, c, (a, b)), pts.astype(np.float32)).astype(np.float64)
    assert np.abs(got - want).max() < 1.5 * step
    # a wrong axis or a full height read as a half height is far off
    other = {2: 3, 3: 4, 4: 2}[kind]
    assert np.abs(O.recipe_eval_points(one(other, c, (a, b)), pts.astype(np.float32)) - want).max() > 1.0
    assert np.abs(O.recipe_eval_points(one(kind, c, (a, b / 2)), pts.astype(np.float32)) - want).max() > 1.0


def test_torus_against_brute_force():
    rng = np.random.default_rng(30)
    R, r = 9.0, 3.0
    c = (-4.0, 5.0, 2.0)
    surf = _surface_torus(R, r, 0.03) + np.asarray(c)
    pts = cloud(rng, c, 8.0, 4000).astype(np.float64)
    loc = pts - np.asarray(c)
    ring = np.hypot(loc[:, 0], loc[:, 1]) - R
    inside = ring ** 2 + loc[:, 2] ** 2 < r * r
    want = _brute(surf, inside, pts)
    got = O.recipe_eval_points(one(5, c, (R, r)), pts.astype(np.float32)).astype(np.float64)
    assert np.abs(got - want).max() < 0.05
    assert np.abs(O.recipe_eval_points(one(5, c, (r, R)), pts.astype(np.float32)) - want).max() > 1.0


@pytest.mark.parametrize("pitch", [1.0, 0.5])
def test_sphere_voxel_volume(pitch):
    r = 20.0
    n = int(round(46 / pitch))
    centre = (23.0 - 0.5 * pitch,) * 3  # a cell centre of the sample lattice
    g = O.grid_record(np.float32(0), np.float32(0), np.float32(0), np.float32(pitch), n, n, n, 8,
                      np.float32(math.sqrt(3.0) * 8 * pitch))
    kind, cma, ids, tiles = O.build_sparse(one(0, centre, (r,)), g)
    dense = O.reconstruct(kind, ids, tiles, n, n, n, 8)
    vol = float(np.count_nonzero(dense < 0)) * pitch ** 3
    exact = 4.0 / 3.0 * math.pi * r ** 3
    assert abs(vol - exact) / exact < 2e-3, (pitch, vol, exact)


def test_fold_order_and_ties():
    """min for union, max(f, -p) for subtraction, max for intersection; equal
    operands return the second one (+0 vs -0 keeps the sign of the later)."""
    pts = np.array([[0.0, 0.0, 0.0], [3.0, 0.0, 0.0], [10.0, 0.0, 0.0]], np.float32)
    s1 = ((0, 0, 0), (5.0,))
    s2 = ((3, 0, 0), (2.0,))
    a = O.recipe_eval_points(one(0, *s1), pts)
    b = O.recipe_eval_points(one(0, *s2), pts)
    u = O.recipe_eval_points(prims([(0, 0, *s1), (0, 0, *s2)]), pts)
    d = O.recipe_eval_points(prims([(0, 0, *s1), (0, 1, *s2)]), pts)
    i = O.recipe_eval_points(prims([(0, 0, *s1), (0, 2, *s2)]), pts)
    assert np.array_equal(u, np.minimum(a, b)) and np.array_equal(d, np.maximum(a, -b))
    assert np.array_equal(i, np.maximum(a, b))
    # at (5,0,0) both spheres are +0; the subtraction compares +0 with -(+0) =
    # -0, which are equal, so the second operand (-0) is the result
    z = np.array([[5.0, 0.0, 0.0]], np.float32)
    tie = O.recipe_eval_points(prims([(0, 0, (0, 0, 0), (5.0,)), (0, 1, (5, 0, 0), (0.0,))]), z)
    assert tie[0] == 0 and np.signbit(tie[0])
