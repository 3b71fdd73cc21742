"""IKARUS trees on the CPU: the fp32 oracle and the float64 ground truth
against closed forms and the reference's own gates, the compiler's program
shape and stack bound, the JSON format.  No GPU.

What pins the oracle (oracle/ikarus_ref.py):
* eval_f64 (FE/ikarus_v1/expr.py:263-385): closed forms -- exact distances of a
  sphere, an axis-aligned box (brute-force nearest point), a halfspace and a
  capped cylinder on its axis; transforms and rotations equal the moved
  primitive; scale(sphere r, s) = sphere(r s); offset; hard booleans = min/max;
  a smooth union lies between min - k/4 and min and equals min where the
  operands differ by >= k; rect_frustum with equal profiles is the box's
  inside distance; a grid leaf of an affine field reproduces the field;
* eval_f32 (the generated fp32 arithmetic): within fp32 rounding of eval_f64
  on random trees of every node type, and inside the reference's own
  acceptance gates on its self-test trees (max error < 1e-3 for
  FE/ikarus_v1/eval_warp.py:737-770, < 1e-2 for eval_warp_uniform_v2.py:57-109).
"""
import json
import math

import numpy as np
import pytest

import ikarus_ref as IR
from ikarus_trees import balanced, chain, linear_grid, random_tree, selftest_v1, selftest_v2


@pytest.fixture(scope="module")
def E(fe):
    return fe.ikarus


def test_f64_closed_forms(E):
    rng = np.random.default_rng(0)
    p = rng.uniform(-10, 10, (2000, 3))
    c = np.array([1.0, -2.0, 0.5])
    d = np.linalg.norm(p - c, axis=1)
    assert np.allclose(IR.eval_f64(E.sphere(3.0, c), p), d - 3.0, atol=1e-12)
    # box: distance to the nearest point of the box (outside), -distance to the faces (inside)
    h = np.array([2.0, 3.0, 1.5])
    q = np.abs(p - c) - h
    nearest = np.clip(p, c - h, c + h)
    outside = np.linalg.norm(p - nearest, axis=1)
    inside = np.max(q, axis=1)
    want = np.where(np.all(q <= 0, axis=1), inside, outside)
    assert np.allclose(IR.eval_f64(E.box(h, c), p), want, atol=1e-12)
    n = np.array([1.0, 2.0, -2.0]) / 3.0
    assert np.allclose(IR.eval_f64(E.halfspace([1.0, 2.0, -2.0], 0.7), p), p @ n - 0.7, atol=1e-12)
    # cylinder on its axis: |z| - h/2 beyond the caps, max(-r, |z| - h/2) inside
    z = np.linspace(-12, 12, 101)
    on_axis = np.stack([np.full_like(z, c[0]), np.full_like(z, c[1]), c[2] + z], 1)
    v = IR.eval_f64(E.cylinder(2.0, 8.0, "z", c), on_axis)
    assert np.allclose(v, np.maximum(-2.0, np.abs(z) - 4.0), atol=1e-12)
    # moved / rotated / scaled primitives
    s = E.sphere(2.5, (0.0, 0.0, 0.0))
    assert np.allclose(IR.eval_f64(E.translate(s, *c), p), d - 2.5, atol=1e-12)
    m = E.mat4_mul(E.mat4_translate(*c), E.mat4_rot_axis("y", 33.0))
    assert np.allclose(IR.eval_f64(E.transform(s, m), p), d - 2.5, atol=1e-12)
    rb = E.rotate(E.box(h, (0.0, 0.0, 0.0)), "z", 90.0)
    assert np.allclose(IR.eval_f64(rb, p), IR.eval_f64(E.box([h[1], h[0], h[2]], (0.0, 0.0, 0.0)), p), atol=1e-12)
    assert np.allclose(IR.eval_f64(E.scale(s, 2.0), p), np.linalg.norm(p, axis=1) - 5.0, atol=1e-12)
    assert np.allclose(IR.eval_f64(E.offset(s, 0.5), p), np.linalg.norm(p, axis=1) - 3.0, atol=1e-12)


def test_f64_booleans_and_frustum(E):
    rng = np.random.default_rng(1)
    p = rng.uniform(-8, 8, (3000, 3))
    a, b = E.sphere(4.0, (1.0, 0.0, 0.0)), E.box((2.0, 3.0, 2.5), (-1.0, 0.5, 0.0))
    da, db = IR.eval_f64(a, p), IR.eval_f64(b, p)
    assert np.array_equal(IR.eval_f64(E.union(a, b), p), np.minimum(da, db))
    assert np.array_equal(IR.eval_f64(E.intersect(a, b), p), np.maximum(da, db))
    assert np.array_equal(IR.eval_f64(E.subtract(a, b), p), np.maximum(da, -db))
    k = 1.5
    su = IR.eval_f64(E.union(a, b, k=k), p)
    assert np.all(su <= np.minimum(da, db) + 1e-12) and np.all(su >= np.minimum(da, db) - k / 4 - 1e-12)
    far = np.abs(da - db) >= k
    assert np.array_equal(su[far], np.minimum(da, db)[far])
    # equal profiles: the frustum is a box of half height (z1 - z0)/2; inside, its planes are the box's faces
    rf = E.rect_frustum(-2.0, 4.0, 3.0, 2.0, 3.0, 2.0)
    bx = E.box((3.0, 2.0, 3.0), (0.0, 0.0, 1.0))
    inner = rng.uniform([-3, -2, -2], [3, 2, 4], (500, 3))
    assert np.allclose(IR.eval_f64(rf, inner), IR.eval_f64(bx, inner), atol=1e-12)
    # axis_project freezes one coordinate before the child
    ap = E.axis_project(a, 2, 0.75)
    frozen = p.copy()
    frozen[:, 2] = 0.75
    assert np.array_equal(IR.eval_f64(ap, p), IR.eval_f64(a, frozen))


def test_f32_grid_leaf_reproduces_an_affine_field(E, tmp_path):
    g = linear_grid()
    origin, pitch = (-6.0, -5.0, -4.5), 1.5
    rng = np.random.default_rng(2)
    p = rng.uniform([-6, -5, -4.5], [6, 4, 6], (4000, 3)).astype(np.float32).astype(np.float64)
    want = -1.0 + 0.5 * p[:, 0] - 0.25 * p[:, 1] + 0.75 * p[:, 2]
    leaf = E.grid_leaf(g, origin, pitch, mode="sdf")
    got = IR.eval_f32(leaf, p)
    assert np.max(np.abs(got - want)) < 2e-5
    far = np.array([[100.0, 0.0, 0.0], [-7.0, 0.0, 0.0]])
    assert np.all(IR.eval_f32(leaf, far) == np.float32(1.0e6))
    # the .npz form (FE/ikarus_v1/eval_warp.py:433-450): occupancy mode reads (iso - v) * pitch
    path = tmp_path / "occ.npz"
    np.savez(path, occupancy=g, global_origin_mm=np.array(origin), pitch_mm=np.float64(pitch))
    occ = E.gridfield(str(path), key="occupancy", mode="occupancy", iso=0.5)
    got = IR.eval_f32(occ, p)
    assert np.max(np.abs(got - (0.5 - want) * pitch)) < 5e-5


def test_f32_within_rounding_of_f64(E):
    rng = np.random.default_rng(5)
    for trial in range(40):
        tree = random_tree(E, rng, 4)
        p = rng.uniform(-10, 10, (1500, 3)).astype(np.float32).astype(np.float64)
        v32 = IR.eval_f32(tree, p).astype(np.float64)
        v64 = IR.eval_f64(tree, p)
        scale = 1.0 + np.abs(v64) + np.abs(p).max()
        assert np.all(np.abs(v32 - v64) <= 2e-5 * scale), trial


def test_reference_selftest_gates(E):
    """The reference's acceptance gates on its own trees and points."""
    tree, pts = selftest_v1(E)
    err = np.abs(IR.eval_f32(tree, pts).astype(np.float64) - IR.eval_f64(tree, pts))
    assert err.max() < 1e-3
    tree, pts = selftest_v2(E)
    err = np.abs(IR.eval_f32(tree, pts).astype(np.float64) - IR.eval_f64(tree, pts))
    assert err.max() < 1e-2


def test_compiler_program_shape(E):
    tree, _ = selftest_v1(E)
    prog = E.compile_tree(tree)
    ops = prog.ins[:, 0].tolist()
    assert ops.count(E.SPHERE) == 1 and ops.count(E.BOX) == 1 and ops.count(E.CYL) == 1
    assert ops.count(E.UNION) == 1 and ops.count(E.SUBTRACT) == 1 and ops.count(E.XFORM) == 1
    assert prog.params.dtype == np.float32 and prog.depth == 2
    # a smooth k of 2.0 and the cylinder's half height 20.0, as fp32
    assert np.float32(2.0) in prog.params and np.float32(20.0) in prog.params
    for right in (True, False):
        p = E.compile_tree(chain(E, 200, right))
        assert p.depth == 2
    assert E.compile_tree(balanced(E, 8)).depth == 9
    assert E.compile_tree(balanced(E, 16)).depth == 17
    with pytest.raises(ValueError):
        E.compile_tree(balanced(E, 17))
    with pytest.raises(ValueError):
        E.compile_tree({"op": "torus"})


def raw(*ins):
    return np.array([(op, a, pofs, 0) for op, a, pofs in ins], np.int32).reshape(-1, 4)


def test_host_program_check(E, fe):
    """fe_ikarus_check accepts every compiled program and rejects each kind of
    malformed one, naming the instruction; the stack bound is the
    interpreter's 16 LDS slots plus the register top."""
    L = fe._lib.lib()

    def check(ins, n_params=64, n_grids=1):
        status = L.fe_ikarus_check(ins.ctypes.data, len(ins), n_params, n_grids)
        return status, L.fe_last_error().decode()

    for tree in (selftest_v1(E)[0], balanced(E, 16), chain(E, 200, True), random_tree(E, np.random.default_rng(3), 6)):
        p = E.compile_tree(tree)
        assert check(p.ins, len(p.params), len(p.grids))[0] == 0
    S, U = E.SPHERE, E.UNION
    two = raw((S, 0, 0), (S, 0, 0), (U, 0, 4))
    assert check(two)[0] == 0
    # deepest legal stack: 17 leaves, then 16 unions
    deep = raw(*([(S, 0, 0)] * 17 + [(U, 0, 4)] * 16))
    assert check(deep)[0] == 0
    cases = {
        "value stack deeper than 17": raw(*([(S, 0, 0)] * 18 + [(U, 0, 4)] * 17)),
        "binary operation without two values": raw((S, 0, 0), (U, 0, 4)),
        "unknown opcode": raw((S, 0, 0), (99, 0, 0)),
        "parameters out of range": raw((S, 0, 62)),
        "grid leaf out of range": raw((E.GRID, 1, 0)),
        "axis out of range": raw((E.CYL, 3, 0)),
        "bad order flag": raw((S, 0, 0), (S, 0, 0), (U, 2, 4)),
        "unary operation without a value": raw((E.OFFSET, 0, 0), (S, 0, 0)),
        "does not end with exactly one value": raw((S, 0, 0), (S, 0, 0)),
    }
    for what, ins in cases.items():
        status, msg = check(ins)
        assert status == 1 and what in msg, (what, msg)
    assert check(raw((S, 0, -1)))[0] == 1
    assert L.fe_ikarus_check(None, 1, 4, 0) == 1 and check(np.zeros((0, 4), np.int32))[0] == 1


def test_json_round_trip(E, tmp_path):
    rng = np.random.default_rng(9)
    tree = random_tree(E, rng, 5)
    path = E.save(tree, tmp_path / "t.json")
    assert E.load(path) == json.loads(json.dumps(tree))
    assert E.node_count(tree) >= 1
    assert E.leaf_ops(E.union(E.sphere(1), E.cone(1, 2, 3))) == {"sphere", "cone"}
    (tmp_path / "bare.json").write_text(json.dumps(tree))
    assert E.load(tmp_path / "bare.json") == json.loads(json.dumps(tree))
    m = E.mat4_mul(E.mat4_translate(1, 2, 3), E.mat4_rot_axis("x", 30))
    inv = E.mat4_inverse_affine(m)
    assert np.allclose(np.array(E.mat4_mul(m, inv)), np.eye(4), atol=1e-14)
    assert inv == IR.inverse_affine(m)


def test_query_helpers(fe, E):
    Q = fe.ikarus_queries
    lo, hi = Q.bbox_of(E.union(E.sphere(2.0, (1, 2, 3)), E.translate(E.box((1, 2, 3)), 10, 0, 0)), pad=0.0)
    assert np.allclose(lo, [-1, -2, -3]) and np.allclose(hi, [11, 4, 5])
    lo, hi = Q.bbox_of(E.rotate(E.box((2, 1, 1)), "z", 90.0), pad=0.5)
    assert np.allclose(lo, [-1.5, -2.5, -1.5]) and np.allclose(hi, [1.5, 2.5, 1.5])
    assert Q.grid_shape([0, 0, 0], [1, 2, 3], 0.5) == (3, 5, 7)
    # a cylinder's box follows its own axis; a rotated sphere keeps its radius
    lo, hi = Q.bbox_of(E.cylinder(2.0, 10.0, "x", (1, 0, 0)), pad=0.0)
    assert np.allclose(lo, [-4, -2, -2]) and np.allclose(hi, [6, 2, 2])
    lo, hi = Q.bbox_of(E.rotate(E.sphere(3.0, (5, 0, 0)), "z", 90.0), pad=0.0)
    assert np.allclose(lo, [-3, 2, -3]) and np.allclose(hi, [3, 8, 3])
    with pytest.raises(ValueError):
        Q.bbox_of(E.halfspace((0, 0, 1), 0.0))


def test_oracle_grid_volume_of_box_overlap(E):
    """Two axis-aligned boxes: the counting restatement against the exact overlap."""
    a = E.box((3.0, 2.0, 2.5), (0.0, 0.0, 0.0))
    b = E.box((2.0, 3.0, 1.5), (1.5, 0.5, 1.0))
    exact = (3.0 - (-0.5)) * (2.0 - (-2.0)) * (2.5 - (-0.5))
    lo, hi = np.array([-0.5, -2.0, -0.5]) - 1, np.array([3.0, 2.0, 2.5]) + 1
    r = IR.grid_volume(E.intersect(a, b), lo, hi, 0.4, 4)
    assert abs(r["volume_mm3"] - exact) < 0.05 * exact
    assert r["n_boundary"] > 0 and r["n_fine_inside"] > 0


def test_oracle_pairs_and_band_pinned(E):
    """The pair and band-tree restatements against the tree-walking
    oracle: a pair's value is intersect(a, b) (FE/ikarus_v1/eval_warp.py:
    504-560); the band plan as data equals the literal tree bit for bit, as
    the reference states for its kernel (FE/ikarus_v1/eval_warp.py:582-594),
    and that tree has 7*n_bands - 1 nodes (:575)."""
    from ikarus_trees import leaf_set
    leaves = leaf_set(E)
    rng = np.random.default_rng(31)
    pts = rng.uniform([-8, -6, -6], [12, 6, 6], (3000, 3))
    la, lb = [0, 1, 2, 3, 0], [1, 2, 0, 0, 4]
    pop = rng.integers(0, len(la), len(pts))
    got = IR.pairs_f32(pts, pop, la, lb, leaves)
    for k in range(len(la)):
        sel = pop == k
        want = IR.eval_f32(E.intersect(leaves[la[k]], leaves[lb[k]]), pts[sel])
        assert got[sel].tobytes() == want.tobytes(), k
    bands = [{"zs": float(z), "z0": float(z - 0.6), "z1": float(z + 0.4)} for z in np.linspace(-4.0, 4.5, 9)]
    for axis in range(3):
        tree = E.band_tree_expr(leaves[0], axis, bands)
        assert E.node_count(tree) == 7 * len(bands) - 1
        assert IR.band_f32(leaves[0], axis, bands, pts).tobytes() == IR.eval_f32(tree, pts).tobytes(), axis


def test_leaf_keys(fe, E):
    from ikarus_trees import leaf_set
    leaves = leaf_set(E)
    assert E._leaf_key(leaves[0]) != E._leaf_key(leaves[1])
    with pytest.raises(ValueError):
        E._leaf_key(E.sphere(1.0))
