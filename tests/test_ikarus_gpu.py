"""IKARUS trees on the GPU: the device interpreter bit-exact against the fp32
oracle (oracle/ikarus_ref.py, which walks the tree itself, not the compiled
program) on the reference's self-test trees, seeded random trees over every
node type (grid leaves included), deep chains and balanced trees at the stack
bound; the reference's acceptance gates against the float64 ground truth;
gradients; and block-sparse fields built from trees, bit-exact (kinds, corner
minima, active ids, tiles) against the oracle's classification and tiles."""
import math

import numpy as np
import pytest

import ikarus_ref as IR
from ikarus_trees import balanced, chain, linear_grid, random_tree, selftest_v1, selftest_v2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E(fe):
    return fe.ikarus


def same_bits(a, b):
    return np.asarray(a, np.float32).tobytes() == np.asarray(b, np.float32).tobytes()


def test_reference_trees(E, gpu):
    for make, gate in ((selftest_v1, 1e-3), (selftest_v2, 1e-2)):
        tree, pts = make(E)
        got = E.eval_batch(tree, pts)
        assert same_bits(got, IR.eval_f32(tree, pts)), make.__name__
        assert np.max(np.abs(got.astype(np.float64) - IR.eval_f64(tree, pts))) < gate


def test_random_trees_match_oracle(E, gpu):
    rng = np.random.default_rng(11)
    leaf = E.grid_leaf(linear_grid(), (-6.0, -5.0, -4.5), 1.5, mode="sdf")
    occ = E.grid_leaf(linear_grid(coef=(0.1, 0.2, -0.1), c0=0.4), (-4.0, -3.0, -2.0), 0.75, mode="occupancy",
                      iso=0.5)
    for trial in range(60):
        tree = random_tree(E, rng, int(rng.integers(1, 6)), grid=leaf if trial % 2 else occ)
        pts = rng.uniform(-12, 12, (5000, 3))
        got = E.eval_batch(tree, pts)
        want = IR.eval_f32(tree, pts)
        if not same_bits(got, want):
            bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
            raise AssertionError((trial, bad[:5], got[bad[:5]], want[bad[:5]]))


def test_stack_bound_and_chains(E, gpu):
    rng = np.random.default_rng(12)
    pts = rng.uniform(-20, 20, (20000, 3))
    for tree in (chain(E, 300, True), chain(E, 300, False), balanced(E, 6), balanced(E, 16)):
        prog = E.compile_tree(tree)
        print("depth", prog.depth, "instructions", len(prog.ins), flush=True)
        sample = pts if len(prog.ins) < 10000 else pts[:256]
        assert same_bits(E.eval_batch(prog, sample), IR.eval_f32(tree, sample))


def test_gradient(E, gpu):
    tree, pts = selftest_v1(E)
    g = E.gradient_batch(tree, pts[:500], h=0.05)
    offsets = np.array([[0.05, 0, 0], [-0.05, 0, 0], [0, 0.05, 0], [0, -0.05, 0], [0, 0, 0.05], [0, 0, -0.05]])
    vals = IR.eval_f32(tree, (pts[:500][None] + offsets[:, None]).reshape(-1, 3)).reshape(6, 500)
    want = np.stack([(vals[0] - vals[1]) / (2 * 0.05), (vals[2] - vals[3]) / (2 * 0.05),
                     (vals[4] - vals[5]) / (2 * 0.05)], 1)
    assert g.tobytes() == want.tobytes()
    # unit length away from the blend, creases and the rotated sphere's neighbourhood
    s = E.sphere(5.0, (1.0, 2.0, 3.0))
    p = np.random.default_rng(3).normal(size=(300, 3))
    p = p / np.linalg.norm(p, axis=1)[:, None] * 8.0 + np.array([1.0, 2.0, 3.0])
    n = np.linalg.norm(E.gradient_batch(s, p), axis=1)
    assert np.allclose(n, 1.0, atol=1e-3)


def field_meta(fe, x0, y0, z0, pitch, n, B):
    return fe.GridMeta(x0, y0, z0, pitch, n[0], n[1], n[2], block=B)


def test_field_build_matches_oracle(fe, E, gpu):
    rng = np.random.default_rng(13)
    leaf = E.grid_leaf(linear_grid(), (-6.0, -5.0, -4.5), 1.5, mode="sdf")
    trees = [selftest_v1(E)[0], selftest_v2(E)[0], random_tree(E, rng, 4, grid=leaf), random_tree(E, rng, 5)]
    grids = [(-30.0, -30.0, -30.0, 0.5, (120, 120, 120), 8), (-30.0, -30.0, -30.0, 0.75, (130, 81, 81), 4),
             (-12.0, -12.0, -12.0, 0.25, (96, 100, 90), 16), (-12.5, -11.0, -10.0, 0.3, (80, 70, 75), 8)]
    for i, (tree, (x0, y0, z0, h, n, B)) in enumerate(zip(trees, grids)):
        meta = field_meta(fe, x0, y0, z0, h, n, B)
        sf = E.klassificera_och_evaluera_ikarus(meta, tree)
        om = dict(x0=meta.x0, y0=meta.y0, z0=meta.z0, pitch=meta.pitch, margin=meta.margin_mm, B=B,
                  nbx=meta.n_bx, nby=meta.n_by, nbz=meta.n_bz)
        kind, cma = IR.classify_f32(tree, om)
        print("field", i, meta.n_blocks, int((kind == 0).sum()), flush=True)
        assert np.array_equal(sf.kind, kind), i
        assert same_bits(sf.corner_min_abs, cma), i
        ids = np.flatnonzero(kind == 0)
        assert np.array_equal(sf.active_ids, ids), i
        assert same_bits(sf.tiles, IR.tiles_f32(tree, om, ids)), i
        assert 0 < len(ids) < meta.n_blocks
        # the existing reconstruction serves tree-built fields
        dense = fe.rekonstruera_tat_gpu(sf)
        assert dense.shape == tuple(n)


def test_grid_volume_counts_match_oracle(fe, E, gpu):
    Q = fe.ikarus_queries
    rng = np.random.default_rng(14)
    trees = [selftest_v1(E)[0], E.intersect(E.sphere(6.0, (0.5, 0.2, 0.1)), E.box((4, 5, 3), (1, 0, 0))),
             random_tree(E, rng, 4)]
    for tree in trees:
        lo, hi = np.array([-14.0, -13.0, -12.5]), np.array([13.0, 12.0, 14.5])
        for pitch, r in ((1.0, 4), (0.7, 3), (1.3, 1)):
            got = Q.grid_volume(tree, lo, hi, pitch, r)
            want = IR.grid_volume(tree, lo, hi, pitch, r)
            for k in ("n_fully_inside", "n_boundary", "n_fine_inside", "n_fine_boundary"):
                assert got[k] == want[k], (k, pitch, r)
            assert got["volume_mm3"] == want["volume_mm3"]


def test_intrusion_volume_closed_forms(fe, E, gpu):
    Q = fe.ikarus_queries
    r1, r2, d = 6.0, 4.0, 7.0
    lens = math.pi * (r1 + r2 - d) ** 2 * (d * d + 2 * d * r2 - 3 * r2 * r2 + 2 * d * r1 + 6 * r1 * r2
                                           - 3 * r1 * r1) / (12 * d)
    res = Q.intrusion_volume(E.sphere(r1, (0, 0, 0)), E.sphere(r2, (d, 0, 0)), max_points=2_000_000)
    assert abs(res["volume_mm3"] - lens) <= res["uncertainty_mm3"] + 0.02 * lens, (res["volume_mm3"], lens)
    assert res["refine_factor"] >= 2 and res["n_points"] <= 2_000_000
    # the passes it ran are grid_volume's at the pitch and factor it chose
    lo, hi = (np.maximum(x, y) for x, y in zip(Q.bbox_of(E.sphere(r1, (0, 0, 0))), Q.bbox_of(E.sphere(r2, (d, 0, 0)))))
    want = IR.grid_volume(E.intersect(E.sphere(r1, (0, 0, 0)), E.sphere(r2, (d, 0, 0))), lo, hi,
                          res["coarse_pitch_mm"], res["refine_factor"])
    for k in ("n_fully_inside", "n_boundary", "n_fine_inside", "n_fine_boundary"):
        assert res[k] == want[k], k
    a = E.box((3.0, 2.0, 2.5), (0.0, 0.0, 0.0))
    b = E.box((2.0, 3.0, 1.5), (1.5, 0.5, 1.0))
    exact = 3.5 * 4.0 * 3.0
    res = Q.intrusion_volume(a, b, max_points=500_000)
    assert abs(res["volume_mm3"] - exact) <= res["uncertainty_mm3"] + 0.02 * exact
    far = Q.intrusion_volume(E.sphere(1.0, (0, 0, 0)), E.sphere(1.0, (50, 0, 0)))
    assert far["volume_mm3"] == 0.0 and far["uncertainty_mm3"] == 0.0


def test_multi_pair_matches_oracle(E, gpu):
    from ikarus_trees import leaf_set
    leaves = leaf_set(E)
    rng = np.random.default_rng(15)
    pts = rng.uniform([-8, -6, -6], [42, 6, 6], (50000, 3))
    la, lb = np.array([0, 1, 2, 3, 0, 4]), np.array([1, 2, 0, 0, 4, 4])
    pop = np.sort(rng.integers(0, len(la), len(pts)))
    got = E.eval_batch_multi_pair(pts, pop, la, lb, leaves)
    assert same_bits(got, IR.pairs_f32(pts, pop, la, lb, leaves))
    # a shuffled pair order (pairs not contiguous) and a second call on the cached leaf table
    perm = rng.permutation(len(pts))
    got2 = E.eval_batch_multi_pair(pts[perm], pop[perm], la, lb, leaves)
    assert same_bits(got2, got[perm])
    assert E.eval_batch_multi_pair(np.zeros((0, 3)), np.zeros(0, np.int32), la, lb, leaves).shape == (0,)
    with pytest.raises(ValueError):
        E.eval_batch_multi_pair(pts[:4], [0, 1, 2, 6], la, lb, leaves)
    with pytest.raises(ValueError):
        E.eval_batch_multi_pair(pts[:4], [0, 0, 0, 0], [5], [0], leaves)


def test_band_tree_evaluator(E, gpu):
    from ikarus_trees import leaf_set
    leaves = leaf_set(E)
    rng = np.random.default_rng(16)
    pts = rng.uniform(-8, 8, (40000, 3))
    for axis in range(3):
        ev = E.BandTreeEvaluator(leaves[1 if axis else 0], axis)
        for nb in (1, 5, 28):
            zs = np.sort(rng.uniform(-6, 6, nb))
            bands = [{"zs": float(z), "z0": float(z - rng.uniform(0.1, 1.0)), "z1": float(z + rng.uniform(0.1, 1.0))}
                     for z in zs]
            got = ev.eval(pts, bands)
            print("band tree", axis, nb, flush=True)
            assert same_bits(got, IR.band_f32(ev.leaf, axis, bands, pts)), (axis, nb)
            # the plan as data equals the interpreter on the literal tree
            assert same_bits(got, E.eval_batch(E.band_tree_expr(ev.leaf, axis, bands), pts)), (axis, nb)
        assert np.all(ev.eval(pts[:10], []) == np.float32(1.0e30))
    with pytest.raises(ValueError):
        E.BandTreeEvaluator(leaves[0], 3)


def test_malformed_programs_stop_and_latch(E, fe, gpu):
    """Raw programs past the host check: stack underflow, a push past the 16
    LDS slots and an unknown opcode evaluate to NaN and latch the fault word,
    which fe_ikarus_fault reports once; a well-formed program after them is
    unaffected and leaves no fault."""
    import torch
    L, P = fe._lib.lib(), fe._lib.ptr
    S, U = E.SPHERE, E.UNION
    params = torch.tensor([1.0, 2.0, 3.0, 4.0, 0.0, 0.0, 0.0, 0.0], device=gpu)
    pts = torch.rand((1000, 3), device=gpu) * 10
    out = torch.empty(1000, device=gpu)
    st = fe._lib.stream_handle(None)
    assert L.fe_ikarus_fault(st) == 0
    bad = {
        "underflow": [(S, 0, 0), (U, 0, 4)],
        "overflow": [(S, 0, 0)] * 18 + [(U, 0, 4)] * 17,
        "opcode": [(S, 0, 0), (99, 0, 0)],
    }
    for name, ins in bad.items():
        prog = np.array([(op, a, pofs, 0) for op, a, pofs in ins], np.int32)
        assert L.fe_ikarus_check(prog.ctypes.data, len(prog), 8, 0) == 1, name
        d = torch.from_numpy(prog.reshape(-1)).to(gpu)
        print("malformed program:", name, flush=True)
        assert L.fe_ikarus_eval(P(d), len(prog), P(params), None, None, P(pts), 1000, P(out), st) == 0
        assert torch.isnan(out).all().item(), name
        assert L.fe_ikarus_fault(st) == 1 and "value stack" in L.fe_last_error().decode(), name
        assert L.fe_ikarus_fault(st) == 0, name
    good = np.array([(S, 0, 0, 0), (S, 0, 0, 0), (U, 0, 4, 0)], np.int32)
    d = torch.from_numpy(good.reshape(-1)).to(gpu)
    assert L.fe_ikarus_eval(P(d), 3, P(params), None, None, P(pts), 1000, P(out), st) == 0
    ref = torch.linalg.norm(pts - params[:3], dim=1) - 4.0
    assert torch.isfinite(out).all().item() and torch.allclose(out, ref, atol=1e-5)
    assert L.fe_ikarus_fault(st) == 0
