"""GPU parity of the block-sparse field build (classify, compaction, tiles,
reconstruction, checksum) against the oracle and the reference's recorded answers."""
import math

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
HX, HY, HZ, R = 20.0, 15.0, 10.0, 6.0


def make_meta(fe, pitch, block=8):
    """FE/faltkarna_v1_g7_fix.py:42-48."""
    pad = 1.0
    nx = int(round((2 * HX + 2 * pad) / pitch)) + 1
    ny = int(round((2 * HY + 2 * pad) / pitch)) + 1
    nz = int(round((2 * HZ + 2 * pad) / pitch)) + 1
    return fe.GridMeta(x0=-HX - pad, y0=-HY - pad, z0=-HZ - pad, pitch=pitch, nx=nx, ny=ny, nz=nz, block=block)


def oracle_build(meta, recipe, corrupt=False):
    g = O.grid_record(np.float32(meta.x0), np.float32(meta.y0), np.float32(meta.z0), np.float32(meta.pitch),
                      meta.nx, meta.ny, meta.nz, meta.block, np.float32(meta.margin_mm))
    return O.build_sparse(np.ascontiguousarray(recipe).view(O.PRIM), g, corrupt)


def assert_same(sf, ref, what):
    kind, cma, ids, tiles = ref
    assert np.array_equal(sf.kind, kind), what
    assert sf.corner_min_abs.tobytes() == cma.tobytes(), what
    assert np.array_equal(sf.active_ids, ids), what
    assert sf.tiles.tobytes() == tiles.tobytes(), what


def test_testfield_known_answers(fe, golden, gpu):
    tf = golden["testfield"]
    meta = make_meta(fe, 0.25)
    sf = fe.klassificera_och_evaluera_testfalt(meta, HX, HY, HZ, R)
    s = sf.stats
    assert (s["n_blocks"], s["n_active"], s["n_interior"], s["n_exterior"]) == \
        (tf["n_blocks"], tf["n_active"], tf["n_interior"], tf["n_exterior"])
    assert s["active_voxels"] == tf["active_voxels"] and s["dense_voxels"] == tf["dense_voxels"]
    dense, t, stats = fe.klassificera_evaluera_rekonstruera_gpu_resident(meta, HX, HY, HZ, R)
    assert stats["n_active"] == tf["n_active"]
    assert float(np.sum(dense < 0.0)) * 0.25 ** 3 == tf["vol_voxel_mm3"]
    assert fe.field.checksum_tiles(sf.tiles) == round(tf["checksum_i64"] * 1e6)
    bad, _, _ = fe.klassificera_evaluera_rekonstruera_gpu_resident(meta, HX, HY, HZ, R, corrupt_index=True)
    assert float(np.sum(bad < 0.0)) * 0.25 ** 3 == tf["vol_corrupt_mm3"]
    for pitch, key in ((0.9, "n_active_pitch_0_9"), (0.09, "n_active_pitch_0_09")):
        _, _, st = fe.klassificera_evaluera_rekonstruera_gpu_resident(make_meta(fe, pitch), HX, HY, HZ, R)
        assert st["n_active"] == tf[key], pitch


def test_testfield_bit_exact_vs_oracle(fe, gpu):
    meta = make_meta(fe, 0.25)
    rec = fe.recipes.testfalt(HX, HY, HZ, R)
    ref = oracle_build(meta, rec)
    assert_same(fe.klassificera_och_evaluera_testfalt(meta, HX, HY, HZ, R), ref, "plain")
    plan = fe.FieldPlan([(meta, rec)])
    plan.launch()
    assert_same(plan.sparse_field(0), ref, "resident")
    dense = plan.reconstruct(0).cpu().numpy()
    assert dense.tobytes() == O.reconstruct(ref[0], ref[2], ref[3], meta.nx, meta.ny, meta.nz, 8).tobytes()
    assert np.array_equal(fe.rekonstruera_tat(plan.sparse_field(0)), dense)
    occ = fe.reconstruct_solid_from_sparse(plan.sparse_field(0), meta.nx, meta.ny, meta.nz, 8)
    assert np.array_equal(occ, dense < 0)


def random_cases(fe):
    rng = np.random.default_rng(8)
    for B in (4, 8, 16):
        for n_prims in (1, 7, 300):
            dims = tuple(int(x) for x in rng.integers(20, 90, 3))
            pitch = float(rng.choice([0.5, 0.7, 1.0, 1.3]))
            origin = tuple(float(x) for x in rng.uniform(-30, 10, 3))
            ext = np.array(dims) * pitch
            rec = fe.recipes.mixed_recipe(rng, n_prims, int(max(dims)))
            for k in ("cx", "cy", "cz"):
                rec[k] = rec[k] * np.float32(pitch) + np.float32(origin["xyz".index(k[1])])
            rec["a"] *= np.float32(pitch)
            rec["b"] *= np.float32(pitch)
            rec["c"] *= np.float32(pitch)
            rec[0]["op"] = 0
            meta = fe.GridMeta(*origin, pitch, *dims, block=B)
            yield f"B{B}_P{n_prims}_{dims}", meta, rec


def test_random_recipes_bit_exact(fe, gpu):
    for name, meta, rec in random_cases(fe):
        print("case", name, flush=True)
        ref = oracle_build(meta, rec)
        plan = fe.FieldPlan([(meta, rec)])
        plan.launch()
        assert_same(plan.sparse_field(0), ref, name)
        assert_same(fe.klassificera_och_evaluera_recept(meta, rec), ref, name + " plain")


def op_order_recipe(fe, rng, n_prims, dims, pattern):
    """Unions and subtractions of 3..15-voxel primitives in the central 40 % of
    the box, intersections with large central spheres (they trim the solid
    rather than empty it), in alternating single steps or in runs of random
    length."""
    R = fe.recipes
    if pattern == "alternate":
        ops = list(rng.choice([0, 1, 2], n_prims, p=[0.55, 0.3, 0.15]))
    else:
        ops = []
        while len(ops) < n_prims:
            ops += [int(rng.choice([0, 1, 2], p=[0.55, 0.3, 0.15]))] * int(rng.integers(1, 90))
        ops = ops[:n_prims]
    ops[0] = 0
    d = np.array(dims, float)
    recs = []
    for op in ops:
        if op == 2:
            r = float(rng.uniform(0.22, 0.3)) * float(d.max())
            recs.append((R.SPHERE, 2, tuple(d / 2 + rng.uniform(-3, 3, 3)), (r,)))
        else:
            k = int(rng.integers(0, 6))
            c = tuple(rng.uniform(0.3 * d, 0.7 * d))
            if k == R.SPHERE:
                recs.append((k, op, c, (float(rng.uniform(3, 15)),)))
            elif k == R.BOX:
                recs.append((k, op, c, tuple(rng.uniform(2, 8, 3))))
            elif k == R.TORUS:
                big = float(rng.uniform(4, 12))
                recs.append((k, op, c, (big, big * float(rng.uniform(0.2, 0.5)))))
            else:
                recs.append((k, op, c, (float(rng.uniform(2, 7)), float(rng.uniform(2, 10)))))
    return R.make(recs)


def test_random_op_orders_bit_exact(fe, gpu):
    """Culling is by runs of same-op steps: recipes whose folds alternate
    union, subtraction and intersection (single steps and runs of random
    length) against the oracle's unculled fold."""
    rng = np.random.default_rng(81)
    for B, n_prims, pattern in ((8, 300, "runs"), (4, 200, "alternate"), (16, 400, "runs"), (8, 257, "alternate")):
        dims = tuple(int(x) for x in rng.integers(110, 170, 3))
        rec = op_order_recipe(fe, rng, n_prims, dims, pattern)
        meta = fe.GridMeta(0.0, 0.0, 0.0, 1.0, *dims, block=B)
        name = f"B{B}_P{n_prims}_{pattern}_{dims}"
        print("case", name, flush=True)
        ref = oracle_build(meta, rec)
        assert 0 < len(ref[2]) < ref[0].size, name
        plan = fe.FieldPlan([(meta, rec)])
        plan.launch()
        assert_same(plan.sparse_field(0), ref, name)


def test_batched_plan_equals_single(fe, gpu):
    cases = [c for c in random_cases(fe) if c[1].block == 8]
    plan = fe.FieldPlan([(m, r) for _, m, r in cases])
    plan.launch()
    for i, (name, meta, rec) in enumerate(cases):
        assert_same(plan.sparse_field(i), oracle_build(meta, rec), name)


def test_plain_part_128_and_fp64_error(fe, gpu):
    import fp64
    rec = fe.recipes.plain_part_128(seed=1)
    meta = fe.GridMeta(0.0, 0.0, 0.0, 1.0, 128, 128, 128, block=8)
    ref = oracle_build(meta, rec)
    plan = fe.FieldPlan([(meta, rec)])
    plan.launch()
    sf = plan.sparse_field(0)
    assert_same(sf, ref, "cfg1")
    dense = plan.reconstruct(0).cpu().numpy()
    idx = np.indices(dense.shape).reshape(3, -1).T.astype(np.float64)
    exact = fp64.eval_fp64(rec.view(O.PRIM), idx).reshape(dense.shape)
    near = np.abs(dense) < 1e5
    err = np.abs(dense[near].astype(np.float64) - exact[near])
    print("cfg1 fp64 max err", err.max(), "mean", err.mean())
    assert err.max() < 1e-4


def test_dense_source_classifier(fe, gpu):
    r = 20.0
    for pitch in (1.0, 0.5):
        n = int(round(2 * (r + 4.0) / pitch)) + 1
        ax = -(r + 4.0) + pitch * np.arange(n)
        X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
        dense = (np.sqrt(X * X + Y * Y + Z * Z) - r).astype(np.float32)
        sf = fe.klassificera_och_evaluera_fran_tatt_falt(None, dense, pitch, 8)
        kind = O.classify_dense(dense, 8, np.float32(math.sqrt(3.0) * 8 * pitch))
        assert np.array_equal(sf["kind"], kind)
        ids = np.where(kind == 0)[0].astype(np.int32)
        assert sf["tiles"].tobytes() == O.tiles_dense(dense, 8, ids).tobytes()


def test_determinism_certificate(fe, gpu):
    meta = make_meta(fe, 0.5)
    sf = fe.klassificera_och_evaluera_testfalt(meta, HX, HY, HZ, R)
    cert = fe.volym_via_gpu_reduktion(sf, n_korningar=3)
    assert cert["i64_fixpunkt_safe"]["bit_identiska"]
    assert cert["i64_fixpunkt_safe"]["varden"][0] * 1e6 == pytest.approx(O.checksum_i64(sf.tiles), abs=0.5)
    # the float32 atomic sum is the order-dependent control: same total to float precision, not bits
    f32 = cert["f32_unsafe"]["varden"]
    assert len(f32) == 3 and all(v == pytest.approx(cert["i64_fixpunkt_safe"]["varden"][0], rel=1e-4) for v in f32)
