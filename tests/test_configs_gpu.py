"""GPU parity at the benchmark configurations: full-size builds and EDTs
against the oracle's digests in tests/golden/config_digests.json (written by
tests/golden/make_config_digests.py), and recipes that reach the build's list
fallbacks.

* cfg1: the 128^3 part through the plain per-call and the resident path, its
  dense field, occupancy and the int32 EDT of the occupancy;
* cfg2: the 512^3 field of 1024 primitives, its dense field, occupancy and
  paired EDT, plus its error against the FP64 closed forms;
* cfg3: the 256 uneven recipes in one autotuned, graph-captured batch (every
  replay checked), and a sample through the plain path;
* cfg5: served builds of the large-field distribution under Poisson arrivals
  at 2000 requests/s, each checked against the oracle's build of the same
  recipe alone;
* superblock lists longer than the 512-entry capacity and recipes of 2500
  primitives (the unrefined tile pass), through every launch path.
Bit-exact everywhere (kinds, corner minima, ids, tiles, dense, EDT).
"""
import hashlib
import json

import numpy as np
import pytest

import config_cases as C
from conftest import ROOT

pytestmark = pytest.mark.gpu
BUILD_KEYS = ("n_active", "kind", "cma", "ids", "tiles")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def digests():
    return json.loads((ROOT / "tests/golden/config_digests.json").read_text())


def meta_of(fe, n, B):
    return fe.GridMeta(*C.grid(n, B))


def plan_digests(plan, j):
    na = plan.n_active(j)
    return dict(n_active=na, kind=sha(plan.kind(j).cpu().numpy()), cma=sha(plan.corner_min_abs(j).cpu().numpy()),
                ids=sha(plan.active_ids(j, na).cpu().numpy()), tiles=sha(plan.tiles(j, na).cpu().numpy()))


def sparse_digests(sf):
    return dict(n_active=len(sf.active_ids), kind=sha(sf.kind), cma=sha(sf.corner_min_abs), ids=sha(sf.active_ids),
                tiles=sha(sf.tiles))


def check(got, ref, what):
    for k in BUILD_KEYS:
        assert got[k] == ref[k], (what, k)


def full_checks(fe, plan, ref, n):
    """Dense field, occupancy and EDTs of a built plan against the digests."""
    import torch
    dev = plan.device
    dense = torch.empty((n, n, n), dtype=torch.float32, device=dev)
    occ = torch.empty((n, n, n), dtype=torch.uint8, device=dev)
    plan.reconstruct(0, dense, occ)
    assert sha(dense.cpu().numpy()) == ref["dense"]
    assert sha(occ.cpu().numpy()) == ref["occupancy"]
    single = fe.edt_device(occ, planes=1)
    assert sha(single.cpu().numpy()) == ref["edt_single"]
    pair = fe.edt_device(occ, planes=2)
    assert sha(pair.cpu().numpy()) == ref["edt_pair"]
    assert sha(fe.signed_from_pair_device(pair, 1.0).cpu().numpy()) == ref["edt_signed"]
    return dense, occ


def test_cfg1_plain_part_and_occupancy_edt(fe, gpu, digests):
    ref = digests["cfg1"]
    rec = fe.recipes.plain_part_128(seed=1)
    meta = meta_of(fe, 128, 8)
    check(sparse_digests(fe.klassificera_och_evaluera_recept(meta, rec)), ref, "cfg1 plain")
    plan = fe.FieldPlan([(meta, rec)])
    plan.launch()
    check(plan_digests(plan, 0), ref, "cfg1 resident")
    _, occ = full_checks(fe, plan, ref, 128)
    # the reference-shaped host call (uncapped edt_squared_large contract)
    host = fe.NativeEDTVolume().squared_distance(occ.cpu().numpy().astype(bool))
    assert sha(host.astype(np.int32)) == ref["edt_single"]


def fp64_error(fe, rec, dense, sample=1 << 18, seed=0):
    """Max and mean |fp32 - FP64 closed form| over a seeded sample of the
    stored (non-background) voxels, and the bound it is held to."""
    import fp64
    import oracle as O
    vals = dense.reshape(-1)
    stored = np.flatnonzero(np.abs(vals) < 1e5)
    pick = np.random.default_rng(seed).choice(stored, min(sample, len(stored)), replace=False)
    n = dense.shape[0]
    pts = np.stack(np.unravel_index(pick, dense.shape), -1).astype(np.float64)
    exact = fp64.eval_fp64(np.ascontiguousarray(rec).view(O.PRIM), pts)
    err = np.abs(vals[pick].astype(np.float64) - exact)
    return float(err.max()), float(err.mean()), n


def test_cfg2_large_field(fe, gpu, digests):
    ref = digests["cfg2"]
    rec = fe.recipes.large_field_512(seed=2)
    meta = meta_of(fe, 512, 8)
    plan = fe.FieldPlan([(meta, rec)])
    plan.launch()
    check(plan_digests(plan, 0), ref, "cfg2 resident")
    dense, _ = full_checks(fe, plan, ref, 512)
    # the leaf-box pack of the 2^27-sample dense field (past the reference's 2^24 cap)
    packed = fe.AdaptiveSampleBlocks(dense, [0, 0, 0], 1.0)
    assert (len(packed.nodes), len(packed.payload)) == (ref["pack_count"], ref["pack_samples"])
    assert sha(packed.nodes) == ref["pack_nodes"] and sha(packed.payload) == ref["pack_payload"]
    del packed
    emax, emean, n = fp64_error(fe, rec, dense.cpu().numpy())
    # 1e-6 relative to the coordinate scale (near the surface the value itself
    # cancels towards 0, so its own ulp is no scale for the error)
    bound = 1e-6 * max(512.0, fe.recipes.extent(rec))
    print(f"cfg2 fp64 error max {emax:.3e} mean {emean:.3e} bound {bound:.3e}", flush=True)
    assert emax <= bound
    del plan, dense
    check(sparse_digests(fe.klassificera_och_evaluera_recept(meta, rec)), ref, "cfg2 plain")


@pytest.mark.parametrize("name", ["crowded1024_B4", "crowded2500_B4", "crowded1024_B8", "crowded2500_B8",
                                  "crowded1024_B16", "crowded2500_B16"])
def test_list_fallbacks(fe, gpu, digests, name):
    case = {c[0]: c for c in C.crowded_cases(fe.recipes)}[name]
    _, n, B, rec = case
    ref = digests["crowded"][name]
    meta = meta_of(fe, n, B)
    plan = fe.FieldPlan([(meta, rec)])
    plan.launch()
    lens = plan._slice(plan.i32, 0, "list_len").cpu().numpy()
    assert (lens == -1).any(), "no superblock list over capacity"
    check(plan_digests(plan, 0), ref, name + " resident")
    check(sparse_digests(fe.klassificera_och_evaluera_recept(meta, rec)), ref, name + " plain")
    # the server's query kernel falls back to the whole recipe where the list overflowed
    import oracle as O
    # (300,000 points: the superblock-sorted query path; 4096: one thread per point)
    pts = np.random.default_rng(9).uniform(30.0, 66.0, (300000, 3)).astype(np.float32)
    import torch
    f = fe.server.resident_fields(plan, torch.cuda.current_stream())[0]
    got = f.query_device(torch.from_numpy(pts).to(gpu)).cpu().numpy()
    assert got.tobytes() == O.recipe_eval_points(np.ascontiguousarray(rec).view(O.PRIM), pts).tobytes()
    few = f.query_device(torch.from_numpy(pts[:4096]).to(gpu)).cpu().numpy()
    assert few.tobytes() == got[:4096].tobytes()


def test_crowded_cases_in_one_plan(fe, gpu, digests):
    """All block-8 fallback recipes with two ordinary ones in one descriptor table."""
    cases = [c for c in C.crowded_cases(fe.recipes) if c[2] == 8]
    items = [(meta_of(fe, n, B), rec) for _, n, B, rec in cases]
    items.append((meta_of(fe, 128, 8), fe.recipes.plain_part_128(seed=1)))
    plan = fe.FieldPlan(items)
    plan.launch()
    for j, (name, *_rest) in enumerate(cases):
        check(plan_digests(plan, j), digests["crowded"][name], name)
    check(plan_digests(plan, len(cases)), digests["cfg1"], "cfg1 in plan")


def test_cfg3_autotuned_graph_batch(fe, gpu, digests):
    import torch
    ref = digests["cfg3"]
    items = fe.batch.batch_workload(seed=3, count=256)
    b = fe.batch.BatchBuilder(items)
    print("cfg3 choice", sorted(set(b.choice)), flush=True)
    b.run()
    torch.cuda.synchronize()
    first = []
    for i in range(len(items)):
        plan, j = b.plan_of(i)
        B = plan.B
        check(plan_digests(plan, j), ref[i][f"B{B}"], f"cfg3 recipe {i} B{B}")
        first.append(plan.tiles(j).clone())
    lens = [int((p._slice(p.i32, j, "list_len") < 0).sum().item()) for p in b.plans for j in range(len(p))]
    print("cfg3 superblocks with lists over capacity:", sum(lens), flush=True)
    # later replays, with eager work in between, reproduce the first bit for bit
    for rep in range(2):
        torch.zeros(1 << 20, device=gpu).add_(1.0)
        torch.cuda.synchronize()
        b.run()
        torch.cuda.synchronize()
        for i in range(len(items)):
            plan, j = b.plan_of(i)
            assert torch.equal(plan.tiles(j), first[i]), (rep, i)
    del b, first
    torch.cuda.empty_cache()
    # a sample through the plain per-call path (the largest recipes among them)
    order = sorted(range(len(items)), key=lambda i: -len(items[i][1]))
    for i in order[:3] + list(range(0, 256, 64)):
        spec, rec = items[i]
        sf = fe.klassificera_och_evaluera_recept(meta_of(fe, spec[4], 8), rec)
        check(sparse_digests(sf), ref[i]["B8"], f"cfg3 plain {i}")


def test_cfg5_served_under_poisson(fe, gpu, digests):
    import oracle as O
    ref = digests["cfg5"]
    recs = fe.recipes.served_requests(5, C.SERVED_COUNT)
    metas = [(meta_of(fe, 512, 8), r) for r in recs]
    srv = fe.server.FieldServer(hbm_budget=200e9, slots=2, max_batch=8)
    try:
        rate, stats, rids = fe.server.poisson_run(srv, metas, rate=2000.0, queries_per_field=1000)
        print("cfg5", rate, stats, flush=True)
        assert stats["completed"] == len(metas)
        for k, rid in enumerate(rids):
            check(sparse_digests(srv.result(rid).sparse_field()), ref[k], f"cfg5 request {k}")
        rng = np.random.default_rng(55)
        for k, npts in ((0, 300000), (len(rids) - 1, 20000)):  # sorted and per-thread query paths
            pts = rng.uniform(0.0, 511.0, (npts, 3)).astype(np.float32)
            got = srv.query(rids[k], pts)
            assert got.tobytes() == O.recipe_eval_points(np.ascontiguousarray(recs[k]).view(O.PRIM), pts).tobytes()
    finally:
        srv.close()
