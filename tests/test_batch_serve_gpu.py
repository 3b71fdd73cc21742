"""Batched (graph-captured) builds and the continuous-batching server, against the
plain per-call path and the oracle."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def mixed_items(fe, count, seed=7, max_n=160, max_p=300):
    rng = np.random.default_rng(seed)
    items = []
    for i in range(count):
        n = int(rng.choice(np.arange(32, max_n + 1, 32)))
        p = int(rng.integers(1, max_p))
        items.append(((0.0, 0.0, 0.0, 1.0, n, n, n), fe.recipes.mixed_recipe(rng, p, n)))
    return items


def same(a, b):
    return (np.array_equal(a.kind, b.kind) and np.array_equal(a.active_ids, b.active_ids)
            and a.tiles.tobytes() == b.tiles.tobytes() and a.corner_min_abs.tobytes() == b.corner_min_abs.tobytes())


def test_graph_batch_equals_plain_every_replay(fe, gpu):
    import torch
    items = mixed_items(fe, 12)
    b = fe.batch.BatchBuilder(items, block=8)
    plain = fe.batch.plain_builds(items, block=8)
    for rep in range(3):
        b.run()
        torch.cuda.synchronize()
        for i in range(len(items)):
            assert same(b.sparse_field(i), plain[i]), (rep, i)
        # eager work between replays
        fe.klassificera_och_evaluera_recept(fe.batch.make_meta(*items[0][:1], 8), items[0][1])


def test_autotuned_batch_equals_plain(fe, gpu):
    items = mixed_items(fe, 9, seed=9, max_n=288)
    tuner = fe.batch.Autotuner(sample=2, tile_waves=(1024, 4096))
    tuner.tune(items)
    b = fe.batch.build_batch(items, tuner=tuner)
    for i, (spec, rec) in enumerate(items):
        B = b.choice[i][0]
        assert B in (4, 8, 16)
        ref = fe.klassificera_och_evaluera_recept(fe.batch.make_meta(spec, B), rec)
        assert same(b.sparse_field(i), ref), i


def test_server_results_alone_and_in_mix(fe, gpu):
    items = mixed_items(fe, 10, seed=11)
    metas = [(fe.batch.make_meta(s, 8), r) for s, r in items]
    srv = fe.server.FieldServer(hbm_budget=8e9, slots=2, max_batch=3)
    try:
        rids = [srv.submit(m, r) for m, r in metas]
        for rid, (m, r) in zip(rids, metas):
            alone = fe.klassificera_och_evaluera_recept(m, r)
            assert same(srv.result(rid, timeout=120).sparse_field(), alone), rid
        rng = np.random.default_rng(3)
        for rid, (m, r) in zip(rids, metas):
            pts = rng.uniform(-2, m.nx + 2, (20001, 3)).astype(np.float32)
            got = srv.query(rid, pts)
            ref = O.recipe_eval_points(np.ascontiguousarray(r).view(O.PRIM), pts)
            assert got.tobytes() == ref.tobytes(), rid
        st = srv.stats()
        assert st["completed"] == len(items) and st["build_p99_ms"] >= st["build_p50_ms"]
        before = srv.reserved
        srv.evict(rids[0])
        assert srv.reserved < before
    finally:
        srv.close()


def test_server_overlapping_arrivals(fe, gpu):
    items = mixed_items(fe, 16, seed=13, max_n=128)
    metas = [(fe.batch.make_meta(s, 8), r) for s, r in items]
    srv = fe.server.FieldServer(hbm_budget=4e9, slots=2, max_batch=2)
    try:
        rate, stats, rids = fe.server.poisson_run(srv, metas, rate=500.0, queries_per_field=1000)
        assert stats["completed"] == len(items) and rate > 0
        for rid, (m, r) in zip(rids, metas):
            assert same(srv.result(rid).sparse_field(), fe.klassificera_och_evaluera_recept(m, r))
    finally:
        srv.close()


def test_plan_on_pools_of_a_finished_plan(fe, gpu):
    """A plan built on the pools of a finished plan (smaller than those pools,
    then larger) equals the plain per-call builds bit for bit; the server's
    slots reuse workspaces this way."""
    import torch
    items = mixed_items(fe, 9, seed=13)
    metas = [(fe.batch.make_meta(s, 8), r) for s, r in items]
    plain = fe.batch.plain_builds(items, block=8)
    first = fe.FieldPlan(metas[:6])
    first.launch()
    torch.cuda.synchronize()
    small = fe.FieldPlan(metas[6:8], pools=first.pools)
    assert small.f32.data_ptr() == first.pools["f32"].data_ptr()  # the fitting pool is reused
    small.launch()
    torch.cuda.synchronize()
    for j in range(2):
        assert same(small.sparse_field(j), plain[6 + j]), j
    large = fe.FieldPlan(metas, pools=small.pools)
    large.launch()
    torch.cuda.synchronize()
    for j in range(len(metas)):
        assert same(large.sparse_field(j), plain[j]), j
