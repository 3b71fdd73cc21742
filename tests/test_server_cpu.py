"""Host-side admission logic of the field server (no GPU work)."""
import numpy as np
import pytest


def test_admission_refuses_over_budget(fe):
    srv = fe.server.FieldServer(hbm_budget=1e6, start=False)
    meta = fe.GridMeta(0.0, 0.0, 0.0, 1.0, 256, 256, 256, block=8)
    rec = fe.recipes.testfalt(10, 10, 10, 3)
    with pytest.raises(ValueError):
        srv.submit(meta, rec)
    small = fe.GridMeta(0.0, 0.0, 0.0, 1.0, 8, 8, 8, block=8)
    rid = srv.submit(small, rec)
    assert rid == 0 and len(srv.pending) == 1
    with pytest.raises(ValueError):
        srv.submit(small, np.zeros(3, np.float32))
    srv.close()
    with pytest.raises(RuntimeError):
        srv.submit(small, rec)


def test_fifo_admission_respects_budget_and_batch(fe):
    srv = fe.server.FieldServer(hbm_budget=1e9, max_batch=3, start=False)
    small = fe.GridMeta(0.0, 0.0, 0.0, 1.0, 64, 64, 64, block=8)
    rec = fe.recipes.testfalt(10, 10, 10, 3)
    for _ in range(5):
        srv.submit(small, rec)
    batch, need = srv._admit()
    assert [r.rid for r in batch] == [0, 1, 2] and need > 0
    srv.reserved = srv.budget - need / 3 * 0.5
    batch2, _ = srv._admit()
    assert batch2 == []
    srv.close()


def test_workspace_estimate_covers_plan_sizes(fe):
    meta = fe.GridMeta(0.0, 0.0, 0.0, 1.0, 100, 90, 70, block=8)
    rec = fe.recipes.mixed_recipe(np.random.default_rng(0), 50, 100)
    nb = meta.n_blocks
    assert fe.server.workspace_bytes(meta, rec) >= 4 * nb * 8 ** 3


def fake_pools(fe, items):
    """CPU tensors of a plan's pool sizes (what a slot keeps after a batch)."""
    import torch
    dt = {"i32": torch.int32, "f32": torch.float32, "u16": torch.int16}
    return {k: torch.empty(v // fe.field.POOL_ITEM_BYTES[k], dtype=dt[k])
            for k, v in fe.field.plan_pool_bytes(items).items()}


def test_kept_pools_stay_in_reserved(fe):
    """A slot's kept pools count against the budget between batches; the next
    batch on that slot is charged only the pools' growth; reserved returns to
    kept pools plus resident fields after each batch."""
    rec = fe.recipes.testfalt(10, 10, 10, 3)
    small = fe.GridMeta(0.0, 0.0, 0.0, 1.0, 64, 64, 64, block=8)
    big = fe.GridMeta(0.0, 0.0, 0.0, 1.0, 128, 128, 128, block=8)
    srv = fe.server.FieldServer(hbm_budget=64e6, max_batch=2, start=False)
    key = 7  # a slot
    for _ in range(2):
        srv.submit(small, rec)
    batch, charge = srv._admit(srv.retained.get(key))
    assert len(batch) == 2 and srv.reserved + charge <= srv.budget
    held = srv._start(key, charge)
    srv.slot_pools[key] = fake_pools(fe, [(r.meta, r.recipe) for r in batch])
    srv._settle(key, held)
    kept = sum(srv.retained[key].values())
    assert kept > 0 and srv.reserved == kept
    # the same batch again reuses every pool: charged upload and residents only
    for _ in range(2):
        srv.submit(small, rec)
    batch, charge = srv._admit(srv.retained.get(key))
    assert charge == 2 * 32 * len(rec) + sum(r.res for r in batch)
    held = srv._start(key, charge)
    assert key not in srv.retained and srv.reserved == kept + charge
    srv._settle(key, held)
    assert srv.reserved == kept
    # a larger request is charged the growth of each pool, not the whole workspace
    srv.submit(big, rec)
    batch, charge = srv._admit(srv.retained.get(key))
    need = fe.field.plan_pool_bytes([(big, rec)])
    growth = sum(need[k] - min(need[k], srv.retained[key][k]) for k in need)
    assert charge == growth + 32 * len(rec) + batch[0].res
    srv._settle(key, srv._start(key, charge))
    # kept pools alone can block admission of a request that fits the budget
    res = fe.server.resident_estimate(small, rec)
    srv.budget = srv.reserved + 32 * len(rec) + res - 1.0
    assert fe.server.workspace_bytes(small, rec) + res <= srv.budget
    srv.submit(small, rec)
    assert srv._admit(srv.retained.get(key))[0] == []
    srv.budget += 1.0
    assert len(srv._admit(srv.retained.get(key))[0]) == 1
    srv.close()
    assert srv.reserved == 0 and srv.retained == {}


def test_plan_releases_too_small_pools(fe):
    """A pool smaller than a plan needs is removed from the given dict before
    its replacement is allocated; a large enough one is reused."""
    import torch
    rec = fe.recipes.testfalt(10, 10, 10, 3)
    small = fe.GridMeta(0.0, 0.0, 0.0, 1.0, 64, 64, 64, block=8)
    big = fe.GridMeta(0.0, 0.0, 0.0, 1.0, 128, 128, 128, block=8)
    pools = fake_pools(fe, [(big, rec)])
    pools["f32"] = torch.empty(16, dtype=torch.float32)
    keep_i32 = pools["i32"]
    plan = fe.FieldPlan([(small, rec)], device="cpu", pools=pools)
    assert "f32" not in pools and plan.pools["i32"] is keep_i32
    assert plan.pools["f32"].numel() == fe.field.plan_pool_bytes([(small, rec)])["f32"] // 4
