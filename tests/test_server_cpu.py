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
