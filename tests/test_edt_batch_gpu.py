"""GPU parity of the thin-feature EDT workloads: the 768^3 obstacle volume and
uneven volumes batched into one launch per axis pass, bit-exact against the
oracle (full-size digests in tests/golden/edt_volumes.json, direct comparison
at small sizes)."""
import hashlib
import json

import numpy as np
import pytest

import oracle as O
from conftest import ROOT

pytestmark = pytest.mark.gpu


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def volumes_golden():
    return json.loads((ROOT / "tests/golden/edt_volumes.json").read_text())


RAGGED = {
    # rows shorter than 16 voxels in the batch: int32 z planes
    "short_rows": [(1, 5, 9), (33, 1, 17), (40, 33, 27), (64, 64, 64), (17, 130, 5), (7, 9, 1), (3, 600, 8),
                   (520, 3, 12), (24, 24, 24), (24, 24, 24)],
    # every row >= 16 voxels: the y pass reads z summaries (rows of 16, 17, 31
    # and 1100 voxels, tiles of 16 cut short at the row end)
    "zsum_rows": [(1, 5, 16), (33, 1, 17), (40, 33, 27), (64, 64, 64), (17, 130, 31), (3, 600, 20), (520, 3, 48),
                  (5, 4, 1100), (24, 24, 24), (24, 24, 24)],
}


def ragged_masks(kind="short_rows"):
    """Small volumes: 1-voxel axes, nz not a multiple of 4, an empty and a full
    volume, lines longer than 128 and 512 (other tile shapes in one batch)."""
    rng = np.random.default_rng(41)
    out = []
    for shape in RAGGED[kind]:
        if len(out) == 8:
            m = np.zeros(shape, np.uint8)
        elif len(out) == 9:
            m = np.ones(shape, np.uint8)
        else:
            from conftest import load_package
            R = load_package().recipes
            m = R.obstacle_volume(shape, rng, max(2, R.obstacle_count(shape, 40000)), scale=12.0)
            m[tuple(rng.integers(0, s) for s in shape)] ^= 1
        out.append(m)
    return out


@pytest.mark.parametrize("kind", sorted(RAGGED))
@pytest.mark.parametrize("planes", [2, 1])
def test_batch_ragged_vs_oracle(fe, gpu, planes, kind):
    masks = ragged_masks(kind)
    print("ragged batch", planes, kind, [m.shape for m in masks], flush=True)
    b = fe.EDTBatch([m.shape for m in masks], planes=planes)
    b.set_masks(masks)
    b.run()
    for i, m in enumerate(masks):
        got = b.result(i).cpu().numpy()
        if planes == 2:
            _, want = O.edt_pair(m, 1.0)
        else:
            want = O.edt_envelope(m)
        assert np.array_equal(got, want), (i, m.shape)
    b.close()


def test_batch_64_uneven_graph(fe, gpu, volumes_golden):
    import torch
    vols = fe.recipes.thin_feature_batch(4, 64)
    for v, g in zip(vols, volumes_golden["batch"]):
        assert list(v.shape) == g["shape"] and sha(v) == g["mask_sha"]
    print("batch of 64, voxels", sum(v.size for v in vols), flush=True)
    b = fe.EDTBatch([v.shape for v in vols], planes=2)
    b.set_masks(vols)
    graph = b.capture()
    for replay in range(2):
        b.out.fill_(-1)
        graph.replay()
        torch.cuda.synchronize()
        for i, g in enumerate(volumes_golden["batch"]):
            assert sha(b.result(i).cpu().numpy()) == g["pair_sha"], (replay, i, g["shape"])
    # the batched volumes equal the single-volume path
    for i in (0, 17, 63):
        single = fe.edt_device(b.mask(i).contiguous(), 2)
        assert torch.equal(single, b.result(i)), i
    b.close()


def test_volume_768(fe, gpu, volumes_golden):
    import torch
    g = volumes_golden["volume_768"]
    vol = fe.recipes.thin_feature_volume_768(4)
    assert sha(vol) == g["mask_sha"]
    print("768^3 pair", flush=True)
    mask = torch.from_numpy(vol).to(gpu)
    del vol
    pair = fe.edt_device(mask, 2)
    torch.cuda.synchronize()
    assert sha(pair.cpu().numpy()) == g["pair_sha"]
    del pair, mask
    torch.cuda.empty_cache()
