"""The thin-feature obstacle workloads on the CPU: the generator reproduces the
committed masks (so the GPU digests refer to the same inputs), the occupancy is
about 12% at 768^3, and the oracle agrees with exhaustive minimisation on small
volumes of the same construction."""
import hashlib
import json

import numpy as np

import oracle as O
from conftest import ROOT


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_generator_reproduces_fixture_masks(fe):
    g = json.loads((ROOT / "tests/golden/edt_volumes.json").read_text())
    shapes = fe.recipes.thin_feature_shapes(4, 64)
    assert [list(s) for s in shapes] == [b["shape"] for b in g["batch"]]
    vols = fe.recipes.thin_feature_batch(4, 8)
    for v, b in zip(vols, g["batch"][:8]):
        assert sha(v) == b["mask_sha"] and int(np.count_nonzero(v)) == b["occupied"]
    occ = g["volume_768"]["occupied"] / 768 ** 3
    assert 0.11 < occ < 0.13


def test_thin_features_present(fe):
    """1-voxel walls and 1-voxel gaps occur (a wall: an occupied voxel whose two
    neighbours along one axis are free; a gap: the converse)."""
    v = fe.recipes.obstacle_volume((96, 96, 96), np.random.default_rng(7), 60)
    walls = gaps = 0
    for ax in range(3):
        a = np.moveaxis(v, ax, 0).astype(bool)
        walls += np.count_nonzero(a[1:-1] & ~a[:-2] & ~a[2:])
        gaps += np.count_nonzero(~a[1:-1] & a[:-2] & a[2:])
    assert walls > 100 and gaps > 100


def test_oracle_exhaustive_on_small_obstacles(fe):
    rng = np.random.default_rng(8)
    for shape in [(12, 17, 9), (20, 6, 31)]:
        m = fe.recipes.obstacle_volume(shape, rng, 6, scale=8.0)
        m[0, 0, 0], m[-1, -1, -1] = 0, 1
        assert np.array_equal(O.edt_envelope(m), O.edt_exhaustive(m).astype(np.int32))
        assert np.array_equal(O.edt_envelope(1 - m), O.edt_exhaustive(1 - m).astype(np.int32))
