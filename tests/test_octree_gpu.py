"""GPU build and lookups of the 2-bit adaptive block tree, bit-exact against
the oracle (FE/field_adaptive_blocks_v1.py restated in oracle/fe_oracle.c)."""
import numpy as np
import pytest

import oracle as O
from octree_fields import noisy, planes, sphere_box

pytestmark = pytest.mark.gpu


def recipe_field(fe):
    meta = fe.GridMeta(0.0, 0.0, 0.0, 1.0, 96, 80, 72, block=8)
    rec = fe.recipes.mixed_recipe(np.random.default_rng(17), 48, 96)
    return fe.rekonstruera_tat(fe.klassificera_och_evaluera_recept(meta, rec))


def cases(fe):
    return [("sphere_box", sphere_box(), 1.0), ("noisy", noisy(), 0.5), ("planes", planes(), 1.0),
            ("sphere_box_fine", sphere_box((21, 9, 30), 0.25), 0.25), ("recipe_96", recipe_field(fe), 1.0),
            ("single", np.full((1, 1, 1), -2.0, np.float32), 1.0), ("thin", sphere_box((3, 70, 5)), 1.0)]


def test_build_and_sample_match_oracle(fe, gpu):
    for name, d, pitch in cases(fe):
        print("case", name, d.shape, flush=True)
        tree, hist = fe.octree.build(None, d, pitch)
        ref = O.octree_build(d, pitch)
        assert tree.node_count == ref["node_count"] and tree.root_shape == tuple(ref["root_shape"]), name
        for key in ("packed", "branch_prefix", "value_prefix", "values"):
            assert getattr(tree, key).tobytes() == ref[key].tobytes(), (name, key)
        want = {}
        for level, side in enumerate((8, 4, 2, 1)):
            for code in (0, 1, 3):
                if ref["level_counts"][level, code]:
                    want[f"side{side}_kind{code}"] = int(ref["level_counts"][level, code])
        assert hist == want, name
        ijk = np.indices(d.shape).reshape(3, -1).T
        got = tree.sample(ijk)
        assert got.tobytes() == O.octree_sample(ref, ijk).tobytes(), name
        if name != "noisy":  # the reference's gates (distance fields)
            r = got.reshape(d.shape)
            narrow = np.abs(d) <= pitch
            assert r[narrow].tobytes() == d[narrow].tobytes() and np.array_equal(r < 0, d < 0), name
        rng = np.random.default_rng(3)
        q = np.stack([rng.integers(0, s, 5000) for s in d.shape], 1)
        assert tree.sample(q).tobytes() == O.octree_sample(ref, q).tobytes(), name


def test_histogram_order_and_determinism(fe, gpu):
    d = sphere_box((64, 48, 40))
    t1, h1 = fe.octree.build(None, d, 1.0)
    t2, h2 = fe.octree.build(None, d, 1.0)
    assert t1.fingerprint() == t2.fingerprint() and list(h1) == list(h2)
    # first-occurrence order: the leaf labels of a level follow those of the levels above
    sides = [int(k.split("_")[0][4:]) for k in h1]
    assert sides == sorted(sides, reverse=True)
    assert t1.stored_bytes() == 80 + t1.packed.nbytes + t1.branch_prefix.nbytes + t1.value_prefix.nbytes + \
        t1.values.nbytes


def test_sample_rejects_out_of_domain(fe, gpu):
    tree, _ = fe.octree.build(None, planes(), 1.0)
    for bad in ([[16, 0, 0]], [[-1, 0, 0]], [[0, 0]], np.zeros((2, 4), np.int64)):
        with pytest.raises(ValueError):
            tree.sample(bad)
