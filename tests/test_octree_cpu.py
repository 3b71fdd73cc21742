"""The oracle of the 2-bit adaptive block tree (FE/field_adaptive_blocks_v1.py)
against an independent breadth-first restatement, brute-force ranks and the
reference's own gates (exact sign, byte-identical narrow band, finite)."""
import numpy as np
import pytest

import oracle as O
from octree_fields import noisy, planes, sphere_box

CASES = [("sphere_box", sphere_box, 1.0), ("noisy", noisy, 0.5), ("planes", planes, 1.0),
         ("sphere_box_fine", lambda: sphere_box((21, 9, 30), 0.25), 0.25)]


def bfs_restatement(dense, pitch):
    """The construction spelled out with Python loops over the oracle classifier."""
    nx, ny, nz = dense.shape
    levels = {}
    for side in (8, 4, 2, 1):
        k = O.classify_dense(dense, side, np.float32(np.sqrt(3.0) * side * pitch))
        levels[side] = k.reshape(-(-nx // side), -(-ny // side), -(-nz // side))
    nodes = [(8 * x, 8 * y, 8 * z, 8) for x, y, z in np.ndindex(levels[8].shape)]
    codes, values = [], []
    i = 0
    while i < len(nodes):
        x, y, z, s = nodes[i]
        if x >= nx or y >= ny or z >= nz:
            code = 0
        else:
            kind = int(levels[s][x // s, y // s, z // s])
            if kind:
                code = 0 if kind > 0 else 1
            elif s == 1:
                code = 3
                values.append(dense[x, y, z])
            else:
                code = 2
                h = s // 2
                nodes.extend((x + a * h, y + b * h, z + c * h, h) for a, b, c in np.ndindex(2, 2, 2))
        codes.append(code)
        i += 1
    return np.array(codes, np.uint8), np.array(values, np.float32)


@pytest.mark.parametrize("name,make,pitch", CASES)
def test_oracle_tree_structure(name, make, pitch):
    d = make()
    t = O.octree_build(d, pitch)
    codes, values = bfs_restatement(d, pitch)
    assert np.array_equal(t["codes"], codes) and np.array_equal(t["values"], values), name
    assert t["node_count"] == len(codes) and t["level_counts"].sum() == len(codes)
    for code in (2, 3):  # checkpoints = counts of the code before every 64th byte
        want = [int(np.sum(codes[:4 * b] == code)) for b in range(0, len(t["packed"]) + 1, 64)]
        got = (t["branch_prefix"] if code == 2 else t["value_prefix"]).tolist()
        assert got == want
    # children of the b-th branch start at root_count + 8 b
    branches = np.nonzero(codes == 2)[0]
    assert len(codes) == t["root_count"] + 8 * len(branches)


# the gates hold for distance fields (1-Lipschitz in voxel units), not for the noisy case
@pytest.mark.parametrize("name,make,pitch", [c for c in CASES if c[0] != "noisy"])
def test_oracle_sample_gates(name, make, pitch):
    d = make()
    t = O.octree_build(d, pitch)
    ijk = np.indices(d.shape).reshape(3, -1).T
    r = O.octree_sample(t, ijk).reshape(d.shape)
    narrow = np.abs(d) <= pitch
    assert r[narrow].tobytes() == d[narrow].tobytes()
    assert np.array_equal(r < 0, d < 0) and np.isfinite(r).all()
    # leaves above side 1 give +/- f32(sqrt(3) * side * pitch); value leaves their sample
    sides = {np.float32(np.sqrt(3.0) * s * pitch) for s in (8, 4, 2, 1)}
    far = ~np.isin(r, d)
    assert set(np.abs(r[far]).tolist()) <= {float(v) for v in sides}


def test_oracle_descent_brute_force():
    d = noisy((13, 10, 9), seed=11)
    t = O.octree_build(d, 0.5)
    codes = t["codes"]
    ijk = np.indices(d.shape).reshape(3, -1).T
    got = O.octree_sample(t, ijk)
    want = []
    for i, j, k in ijk:
        node = ((i // 8) * t["root_shape"][1] + j // 8) * t["root_shape"][2] + k // 8
        size = 8
        while True:
            c = codes[node]
            if c < 2:
                want.append(np.float32((1.0 if c == 0 else -1.0) * np.sqrt(3.0) * size * 0.5))
                break
            if c == 3:
                want.append(t["values"][int(np.sum(codes[:node] == 3))])
                break
            half = size // 2
            node = t["root_count"] + 8 * int(np.sum(codes[:node] == 2)) + 4 * ((i // half) % 2) + \
                2 * ((j // half) % 2) + (k // half) % 2
            size = half
    assert np.array(want, np.float32).tobytes() == got.tobytes()


def test_package_rank_matches_counts(fe):
    d = sphere_box((33, 30, 41))
    t = O.octree_build(d, 1.0)
    tree = fe.octree.AdaptiveBlocks(t["shape"], 1.0, t["root_shape"], t["root_count"], t["node_count"], t["packed"],
                                    t["branch_prefix"], t["value_prefix"], t["values"])
    nodes = np.arange(t["node_count"])
    for code in (2, 3):
        cum = np.concatenate(([0], np.cumsum(t["codes"] == code)))[:-1]
        assert np.array_equal(tree.rank_before(nodes, code), cum)
    assert np.array_equal(tree.kinds(nodes), t["codes"])
