"""Pins the CPU oracle against the reference's own golden data (CPU only).

What pins what:
  or_edt_exhaustive / or_edt_envelope   16 integer mask families (both polarities, int64 squares) and
                                        23 pair cases (digests of both int32 planes) + brute force
  or_edt_signed / or_edt_pair           signed digests at 3 + 5 + 3 + 1 pitches
  or_csg_execute                        frozen CSG outputs of 4 cases (digest)
  or_pack                               node/payload/origin digests of the reference's packer,
                                        the reference's own pack.cpp (oracle/_ref) on random fields,
                                        and the lossless invariant
  or_query / or_samples_at              4 x 4097 frozen query results (bit-exact)
  or_gather                             status codes and bits of the reference's gather.cpp (oracle/_ref)
  or_classify/eval/reconstruct/checksum the test field's recorded counts, voxel volume, int64
                                        checksum and the planted-fault volume
  or_recipe_eval                        FP64 closed forms of every primitive
"""
import hashlib
import math
import ctypes as ct

import numpy as np
import pytest

import oracle as O
import csg_inputs

FIELD_ORDER = ["bracket_plate_rectangular_holes_0.5_0_field", "bracket_plate_rectangular_holes_0.25_0_field",
               "sphere_plate_round_hole_0.5_0_field", "sphere_plate_round_hole_0.25_0_field"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def unpack(bits, shape):
    return np.unpackbits(bits)[:int(np.prod(shape))].reshape(shape).astype(bool)


@pytest.fixture(scope="module")
def edt_int():
    from conftest import ROOT
    return np.load(ROOT / "tests/golden/edt_integer.npz")


@pytest.fixture(scope="module")
def edt_pairs():
    from conftest import ROOT
    return np.load(ROOT / "tests/golden/edt_pairs.npz")


def test_edt_integer_families(edt_int, golden):
    for key, meta in golden["edt_integer"].items():
        shape = tuple(meta["shape"])
        mask = unpack(edt_int[f"{key}_mask"], shape)
        s0, s1 = edt_int[f"{key}_sq0"], edt_int[f"{key}_sq1"]
        assert np.array_equal(O.edt_exhaustive(~mask), s0), meta["name"]
        assert np.array_equal(O.edt_exhaustive(mask), s1), meta["name"]
        e0, e1 = O.edt_envelope(~mask), O.edt_envelope(mask)
        assert np.array_equal(e0, s0) and np.array_equal(e1, s1), meta["name"]
        for pitch, digest in meta["signed"].items():
            assert sha(O.edt_signed_from_squares(e0, e1, float(pitch))) == digest, (meta["name"], pitch)


def test_edt_pair_cases(edt_pairs, golden):
    for key, meta in golden["edt_pairs"].items():
        shape = tuple(meta["shape"])
        mask = unpack(edt_pairs[f"{key}_mask"], shape)
        for pitch, digest in meta["signed"].items():
            out, sq = O.edt_pair(mask, float(pitch))
            assert sha(sq[0]) == meta["squared0"] and sha(sq[1]) == meta["squared1"], meta["name"]
            assert sha(out) == digest, (meta["name"], pitch)


def test_edt_brute_force_tiny():
    rng = np.random.default_rng(5)
    for _ in range(40):
        shape = tuple(int(x) for x in rng.integers(1, 7, 3))
        mask = rng.random(shape) < rng.uniform(0.05, 0.95)
        if mask.all() or not mask.any():
            continue
        sites = np.argwhere(~mask)
        grid = np.indices(shape).reshape(3, -1).T
        brute = np.where(mask.ravel(), ((grid[:, None, :] - sites[None]) ** 2).sum(-1).min(1), 0).reshape(shape)
        assert np.array_equal(O.edt_envelope(mask), brute)
        assert np.array_equal(O.edt_exhaustive(mask), brute)


def test_csg_frozen_outputs(golden):
    for curved in (False, True):
        for pitch in (0.5, 0.25):
            fields, prog, origin, pitch, shape = csg_inputs.case(curved, pitch)
            packed = [O.pack(a) for a in fields]
            code, out = O.csg_execute(packed, encode(prog), shape)
            name = f"{csg_inputs.NAMES[curved]}_{pitch}_0_field"
            assert code == 0
            assert sha(out) == golden["csg_pack"][name]["field_sha"], name


def encode(prog):
    P = np.zeros(len(prog), O.INSTR)
    for i, (op, l, r) in enumerate(prog):
        P[i]["op"] = {"union": 0, "difference": 1, "offset": 2}[op]
        P[i]["left"] = l
        if op == "offset":
            P[i]["offset"] = r
        else:
            P[i]["right"] = r
    return P


def frozen_fields():
    out = {}
    for curved in (False, True):
        for pitch in (0.5, 0.25):
            fields, prog, origin, pitch, shape = csg_inputs.case(curved, pitch)
            code, dense = O.csg_execute([O.pack(a) for a in fields], encode(prog), shape)
            out[f"{csg_inputs.NAMES[curved]}_{pitch}_0_field"] = dense
    return out


def test_pack_hashes_and_queries(golden):
    from conftest import ROOT
    q = np.load(ROOT / "tests/golden/pack_query.npz")
    dense = frozen_fields()
    points = csg_inputs.query_points([dense[n].shape for n in FIELD_ORDER])
    for name, pts in zip(FIELD_ORDER, points):
        raw = dense[name]
        nodes, payload = O.pack(raw, 512)
        assert [sha(nodes), sha(payload), sha(np.zeros(3))] == golden["csg_pack"][name]["pack_hashes"], name
        idx = np.indices(raw.shape).reshape(3, -1).T
        assert O.samples_at(nodes, payload, idx).tobytes() == raw.tobytes()
        got = O.query(nodes, payload, raw.shape, np.zeros(3), 1.0, pts)
        assert got.tobytes() == q[name + "_query"].tobytes(), name


def ref_pack(lib, a, leaf):
    lib.adaptive_pack.restype = ct.c_void_p
    lib.adaptive_pack.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int]
    for n in ("adaptive_nodes", "adaptive_samples"):
        getattr(lib, n).restype = ct.c_int64
        getattr(lib, n).argtypes = [ct.c_void_p]
    lib.adaptive_copy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_int64]
    lib.adaptive_free.argtypes = [ct.c_void_p]
    sh = np.array(a.shape, np.int32)
    p = lib.adaptive_pack(a.ctypes.data, sh.ctypes.data, leaf)
    nodes = np.empty(lib.adaptive_nodes(p), O.NODE)
    pay = np.empty(lib.adaptive_samples(p), np.float32)
    assert lib.adaptive_copy(p, nodes.ctypes.data, len(nodes), pay.ctypes.data, len(pay)) == 0
    lib.adaptive_free(p)
    return nodes, pay


def random_fields(rng):
    cases = [rng.normal(size=(19, 23, 17)).astype(np.float32), np.ones((29, 1, 17), np.float32),
             np.zeros((17, 19, 23), np.float32)]
    cases[-1].ravel()[::2] = np.float32(-0.)
    blocky = np.repeat(np.repeat(rng.integers(0, 3, (6, 5, 7)).astype(np.float32), 5, 0), 4, 1)
    cases.append(np.ascontiguousarray(blocky[:, :, :]))
    cases.append(np.ascontiguousarray(cases[0][::2, ::2, ::-1]))
    return cases


def test_pack_matches_reference_native():
    lib = O.ref_lib("pack")
    if lib is None:
        pytest.skip("reference pack.cpp not built here (oracle/_ref)")
    rng = np.random.default_rng(17)
    for a in random_fields(rng):
        for leaf in (8, 64, 512, 4096):
            n1, p1 = O.pack(a, leaf)
            n2, p2 = ref_pack(lib, a, leaf)
            assert n1.tobytes() == n2.tobytes() and p1.tobytes() == p2.tobytes()


def test_gather_matches_reference_native():
    lib = O.ref_lib("gather")
    if lib is None:
        pytest.skip("reference gather.cpp not built here (oracle/_ref)")
    lib.csg_gather.argtypes = [ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_void_p,
                               ct.c_void_p]
    rng = np.random.default_rng(3)
    a = np.repeat(rng.normal(size=(9, 13, 6)).astype(np.float32), 3, axis=2)
    nodes, pay = O.pack(a, 8)
    for _ in range(60):
        lo = rng.integers(0, a.shape)
        hi = lo + rng.integers(1, 12, 3)
        hi = np.minimum(hi, np.array(a.shape) + rng.integers(0, 2, 3))
        lo32, hi32 = lo.astype(np.int32), hi.astype(np.int32)
        c1, o1 = O.gather(nodes, pay, lo32, hi32)
        o2 = np.empty_like(o1)
        c2 = lib.csg_gather(nodes.ctypes.data, len(nodes), pay.ctypes.data, len(pay), lo32.ctypes.data,
                            hi32.ctypes.data, o2.ctypes.data)
        assert c1 == c2
        if c1 == 0:
            assert o1.tobytes() == o2.tobytes()
    broken = nodes.copy()
    broken["offset"][np.argmax(broken["children"][:, 0] < 0)] = 10 ** 6
    c1, _ = O.gather(broken, pay, np.zeros(3, np.int32), np.array([4, 4, 4], np.int32))
    assert c1 == 3


HX, HY, HZ, R = 20.0, 15.0, 10.0, 6.0


def _testfield_grid(pitch, block=8):
    pad = 1.0
    nx = int(round((2 * HX + 2 * pad) / pitch)) + 1
    ny = int(round((2 * HY + 2 * pad) / pitch)) + 1
    nz = int(round((2 * HZ + 2 * pad) / pitch)) + 1
    g = O.grid_record(np.float32(-HX - pad), np.float32(-HY - pad), np.float32(-HZ - pad), np.float32(pitch),
                      nx, ny, nz, block, np.float32(math.sqrt(3.0) * block * pitch))
    prims = np.zeros(2, O.PRIM)
    prims[0] = (1, 0, 0, 0, 0, HX, HY, HZ)
    prims[1] = (6, 1, 0, 0, 0, R, 0, 0)
    return prims, g, (nx, ny, nz)


def test_testfield_known_answers(golden):
    tf = golden["testfield"]
    prims, g, (nx, ny, nz) = _testfield_grid(0.25)
    kind, cma, ids, tiles = O.build_sparse(prims, g)
    assert (len(kind), len(ids), int((kind == -1).sum()), int((kind == 1).sum())) == \
        (tf["n_blocks"], tf["n_active"], tf["n_interior"], tf["n_exterior"])
    dense = O.reconstruct(kind, ids, tiles, nx, ny, nz, 8)
    assert dense.size == tf["dense_voxels"] and tiles.size == tf["active_voxels"]
    assert int((dense < 0).sum()) * 0.25 ** 3 == tf["vol_voxel_mm3"]
    assert O.checksum_i64(tiles) == round(tf["checksum_i64"] * 1e6)
    _, _, _, bad = O.build_sparse(prims, g, corrupt=True)
    assert int((O.reconstruct(kind, ids, bad, nx, ny, nz, 8) < 0).sum()) * 0.25 ** 3 == tf["vol_corrupt_mm3"]
    for pitch, key in ((0.9, "n_active_pitch_0_9"),):
        prims, g, _ = _testfield_grid(pitch)
        assert len(O.build_sparse(prims, g)[2]) == tf[key]


def test_classifier_never_lies_and_fp64_error():
    """Interior blocks all negative, exterior all >= 0; fp32 values near FP64 closed forms."""
    import fp64 as recipes_fp64
    rng = np.random.default_rng(11)
    prims = recipes_fp64.random_recipe(rng, 24, 48.0)
    g = O.grid_record(0.0, 0.0, 0.0, 1.0, 48, 48, 48, 8, np.float32(math.sqrt(3.0) * 8))
    kind, cma, ids, tiles = O.build_sparse(prims, g)
    pts = np.stack(np.meshgrid(*(np.arange(48, dtype=np.float32),) * 3, indexing="ij"), -1).reshape(-1, 3)
    dense = O.recipe_eval_points(prims, pts).reshape(48, 48, 48)
    kb = np.repeat(np.repeat(np.repeat(kind.reshape(6, 6, 6), 8, 0), 8, 1), 8, 2)
    assert np.all(dense[kb == -1] < 0) and np.all(dense[kb == 1] >= 0)
    exact = recipes_fp64.eval_fp64(prims, pts.astype(np.float64))
    assert np.max(np.abs(dense.ravel() - exact)) < 1e-4
