"""GPU parity of the leaf-box pack, lookups, trilinear query, tile reader and
tiled CSG against the reference's frozen data, the oracle and (where built)
the reference's own C++ natives."""
import ctypes as ct
import hashlib

import numpy as np
import pytest

import oracle as O
import csg_inputs

pytestmark = pytest.mark.gpu
FIELD_ORDER = ["bracket_plate_rectangular_holes_0.5_0_field", "bracket_plate_rectangular_holes_0.25_0_field",
               "sphere_plate_round_hole_0.5_0_field", "sphere_plate_round_hole_0.25_0_field"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


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


def test_frozen_csg_pack_query(fe, golden, gpu):
    from conftest import ROOT
    q = np.load(ROOT / "tests/golden/pack_query.npz")
    results = {}
    for curved in (False, True):
        for pitch in (0.5, 0.25):
            ins, prog, origin, pitch, shape = csg_inputs.case(curved, pitch)
            fields = [fe.AdaptiveSampleBlocks(a, origin, pitch) for a in ins]
            for f, a in zip(fields, ins):
                n, p = O.pack(a)
                assert f.nodes.tobytes() == n.tobytes() and f.payload.tobytes() == p.tobytes()
            name = f"{csg_inputs.NAMES[curved]}_{pitch}_0_field"
            first, dense = fe.csg.evaluate_dense(fields, prog)
            dense = dense.cpu().numpy()
            print(name, flush=True)
            assert sha(dense) == golden["csg_pack"][name]["field_sha"], name
            out = fe.evaluate(fields, prog)
            results[name] = (dense, out)
    points = csg_inputs.query_points([results[n][0].shape for n in FIELD_ORDER])
    for name, pts in zip(FIELD_ORDER, points):
        dense, _ = results[name]
        packed = fe.AdaptiveSampleBlocks(dense, [0, 0, 0], 1)
        assert [sha(packed.nodes), sha(packed.payload), sha(packed.origin)] == golden["csg_pack"][name]["pack_hashes"]
        assert packed.query(pts).tobytes() == q[name + "_query"].tobytes(), name
        idx = np.indices(dense.shape).reshape(3, -1).T
        assert packed.samples_at(idx).tobytes() == dense.tobytes()
        assert packed.dense_device().cpu().numpy().tobytes() == dense.tobytes()


def control_fields(rng):
    cases = [rng.normal(size=(19, 23, 17)).astype(np.float32), np.ones((29, 1, 17), np.float32),
             np.ones((1, 19, 31), np.float32), np.ones((17, 29, 1), np.float32), np.zeros((17, 19, 23), np.float32)]
    cases[-1].ravel()[::2] = np.float32(-0.)
    cases.append(np.ascontiguousarray(cases[0][::2, ::2, ::-1]))
    blocky = np.repeat(np.repeat(rng.integers(0, 3, (6, 5, 7)).astype(np.float32), 5, 0), 4, 1)
    cases.append(np.ascontiguousarray(blocky))
    cases.append(np.repeat(rng.normal(size=(40, 3, 50)).astype(np.float32), 7, axis=1))
    return cases


def test_pack_matches_oracle_and_reference(fe, gpu):
    rng = np.random.default_rng(445)
    ref = O.ref_lib("pack")
    for a in control_fields(rng):
        for leaf in (8, 64, 512, 4096):
            g = fe.AdaptiveSampleBlocks(a, [.1, -2, 3], .7, leaf_samples=leaf)
            n, p = O.pack(a, leaf)
            assert g.nodes.tobytes() == n.tobytes() and g.payload.tobytes() == p.tobytes(), (a.shape, leaf)
            h = fe.NativeAdaptiveBlocks(a, [.1, -2, 3], .7, leaf_samples=leaf)
            assert h.nodes.tobytes() == n.tobytes() and h.payload.tobytes() == p.tobytes()
            if ref is not None:
                r = fe.NativeAdaptiveBlocks(a, [.1, -2, 3], .7, leaf_samples=leaf, library=str(O.REF_DIR / "libref_pack.so"))
                assert r.nodes.tobytes() == n.tobytes()


def test_pack_rejections(fe, gpu):
    bad = [(np.ones((2, 2, 2), np.float64), [0, 0, 0], 1, 512), (np.ones((2, 2), np.float32), [0, 0, 0], 1, 512),
           (np.full((2, 2, 2), np.nan, np.float32), [0, 0, 0], 1, 512), (np.ones((2, 2, 2), np.float32), [0, 0], 1, 512),
           (np.ones((2, 2, 2), np.float32), [0, 0, 0], 0, 512), (np.ones((2, 2, 2), np.float32), [0, 0, 0], True, 512),
           (np.ones((2, 2, 2), np.float32), [0, 0, 0], 1, True), (np.ones((2, 2, 2), np.float32), [0, 0, 0], 1, 4)]
    for cls in (fe.AdaptiveSampleBlocks, fe.NativeAdaptiveBlocks):
        n = 0
        for raw, origin, pitch, leaf in bad:
            try:
                cls(raw, origin, pitch, leaf_samples=leaf)
            except ValueError:
                n += 1
        assert n == len(bad)
    lib = fe._lib.lib()
    a = np.full((3, 3, 3), np.inf, np.float32)
    assert not lib.adaptive_pack(a.ctypes.data, np.array(a.shape, np.int32).ctypes.data, 512)


def test_csg_controls_vs_oracle_and_reference(fe, gpu):
    rng = np.random.default_rng(77)
    ref = O.ref_lib("csg_batch")
    if ref is not None:
        ref.csg_execute.argtypes = [ct.c_void_p] * 5 + [ct.c_void_p, ct.c_void_p, ct.c_int32, ct.c_void_p]
        ref.csg_execute.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int32, ct.c_void_p,
                                    ct.c_void_p, ct.c_int32, ct.c_void_p]
    cases = []
    for shape in ((1, 17, 19), (19, 1, 17), (17, 19, 1), (19, 21, 17), (33, 70, 40)):
        raw = rng.normal(size=shape).astype(np.float32)
        raw.ravel()[::2] = np.float32(-0.)
        cases.append([raw, np.full(shape, .25, np.float32)])
    zeros = np.zeros((17, 19, 23), np.float32)
    neg = zeros.copy()
    neg.ravel()[::2] = np.float32(-0.)
    cases.append([zeros, neg, np.full_like(zeros, np.nextafter(np.float32(0), np.float32(1)))])
    programs = [[("union", 0, 1)], [("difference", 0, 1)], [("offset", 0, .1), ("offset", 2, -.1)],
                [("union", 1, 0)], [("offset", 1, .1), ("offset", 3, -.1)], [("union", 0, 1), ("difference", 2, 0),
                                                                          ("union", 3, 2)]]
    for arrays in cases:
        fields = [fe.AdaptiveSampleBlocks(a, [0, 0, 0], 1) for a in arrays]
        packed = [O.pack(a) for a in arrays]
        for prog in programs:
            if max(max(x for x in (l, r) if isinstance(x, int)) for _, l, r in prog) >= len(arrays) + len(prog):
                continue
            code, want = O.csg_execute(packed, encode(prog), arrays[0].shape)
            if code:
                continue
            _, got = fe.csg.evaluate_dense(fields, prog)
            assert got.cpu().numpy().tobytes() == want.tobytes(), (arrays[0].shape, prog)
            # the host ABI, as the reference batch evaluator calls it
            lib = fe._lib.lib()
            P = encode(prog)
            nodes = (ct.c_void_p * len(packed))(*[p[0].ctypes.data for p in packed])
            pays = (ct.c_void_p * len(packed))(*[p[1].ctypes.data for p in packed])
            counts = np.array([len(p[0]) for p in packed], np.int64)
            lens = np.array([len(p[1]) for p in packed], np.int64)
            sh = np.array(arrays[0].shape, np.int32)
            out = np.empty(arrays[0].shape, np.float32)
            assert lib.csg_execute(nodes, counts.ctypes.data, pays, lens.ctypes.data, len(packed), sh.ctypes.data,
                                   P.ctypes.data, len(P), out.ctypes.data) == 0
            assert out.tobytes() == want.tobytes()


def run_codes(lib, packed, prog, shape):
    nodes = (ct.c_void_p * len(packed))(*[p[0].ctypes.data for p in packed])
    pays = (ct.c_void_p * len(packed))(*[p[1].ctypes.data for p in packed])
    counts = np.array([len(p[0]) for p in packed], np.int64)
    lens = np.array([len(p[1]) for p in packed], np.int64)
    sh = np.array(shape, np.int32)
    out = np.empty(shape, np.float32)
    return lib.csg_execute(nodes, counts.ctypes.data, pays, lens.ctypes.data, len(packed), sh.ctypes.data,
                           prog.ctypes.data, len(prog), out.ctypes.data)


def test_status_codes_match_oracle(fe, gpu):
    lib = fe._lib.lib()
    rng = np.random.default_rng(5)
    a = rng.normal(size=(40, 50, 30)).astype(np.float32)
    b = (rng.normal(size=(40, 50, 30)) * 1e37).astype(np.float32)
    good = [O.pack(a), O.pack(b)]
    shape = a.shape
    progs = {
        "bad_ref": encode([("union", 0, 99)]),
        "nan_offset": encode([("offset", 0, float("nan"))]),
        "overflow": encode([("offset", 1, -3e38)]),
        "ok": encode([("union", 0, 1), ("difference", 2, 0)]),
    }
    bad_op = encode([("union", 0, 1)])
    bad_op[0]["op"] = 7
    progs["bad_op"] = bad_op
    for name, prog in progs.items():
        want, _ = O.csg_execute(good, prog, shape)
        got = run_codes(lib, good, prog, shape)
        assert got == want, name
    # malformed and coverage faults on the first operand
    n0, p0 = good[0]
    for mode in ("offset", "shape", "overlap", "gap", "gap_late"):
        nodes = n0.copy()
        leaf = int(np.argmax(nodes["children"][:, 0] < 0))
        if mode == "offset":
            nodes["offset"][leaf] = 10 ** 9
        elif mode == "shape":
            nodes["stored_shape"][leaf] = nodes["shape"][leaf] + 1
        elif mode == "overlap":
            nodes = np.concatenate((nodes, nodes[leaf:leaf + 1]))
        elif mode == "gap":
            nodes["shape"][leaf] = [1, 1, 1]
            nodes["stored_shape"][leaf] = [1, 1, 1]
        else:
            last = len(nodes) - 1 - int(np.argmax(nodes["children"][::-1, 0] < 0))
            nodes["shape"][last] = [1, 1, 1]
            nodes["stored_shape"][last] = [1, 1, 1]
        for prog in (progs["ok"], progs["overflow"]):
            want, _ = O.csg_execute([(nodes, p0), good[1]], prog, shape)
            assert run_codes(lib, [(nodes, p0), good[1]], prog, shape) == want, mode


def test_tile_reader_codes_and_bits(fe, gpu):
    rng = np.random.default_rng(3)
    a = np.repeat(rng.normal(size=(9, 13, 6)).astype(np.float32), 3, axis=2)
    f = fe.AdaptiveSampleBlocks(a, [0, 0, 0], 1, leaf_samples=8)
    lib = fe._lib.lib()
    for _ in range(40):
        lo = rng.integers(0, a.shape)
        hi = np.minimum(lo + rng.integers(1, 12, 3), np.array(a.shape) + rng.integers(0, 2, 3)).astype(np.int32)
        lo = lo.astype(np.int32)
        c1, o1 = O.gather(f.nodes, f.payload, lo, hi)
        o2 = np.empty_like(o1)
        c2 = lib.csg_gather(f.nodes.ctypes.data, len(f.nodes), f.payload.ctypes.data, len(f.payload), lo.ctypes.data,
                            hi.ctypes.data, o2.ctypes.data)
        assert c1 == c2
        if c1 == 0:
            assert o1.tobytes() == o2.tobytes()
            assert fe.gather_tile(f, lo, np.minimum(hi, a.shape)).tobytes() == \
                a[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]].tobytes() or np.any(hi > a.shape)
    for mode in ("offset", "shape", "overlap", "gap"):
        nodes = f.nodes.copy()
        if mode == "offset":
            nodes["offset"][np.argmax(nodes["children"][:, 0] < 0)] = 1000000
        if mode == "shape":
            nodes["stored_shape"][np.argmax(nodes["children"][:, 0] < 0)] = [5, 4, 4]
        if mode == "overlap":
            nodes = np.concatenate((nodes, nodes))
        if mode == "gap":
            nodes["shape"][np.argmax(nodes["children"][:, 0] < 0)] = [1, 1, 1]
        lo, hi = np.zeros(3, np.int32), np.array([4, 4, 4], np.int32)
        c1, _ = O.gather(nodes, f.payload, lo, hi)
        o2 = np.empty((4, 4, 4), np.float32)
        c2 = lib.csg_gather(nodes.ctypes.data, len(nodes), f.payload.ctypes.data, len(f.payload), lo.ctypes.data,
                            hi.ctypes.data, o2.ctypes.data)
        assert c1 == c2 and c1 != 0, mode


def box_table(a, leaf):
    """A valid leaf-box table over `a`: bisection of the longest axis down to
    boxes of at most `leaf` samples, in pre-order; a leaf broadcasts every
    axis its samples are constant along and stores the rest."""
    rows, pay, used = [], [], [0]

    def build(lo, sh):
        i = len(rows)
        rows.append(None)
        if sh[0] * sh[1] * sh[2] <= leaf:
            sub = a[lo[0]:lo[0] + sh[0], lo[1]:lo[1] + sh[1], lo[2]:lo[2] + sh[2]]
            st = list(sh)
            for ax in range(3):
                first = np.take(sub, [0], axis=ax)
                if np.array_equal(sub.view(np.uint32), np.broadcast_to(first, sub.shape).view(np.uint32)):
                    st[ax] = 1
                    sub = first
            pay.append(np.ascontiguousarray(sub).ravel())
            rows[i] = (lo, sh, st, (-1, -1), used[0])
            used[0] += pay[-1].size
            return i
        ax = int(np.argmax(sh))
        ls, rs, rl = list(sh), list(sh), list(lo)
        ls[ax] = sh[ax] // 2
        rs[ax] -= ls[ax]
        rl[ax] += ls[ax]
        c0 = build(lo, ls)
        c1 = build(rl, rs)
        rows[i] = (lo, sh, [0, 0, 0], (c0, c1), -1)
        return i

    build([0, 0, 0], list(a.shape))
    nodes = np.zeros(len(rows), O.NODE)
    for i, (lo, sh, st, ch, off) in enumerate(rows):
        nodes[i] = (lo, sh, st, ch, off)
    return nodes, np.concatenate(pay).astype(np.float32)


@pytest.mark.parametrize("leaf", [1, 8, 32, 256, 1 << 20])
def test_csg_every_tile_read_path(fe, gpu, leaf):
    """Leaf sizes that send a tile's operand reads down every path of the
    device reader: a few leaves per 4x8x64 list tile, gathered at the tile's
    start (256 and one leaf for the whole grid), lists that overflow their 63
    entries so the read walks the tree from its start node (32 and 8; at 8
    the walk's leaves still fit the workgroup's list), and walks that overflow
    that list into the per-sample descent (1); three operands, so one read
    always comes after the gathered ones, and odd extents leave partial tiles
    on every axis.  Device and oracle run the same tables."""
    rng = np.random.default_rng(1000 + leaf)
    shape = (9, 21, 133)
    arrays = [rng.normal(size=shape).astype(np.float32) for _ in range(3)]
    arrays[1][:, :, 40:90] = 0.5  # constant boxes broadcast every axis
    arrays[2][:, 3:11, :] = arrays[2][:, 3:4, :]  # boxes constant along y
    prog = encode([("union", 0, 1), ("difference", 3, 2), ("offset", 4, 0.25)])
    tables = [box_table(a, leaf) for a in arrays]
    code, want = O.csg_execute(tables, prog, shape)
    assert code == 0
    # the program in closed form (ties give the second operand)
    a, b, c = arrays
    u = np.where(a < b, a, b)
    d = np.where(u > -c, u, -c)
    assert want.tobytes() == (d.astype(np.float64) - 0.25).astype(np.float32).tobytes()
    got = np.empty(shape, np.float32)
    lib = fe._lib.lib()
    nodes = (ct.c_void_p * 3)(*[t[0].ctypes.data for t in tables])
    pays = (ct.c_void_p * 3)(*[t[1].ctypes.data for t in tables])
    counts = np.array([len(t[0]) for t in tables], np.int64)
    lens = np.array([len(t[1]) for t in tables], np.int64)
    sh = np.array(shape, np.int32)
    assert lib.csg_execute(nodes, counts.ctypes.data, pays, lens.ctypes.data, 3, sh.ctypes.data, prog.ctypes.data,
                           len(prog), got.ctypes.data) == 0
    assert got.tobytes() == want.tobytes(), leaf


@pytest.mark.parametrize("leaf", [8, 4096, 1 << 17])
def test_unpack_box_tables(fe, gpu, leaf):
    """fe_unpack of explicit box tables: many tiny leaves (8), leaves handled
    by one workgroup each (4096) and leaves above 65536 samples spread over
    the whole grid (2^17), with leaves broadcasting every combination of axes
    (constant slabs along x, y and z) and odd extents."""
    import torch
    rng = np.random.default_rng(77 + leaf)
    shape = (40, 33, 301)
    a = rng.normal(size=shape).astype(np.float32)
    a[:, :, 100:220] = a[:, :, 100:101]  # constant along z
    a[10:35, 5:30, :] = a[10:35, 5:6, :]  # constant along y
    a[20:40, :, 250:] = 1.5                # constant along every axis
    nodes, payload = box_table(a, leaf)
    dn = torch.from_numpy(nodes.view(np.uint8).copy()).to(gpu)
    dp = torch.from_numpy(payload.view(np.uint32).copy()).to(gpu)
    out = torch.empty(shape, dtype=torch.float32, device=gpu)
    assert fe._lib.lib().fe_unpack(dn.data_ptr(), len(nodes), dp.data_ptr(), *shape, out.data_ptr(), None) == 0
    torch.cuda.synchronize()
    assert out.cpu().numpy().tobytes() == a.tobytes(), leaf
