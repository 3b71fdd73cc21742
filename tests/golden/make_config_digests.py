"""Writes tests/golden/config_digests.json: the oracle's builds at the benchmark
configurations (cases in config_cases.py), as sha256 digests.

Per build: n_active and the sha256 of kind (int32), corner_min_abs (fp32),
the ascending active ids (int32) and the tiles (fp32 [n_active][B][B][B])
from oracle/fe_oracle.c (or_classify, or_compact, or_eval_tiles: the
restatement of FE/faltkarna_v1.py:128-190, 277, pinned by
tests/test_oracle_golden.py).  cfg1 and cfg2 add the cropped dense field
(or_reconstruct), its uint8 occupancy (dense < 0) and the int32 EDT of the
occupancy: the single-plane squares (or_edt_envelope,
FE/edt_native_linear_v1/edt.cu:8-29) and the paired squares with the signed
field at pitch 1 (or_edt_pair); cfg2 also the leaf-box pack (leaf 512) of its
dense field (or_pack, the restatement of
FE/adaptive_pack_native_v1/pack.cpp:18-55, without the 2^24-sample cap).  GPU tests compare the device results' digests with these:
a bit-exact check at full size.

  python tests/golden/make_config_digests.py [part ...]   (parts: cfg1 cfg2
  cfg2pack crowded cfg3 cfg5; default all; about 40 minutes on 8 cores)
"""
import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(HERE))
OUT = HERE / "config_digests.json"


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def oracle_build(recipe, n, B):
    import oracle as O
    g = O.grid_record(np.float32(0), np.float32(0), np.float32(0), np.float32(1.0), n, n, n, B,
                      np.float32(np.sqrt(3.0) * B))
    return O.build_sparse(np.ascontiguousarray(recipe).view(O.PRIM), g)


def build_record(recipe, n, B):
    kind, cma, ids, tiles = oracle_build(recipe, n, B)
    rec = dict(n=n, block=B, prims=int(len(recipe)), n_active=int(len(ids)), kind=sha(kind), cma=sha(cma),
               ids=sha(ids), tiles=sha(tiles))
    return rec, (kind, ids, tiles)


def full_record(recipe, n, B):
    """A build plus its dense field, occupancy and EDTs."""
    import oracle as O
    rec, (kind, ids, tiles) = build_record(recipe, n, B)
    dense = O.reconstruct(kind, ids, tiles, n, n, n, B)
    occ = (dense < 0).astype(np.uint8)
    signed, pair = O.edt_pair(occ, 1.0)
    rec.update(dense=sha(dense), occupancy=sha(occ), occupied=int(occ.sum()), edt_single=sha(O.edt_envelope(occ)),
               edt_pair=sha(pair), edt_signed=sha(signed))
    return rec


def save(data):
    """Writes the digests (a partial list is kept while a part runs, so an
    interrupted run resumes where it stopped)."""
    tmp = OUT.with_suffix(".tmp")
    tmp.write_text(json.dumps(data, indent=0) + "\n")
    tmp.replace(OUT)


def main(parts):
    from bench import load_package
    import config_cases as C
    R = load_package().recipes
    data = json.loads(OUT.read_text()) if OUT.is_file() else {}
    t0 = time.time()
    if "cfg1" in parts:
        data["cfg1"] = full_record(R.plain_part_128(1), 128, 8)
        save(data)
        print("cfg1", round(time.time() - t0, 1), flush=True)
    if "cfg2" in parts:
        data["cfg2"] = full_record(R.large_field_512(2), 512, 8)
        save(data)
        print("cfg2", round(time.time() - t0, 1), flush=True)
    if "cfg2pack" in parts:
        # the leaf-box pack (leaf 512) of cfg2's dense field: 2^27 samples, past
        # the reference packer's 2^24 cap (the oracle restates it uncapped)
        import oracle as O
        kind, ids, tiles = (lambda r: (r[0], r[2], r[3]))(oracle_build(R.large_field_512(2), 512, 8))
        dense = O.reconstruct(kind, ids, tiles, 512, 512, 512, 8)
        nodes, payload = O.pack(dense, 512)
        data["cfg2"].update(pack_nodes=sha(nodes), pack_payload=sha(payload), pack_count=int(len(nodes)),
                            pack_samples=int(len(payload)))
        save(data)
        print("cfg2pack", round(time.time() - t0, 1), flush=True)
    if "crowded" in parts:
        data["crowded"] = {name: build_record(rec, n, B)[0] for name, n, B, rec in C.crowded_cases(R)}
        print("crowded", round(time.time() - t0, 1), flush=True)
    if "cfg3" in parts:
        out = data.get("cfg3_partial", [])
        for i, (rec, n) in enumerate(R.batch_recipes(3, 256)):
            if i < len(out):
                continue
            out.append({f"B{B}": build_record(rec, n, B)[0] for B in C.CFG3_BLOCKS})
            if i % 16 == 15:
                data["cfg3_partial"] = out
                save(data)
                print("cfg3", i + 1, round(time.time() - t0, 1), flush=True)
        data["cfg3"] = out
        data.pop("cfg3_partial", None)
        save(data)
    if "cfg5" in parts:
        out = data.get("cfg5_partial", [])
        for i, rec in enumerate(R.served_requests(5, C.SERVED_COUNT)):
            if i < len(out):
                continue
            out.append(build_record(rec, 512, 8)[0])
            data["cfg5_partial"] = out
            save(data)
            print("cfg5", i + 1, round(time.time() - t0, 1), flush=True)
        data["cfg5"] = out
        data.pop("cfg5_partial", None)
    save(data)
    return 0


if __name__ == "__main__":
    raise SystemExit(main(sys.argv[1:] or ["cfg1", "cfg2", "cfg2pack", "crowded", "cfg3", "cfg5"]))
