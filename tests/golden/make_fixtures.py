"""Writes the golden fixtures under tests/golden/ from data files the reference holds.

Run in a container where the reference is mounted:  python tests/golden/make_fixtures.py
It reads the reference's frozen .npz/.json captures as data (numpy.load with
allow_pickle=False, json) and writes compact fixtures: input masks as packed
bits, expected integer arrays compressed, and sha256 digests of the expected
float outputs (a digest comparison is a bit-exact check).  It imports no
reference code.

Sources (paths relative to the reference root):
  artifacts/field_edt_integer_v1_arrays.npz   16 mask families, int64 squares of both
                                              polarities, signed at 3 pitches
                                              (probes/field_engine/field_edt_integer_probe.py:22-47)
  reports/field_quotient_pair_arrays.npz      9 pair cases, int32 squares (2,nx,ny,nz) and
                                              signed at 5 pitches
  reports/field_linear_edt_arrays.npz         9 further random/plate cases, squares + signed at 3 pitches
  reports/edt_large_arrays.npz                plate 2/1/0.5 mm, single site 512x2x1, checker 129x3x7
                                              (pitches from probes/field_engine/field_edt_large_probe.py:21-24)
  reports/field_csg_curved_arrays.npz         frozen CSG outputs (digests only; inputs are the
                                              closed forms restated in csg_inputs.py)
  reports/field_native_pack.json              node/payload/origin sha256 of the packed outputs
  reports/field_native_pack_arrays.npz        4097 trilinear query results per field
  src/field_engine/artifacts/faltkarna_v1_selftest.json   test-field known answers
"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def packmask(m):
    return np.packbits(np.ascontiguousarray(m, dtype=bool).ravel())


def edt_integer():
    z = np.load(REF / "artifacts/field_edt_integer_v1_arrays.npz", allow_pickle=False)
    names = sorted({k.rsplit("_square_", 1)[0] for k in z.files if "_square_" in k})
    out, meta = {}, {}
    for i, name in enumerate(names):
        s0, s1 = z[f"{name}_square_0"], z[f"{name}_square_1"]
        mask = s0 == 0
        out[f"c{i}_mask"] = packmask(mask)
        out[f"c{i}_sq0"] = s0.astype(np.int32)
        out[f"c{i}_sq1"] = s1.astype(np.int32)
        meta[f"c{i}"] = dict(name=name, shape=list(mask.shape),
                             signed={repr(p): sha(z[f"{name}_signed_{j}"])
                                     for j, p in enumerate((0.125, 0.3, 0.7483641718372458))})
    np.savez_compressed(OUT / "edt_integer.npz", **out)
    return meta


def edt_pairs():
    out, meta = {}, {}
    z = np.load(REF / "reports/field_quotient_pair_arrays.npz", allow_pickle=False)
    cases = sorted({k.rsplit("_0_", 1)[0] for k in z.files if "_0_" in k and not k.endswith("_squared")})
    for i, name in enumerate(cases):
        pitches = ["1e-12", "0.125", "0.7", "2.0", "1000000000000.0"]
        sq = z[f"{name}_0_0.125_squared"]
        mask = sq[1] > 0
        key = f"q{i}"
        out[f"{key}_mask"] = packmask(mask)
        meta[key] = dict(name=name, source="field_quotient_pair", shape=list(mask.shape),
                         squared0=sha(sq[0]), squared1=sha(sq[1]),
                         signed={p: sha(z[f"{name}_0_{p}"]) for p in pitches})
    z = np.load(REF / "reports/field_linear_edt_arrays.npz", allow_pickle=False)
    cases = sorted({k.rsplit("_0_squared_", 1)[0] for k in z.files if "_0_squared_" in k})
    for i, name in enumerate(cases):
        # probes/field_engine/field_linear_edt_probe.py:55 stores squared_distance(mask) as
        # squared_0 and squared_distance(~mask) as squared_1.
        s_in, s_out = z[f"{name}_0_squared_0"], z[f"{name}_0_squared_1"]
        mask = s_in != 0
        key = f"l{i}"
        out[f"{key}_mask"] = packmask(mask)
        meta[key] = dict(name=name, source="field_linear_edt", shape=list(mask.shape),
                         squared0=sha(s_out), squared1=sha(s_in),
                         signed={p: sha(z[f"{name}_0_signed_{p}"]) for p in ("0.125", "0.7", "2.0")})
    z = np.load(REF / "reports/edt_large_arrays.npz", allow_pickle=False)
    for i, (name, pitch) in enumerate((("plate_2.0", 2.0), ("plate_1.0", 1.0), ("plate_0.5", 0.5),
                                       ("single_site", 1.0), ("checker", 0.25))):
        s0, s1 = z[f"{name}_0_squared0"], z[f"{name}_0_squared1"]
        mask = s0 == 0
        key = f"g{i}"
        out[f"{key}_mask"] = packmask(mask)
        meta[key] = dict(name=name, source="edt_large", shape=list(mask.shape),
                         squared0=sha(s0), squared1=sha(s1), signed={repr(pitch): sha(z[f"{name}_0_signed"])})
    np.savez_compressed(OUT / "edt_pairs.npz", **out)
    return meta


def csg_and_pack():
    z = np.load(REF / "reports/field_csg_curved_arrays.npz", allow_pickle=False)
    rep = json.loads((REF / "reports/field_native_pack.json").read_text())
    q = np.load(REF / "reports/field_native_pack_arrays.npz", allow_pickle=False)
    meta, out = {}, {}
    for name in z.files:
        if not name.endswith("_0_field"):
            continue
        hashes = [r["hashes"] for r in rep["rows"] if r["case"] == name][0]
        meta[name] = dict(shape=list(z[name].shape), field_sha=sha(z[name]), pack_hashes=hashes)
        out[name + "_query"] = q[name + "_0"]
    np.savez_compressed(OUT / "pack_query.npz", **out)
    return meta


def testfield():
    s = json.loads((REF / "src/field_engine/artifacts/faltkarna_v1_selftest.json").read_text())
    m = s["substrat_minne"]
    return dict(pitch=0.25, n_blocks=m["n_blocks"], n_active=m["n_active"], n_interior=m["n_interior"],
                n_exterior=m["n_exterior"], active_voxels=m["active_voxels"], dense_voxels=m["dense_voxels"],
                vol_voxel_mm3=s["fallbevis_iii_normal"]["vol_voxel_mm3"],
                vol_corrupt_mm3=s["fallbevis_i_korrupt_index"]["vol_voxel_mm3"],
                checksum_i64=s["fallbevis_ii_determinism"]["i64_fixpunkt_safe"]["varden"][0],
                n_active_pitch_0_9=120, n_active_pitch_0_09=43658,
                recorded_dense_sha16_rtx5070="49100bcf61ab5a5f")


def main():
    if not REF.is_dir():
        print("reference not mounted; fixtures unchanged", file=sys.stderr)
        return 1
    meta = dict(edt_integer=edt_integer(), edt_pairs=edt_pairs(), csg_pack=csg_and_pack(),
                testfield=testfield())
    (OUT / "golden.json").write_text(json.dumps(meta, indent=1, sort_keys=True) + "\n")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
