"""Writes tests/golden/edt_volumes.json: the oracle's paired EDT of the thin-feature
obstacle volumes (one 768^3 volume, seed 4, and the 64-volume uneven batch).

For every volume it records the shape, the occupied voxel count, the sha256 of
the uint8 mask and the sha256 of the oracle's int32 squares (2,nx,ny,nz) from
oracle/fe_oracle.c or_edt_pair (the envelope restatement of
FE/edt_native_linear_v1/edt.cu:8-29, itself pinned to the reference's EDT
captures by tests/test_oracle_golden.py).  GPU tests compare the device pair's
digest with these, which is a bit-exact check at full size.

  python tests/golden/make_edt_volumes.py      (about a minute on 8 cores)
"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def record(vol):
    import oracle as O
    _, sq = O.edt_pair(vol, 1.0)
    return dict(shape=list(vol.shape), occupied=int(np.count_nonzero(vol)), mask_sha=sha(vol), pair_sha=sha(sq))


def main():
    from bench import load_package
    fe = load_package()
    R = fe.recipes
    meta = dict(seed=4, volume_768=record(R.thin_feature_volume_768(4)),
                batch=[record(v) for v in R.thin_feature_batch(4, 64)])
    (Path(__file__).resolve().parent / "edt_volumes.json").write_text(json.dumps(meta, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
