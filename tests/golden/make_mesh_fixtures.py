"""Writes tests/golden/mesh_stage.npz and mesh_stage.json from the reference's
mesh-stage captures (reports/cuda_columns.json and cuda_columns_arrays.npz,
written by probes/field_engine/field_cuda_columns_probe.py).  Run in a
container that holds the reference:

    python tests/golden/make_mesh_fixtures.py [/path/to/reference]

Only data is taken: the per-case sha256 digests of the reference's outputs,
the plate's column hits (the plate mesh itself is an external STL that the
reference does not hold), and the grid parameters of each case.  The box
meshes of the rotated, open and capacity cases are rebuilt here the way the
probe builds them (probes/field_engine/field_mesh_native_mask_stage_probe.py:23-33,
field_cuda_columns_probe.py:77-84): a unit box with trimesh's fixed corner and
face tables scaled to the extents, rotated about [1, 2, 3] by 0.37 rad with
the homogeneous matrix of the textbook axis-angle formula (applied with
np.dot, as trimesh's transform_points does), then translated.  The rebuilt
vertices are stored, so tests do not depend on this machine's BLAS; the
tests check them by reproducing the reference's hit digests exactly.
"""
import json
import math
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent

# trimesh.creation.box: corners of the unit cube (then -0.5 and scaled) and its
# 12 triangles with their outward normals.
BOX_CORNERS = np.array([0, 0, 0, 0, 0, 1, 0, 1, 0, 0, 1, 1, 1, 0, 0, 1, 0, 1, 1, 1, 0, 1, 1, 1],
                       dtype=np.float64).reshape(-1, 3)
BOX_FACES = np.array([1, 3, 0, 4, 1, 0, 0, 3, 2, 2, 4, 0, 1, 7, 3, 5, 1, 4, 5, 7, 1, 3, 7, 2, 6, 4, 2, 2, 7, 6, 6, 5,
                      4, 7, 5, 6], dtype=np.int64).reshape(-1, 3)
BOX_NORMAL_Z = np.array([0, 0, 0, -1, 1, 0, 1, 0, -1, 0, 0, 0], dtype=np.float64)


def box(extents):
    v = BOX_CORNERS - 0.5
    v = v * np.asarray(extents, dtype=np.float64)
    return v, BOX_FACES.copy()


def rotation(angle, direction):
    """4x4 rotation about a unit axis: cos*I + (1-cos)*d d^T + sin*[d]x."""
    sina, cosa = np.sin(angle), np.cos(angle)
    d = np.array(direction[:3], dtype=np.float64)
    d /= math.sqrt(np.dot(d, d))
    m = np.diag([cosa, cosa, cosa, 1.0])
    m[:3, :3] += np.outer(d, d) * (1.0 - cosa)
    d = d * sina
    m[:3, :3] += np.array([[0.0, -d[2], d[1]], [d[2], 0.0, -d[0]], [-d[1], d[0], 0.0]])
    return m


def transform(v, m):
    return np.dot(m, np.column_stack((v, np.ones(len(v)))).T).T[:, :3]


def translate(v, t):
    m = np.eye(4)
    m[:3, 3] = t
    return transform(v, m)


def main(ref):
    report = json.loads((ref / "reports/cuda_columns.json").read_text())
    arrays = np.load(ref / "reports/cuda_columns_arrays.npz", allow_pickle=False)
    rows = {r["case"]: r for r in report["rows"]}
    out, meta = {}, {"cases": {}, "capacity": {}}
    for name, row in rows.items():
        legs = row["legs"]
        if legs[0]["hashes"] != legs[1]["hashes"]:
            raise SystemExit(f"{name}: the reference's two legs differ")
        meta["cases"][name] = {"digests": legs[0]["hashes"], "method": row["method"], "closed": row["closed"],
                               "shape": list(arrays[f"{name}_0_solid"].shape),
                               "gmin": arrays[f"{name}_0_gmin"].tolist()}
    for p in ("2.0", "1.0", "0.5"):
        name = f"plate_{p}"
        out[f"{name}_columns"] = arrays[f"{name}_0_columns"].astype(np.int32)
        out[f"{name}_depths"] = arrays[f"{name}_0_depths"]
        out[f"{name}_signs"] = arrays[f"{name}_0_signs"].astype(np.int8)
        # The plate spans z in [-3, 3] (every stored depth is +-3 to 3e-13) and
        # its grid origin is lo = vertices.min(0) - 3 pitch, so the local
        # anchor's z is -3 - 3 pitch + gmin_z pitch.
        pitch = float(p)
        gz = int(arrays[f"{name}_0_gmin"][2])
        meta["cases"][name].update(pitch=pitch, anchor_z=(-3.0 - 3 * pitch) + gz * pitch)
    base, faces = box([4, 3, 2])
    rotated = translate(transform(base, rotation(.37, [1, 2, 3])), [.13, -.21, .37])
    for shift in (0.0, 1e6):
        v = translate(rotated, [shift] * 3) if shift else rotated
        name = f"rotated_{shift}"
        out[f"{name}_vertices"] = v
        out[f"{name}_faces"] = faces
        meta["cases"][name].update(pitch=.25, lo=(v.min(0) - 1).tolist())
    v, faces = box([4, 4, 4])
    opened = faces[BOX_NORMAL_Z < .5]
    out["open_vertices"] = v
    out["open_faces"] = opened
    out["reversed_open_vertices"] = v
    out["reversed_open_faces"] = opened[:, ::-1].copy()
    for name in ("open", "reversed_open"):
        meta["cases"][name].update(pitch=.25, lo=[-4.031, -4.043, -4.057])
    out["capacity_vertices"] = v
    out["capacity_faces"] = faces
    cap = report["capacity_rows"][0]
    meta["capacity"] = {"stack": 32, "step": 10.0, "xy": [.123, .234], "digests": cap["hashes"],
                        "hits": cap["hits"], "analytic_depth_differences": cap["analytic_depth_differences"],
                        "analytic_max_delta": cap["analytic_max_delta"]}
    np.savez_compressed(HERE / "mesh_stage.npz", **out)
    (HERE / "mesh_stage.json").write_text(json.dumps(meta, indent=1) + "\n")


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
