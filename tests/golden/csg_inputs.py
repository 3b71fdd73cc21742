"""Closed-form CSG operand fields of the reference's CSG probes, restated.

The reference's frozen CSG outputs (reports/field_csg_curved_arrays.npz,
digests in golden.json) were computed from these inputs:
probes/field_engine/field_csg_probe.py:20-33 (BOXES, box_distance) and
probes/field_engine/field_csg_curved_probe.py:28-36 (distances).  The same
numpy float64 formulas are written out here, rounded to float32 once.
"""
import numpy as np

BOXES = [((0, 0, 0), (20, 6, 4)), ((0, 0, 0), (4, 6, 16)), ((0, 0, 0), (20, 12, 2)),
         ((5, 3, -2), (8, 5, 6)), ((12, 3, -2), (15, 5, 6))]

PROGRAMS = {False: [('union', 0, 1), ('union', 5, 2), ('union', 3, 4), ('difference', 6, 7)],
            True: [('union', 0, 1), ('difference', 3, 2)]}
NAMES = {False: 'bracket_plate_rectangular_holes', True: 'sphere_plate_round_hole'}


def box_distance(points, bounds):
    lo, hi = np.asarray(bounds[0]), np.asarray(bounds[1])
    q = np.abs(points - (lo + hi) / 2) - (hi - lo) / 2
    return np.linalg.norm(np.maximum(q, 0), axis=-1) + np.minimum(q.max(axis=-1), 0)


def distances(points, curved):
    if not curved:
        return [box_distance(points, b) for b in BOXES]
    sphere = np.linalg.norm(points - np.array([10, 6, 6]), axis=-1) - 6
    plate = box_distance(points, ((0, 0, 0), (20, 12, 2)))
    q = np.stack((np.linalg.norm(points[:, :2] - [10, 6], axis=-1) - 1.5, np.abs(points[:, 2] - 6) - 10), axis=-1)
    cylinder = np.linalg.norm(np.maximum(q, 0), axis=-1) + np.minimum(q.max(axis=-1), 0)
    return [sphere, plate, cylinder]


def case(curved, pitch):
    """(input float32 fields, program, origin, pitch, shape) of one frozen case."""
    origin = np.array([-2., -2., -4.])
    shape = tuple(np.rint(np.array([24, 16, 22]) / pitch).astype(int) + 1)
    indices = np.indices(shape).reshape(3, -1).T
    fields = [a.reshape(shape).astype(np.float32) for a in distances(origin + indices * pitch, curved)]
    return fields, PROGRAMS[curved], origin, pitch, shape


def query_points(shapes):
    """Query points of probes/field_engine/field_native_pack_probe.py:28-34 (rng 445, one draw per field)."""
    rng = np.random.default_rng(445)
    return [rng.uniform(-1, np.array(s) + 1, (4097, 3)) for s in shapes]
