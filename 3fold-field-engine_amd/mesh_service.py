"""Result layout of the mesh stage's shared service (FE/mesh_field_shared_service_v1.py:16-19).

One byte region holds a query's three arrays back to back: the solid mask
(N bytes, bool), the surface mask (N bytes, bool), zero padding up to the next
64-byte boundary, then the signed distance (N float32), N = voxels of the
grid.  Any process that maps such a region reads the arrays through
`result_views`; `fill_region` writes one query of a prepared mesh stage into
a region (the device copies land in it directly).  The service process around
the region (spawning, leases, IPC) is not part of this package.
"""
import numpy as np

ALIGN = 64


def layout(shape):
    """(voxels, byte offset of the distance array, region bytes) for a grid shape."""
    n = 1
    for s in shape:
        n *= int(s)
    start = -(-2 * n // ALIGN) * ALIGN
    return n, start, start + 4 * n


def result_views(region, shape):
    """(solid, surface, distance) arrays of `shape` viewing a buffer of the
    layout's exact size; views of a read-only buffer are read-only."""
    n, start, total = layout(shape)
    raw = np.frombuffer(region, dtype=np.uint8)
    if raw.size != total:
        raise ValueError(f"region of {total} bytes required for shape {tuple(shape)}, got {raw.size}")
    shape = tuple(int(s) for s in shape)
    solid = raw[:n].view(np.bool_).reshape(shape)
    surface = raw[n:2 * n].view(np.bool_).reshape(shape)
    distance = raw[start:].view(np.float32).reshape(shape)
    return solid, surface, distance


def fill_region(field, region):
    """Runs one query of a PreparedMeshField into a writable region of its
    grid's layout (padding left as it is); returns the three views."""
    solid, surface, distance = result_views(region, field.shape)
    if not solid.flags.writeable:
        raise ValueError("writable region required")
    field.evaluate_into(solid, surface, distance)
    return solid, surface, distance
