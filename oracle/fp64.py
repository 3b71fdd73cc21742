"""FP64 closed forms of the recipe primitives.  TEST INFRASTRUCTURE.

Used to report the error of stored fp32 field values against exact geometry
(sphere, box, capped cylinder, torus, infinite z-cylinder) and the ordered
min/max fold.  Parameters are the fp32 values stored in the recipe, widened.
"""
import numpy as np


def prim_fp64(p, pts):
    x = pts[:, 0] - float(p["cx"])
    y = pts[:, 1] - float(p["cy"])
    z = pts[:, 2] - float(p["cz"])
    a, b, c = float(p["a"]), float(p["b"]), float(p["c"])
    k = int(p["kind"])
    if k == 0:
        return np.sqrt(x * x + y * y + z * z) - a
    if k == 1:
        q = np.stack((np.abs(x) - a, np.abs(y) - b, np.abs(z) - c), -1)
        return np.linalg.norm(np.maximum(q, 0), axis=-1) + np.minimum(q.max(-1), 0)
    if k in (2, 3, 4):
        rad, ax = {4: (np.hypot(x, y), z), 3: (np.hypot(x, z), y), 2: (np.hypot(y, z), x)}[k]
        q = np.stack((rad - a, np.abs(ax) - b), -1)
        return np.linalg.norm(np.maximum(q, 0), axis=-1) + np.minimum(q.max(-1), 0)
    if k == 5:
        return np.hypot(np.hypot(x, y) - a, z) - b
    if k == 6:
        return np.hypot(x, y) - a
    raise ValueError(k)


def eval_fp64(prims, pts):
    pts = np.asarray(pts, np.float64)
    f = prim_fp64(prims[0], pts)
    for p in prims[1:]:
        v = prim_fp64(p, pts)
        op = int(p["op"])
        f = np.minimum(f, v) if op == 0 else (np.maximum(f, -v) if op == 1 else np.maximum(f, v))
    return f


def random_recipe(rng, n, extent):
    """Small mixed recipe inside [0, extent)^3 for oracle self-checks."""
    from oracle import PRIM
    prims = np.zeros(n, PRIM)
    for i in range(n):
        k = int(rng.integers(0, 6))
        prims[i]["kind"] = k
        prims[i]["op"] = 0 if i == 0 or rng.random() > 0.25 else 1
        prims[i]["cx"], prims[i]["cy"], prims[i]["cz"] = rng.uniform(0, extent, 3)
        s = rng.uniform(2, extent / 5, 3)
        if k == 5:
            s[1] = s[0] * rng.uniform(0.2, 0.5)
        prims[i]["a"], prims[i]["b"], prims[i]["c"] = s
    return prims
