"""IKARUS test trees: the reference's own self-test trees and seeded random
trees over every node type.  Test infrastructure only."""
import math
import random

import numpy as np


def selftest_v1(E):
    """FE/ikarus_v1/eval_warp.py:745-749 and its 4000 points (random.seed(3))."""
    c = E.cylinder(radius=12.5, height=40.0)
    b = E.box(half_extents=(8.0, 8.0, 8.0), center=(3.0, 0.0, 5.0))
    s = E.sphere(radius=6.0, center=(-4.0, 2.0, -3.0))
    tree = E.union(E.subtract(c, b, k=0.0), E.rotate(s, "y", 37.0), k=2.0)
    random.seed(3)
    pts = np.array([[random.uniform(-30, 30), random.uniform(-30, 30), random.uniform(-30, 30)]
                    for _ in range(4000)], dtype=np.float64)
    return tree, pts


def selftest_v2(E):
    """FE/ikarus_v1/eval_warp_uniform_v2.py:364-372 and its 4000 points (random.seed(7))."""
    c = E.cylinder(radius=12.5, height=40.0)
    b = E.box(half_extents=(8.0, 8.0, 8.0), center=(3.0, 0.0, 5.0))
    rf = E.rect_frustum(z0=-10.0, z1=30.0, hx0=20.0, hy0=15.0, hx1=12.0, hy1=9.0)
    tree = E.union(E.subtract(c, b, k=0.0), E.translate(rf, 40.0, 0.0, 0.0), k=0.0)
    random.seed(7)
    pts = np.array([[random.uniform(-30, 60), random.uniform(-30, 30), random.uniform(-30, 30)]
                    for _ in range(4000)], dtype=np.float64)
    return tree, pts


def linear_grid(shape=(9, 7, 8), origin=(-6.0, -5.0, -4.5), pitch=1.5, coef=(0.5, -0.25, 0.75), c0=-1.0):
    """Samples of the affine field c0 + coef . p, which trilinear sampling reproduces."""
    g = [origin[a] + np.arange(shape[a]) * pitch for a in range(3)]
    X, Y, Z = np.meshgrid(*g, indexing="ij")
    return (c0 + coef[0] * X + coef[1] * Y + coef[2] * Z).astype(np.float32)


def random_tree(E, rng, depth, grid=None):
    """A seeded random tree over every node type (grid leaves when `grid` is given)."""
    if depth == 0 or rng.random() < 0.25:
        kind = rng.integers(0, 7 if grid is not None else 6)
        c = rng.uniform(-8, 8, 3)
        if kind == 0:
            return E.sphere(rng.uniform(1, 6), c)
        if kind == 1:
            return E.box(rng.uniform(0.5, 5, 3), c)
        if kind == 2:
            return E.cylinder(rng.uniform(0.5, 4), rng.uniform(1, 10), "xyz"[rng.integers(3)], c)
        if kind == 3:
            return E.cone(rng.uniform(0.5, 4), rng.uniform(0.2, 4), rng.uniform(1, 10), "xyz"[rng.integers(3)], c)
        if kind == 4:
            z0 = rng.uniform(-8, 0)
            return E.rect_frustum(z0, z0 + rng.uniform(1, 10), *rng.uniform(0.5, 5, 4), *rng.uniform(-2, 2, 4))
        if kind == 5:
            return E.halfspace(rng.normal(size=3), rng.uniform(-5, 5))
        return grid
    kind = rng.integers(0, 8)
    if kind <= 2:
        op = (E.union, E.intersect, E.subtract)[kind]
        k = 0.0 if rng.random() < 0.5 else rng.uniform(0.1, 3)
        return op(random_tree(E, rng, depth - 1, grid), random_tree(E, rng, depth - 1, grid), k=k)
    child = random_tree(E, rng, depth - 1, grid)
    if kind == 3:
        m = E.mat4_mul(E.mat4_translate(*rng.uniform(-3, 3, 3)),
                       E.mat4_rot_axis("xyz"[rng.integers(3)], rng.uniform(-180, 180)))
        return E.transform(child, m)
    if kind == 4:
        return E.rotate(child, "xyz"[rng.integers(3)], rng.uniform(-90, 90), about=tuple(rng.uniform(-2, 2, 3)))
    if kind == 5:
        return E.axis_project(child, int(rng.integers(3)), rng.uniform(-3, 3))
    if kind == 6:
        return E.offset(child, rng.uniform(-1.5, 1.5))
    return E.scale(child, rng.uniform(0.5, 2.0))


def chain(E, n, right=True):
    """A union chain of n spheres, right- or left-deep (value-stack depth test)."""
    leaves = [E.sphere(1.0 + 0.1 * i, (math.cos(i), math.sin(i), 0.3 * i)) for i in range(n)]
    node = leaves[-1] if right else leaves[0]
    for s in (reversed(leaves[:-1]) if right else leaves[1:]):
        node = E.union(s, node) if right else E.union(node, s)
    return node


def balanced(E, n_levels):
    """A full binary union tree with 2**n_levels sphere leaves (needs n_levels + 1 slots)."""
    def build(level, idx):
        if level == 0:
            return E.sphere(0.5 + 0.01 * idx, (idx * 0.1, 0.0, 0.0))
        return E.union(build(level - 1, 2 * idx), build(level - 1, 2 * idx + 1))
    return build(n_levels, 0)


def sampled_leaf(E, sdf, lo, pitch, shape, mode="sdf", iso=0.5):
    """A grid leaf holding the samples of a float64 field `sdf(X, Y, Z)` at
    lo + pitch*i (occupancy mode: 1 inside, 0 outside)."""
    g = [lo[a] + np.arange(shape[a]) * pitch for a in range(3)]
    X, Y, Z = np.meshgrid(*g, indexing="ij")
    v = sdf(X, Y, Z)
    if mode != "sdf":
        v = (v < 0).astype(np.float64)
    return E.grid_leaf(v.astype(np.float32), lo, pitch, mode=mode, iso=iso)


def sphere_sdf(c, r):
    return lambda X, Y, Z: np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) - r


def box_sdf(c, h):
    def f(X, Y, Z):
        q = [np.abs(X - c[0]) - h[0], np.abs(Y - c[1]) - h[1], np.abs(Z - c[2]) - h[2]]
        out = np.sqrt(sum(np.maximum(a, 0.0) ** 2 for a in q))
        return out + np.minimum(np.maximum(q[0], np.maximum(q[1], q[2])), 0.0)
    return f


def leaf_set(E):
    """Five grid leaves of different shapes, pitches and modes."""
    return [sampled_leaf(E, sphere_sdf((0.0, 0.0, 0.0), 5.0), (-7.0, -7.0, -7.0), 0.5, (29, 29, 29)),
            sampled_leaf(E, box_sdf((4.0, 1.0, 0.5), (3.0, 2.5, 4.0)), (-1.0, -3.0, -5.0), 0.25, (41, 27, 45)),
            sampled_leaf(E, sphere_sdf((6.0, 0.0, 1.0), 3.0), (2.0, -4.0, -3.0), 0.4, (21, 21, 21), mode="occupancy"),
            sampled_leaf(E, box_sdf((-3.0, -2.0, 0.0), (2.0, 2.0, 2.0)), (-6.0, -5.0, -3.0), 0.3, (21, 21, 21)),
            sampled_leaf(E, sphere_sdf((40.0, 0.0, 0.0), 2.0), (37.0, -3.0, -3.0), 0.5, (13, 13, 13))]
