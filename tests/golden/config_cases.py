"""Build cases at the benchmark configurations, and recipes made to reach the
build's list fallbacks.  TEST DATA: imported by make_config_digests.py (which
writes config_digests.json from the oracle) and by the GPU tests.

Grids are cubic at 1 mm with the origin at 0 (the benchmark's grids).

* cfg1: the 128^3 part (40 spheres + 24 boxes, seed 1), block 8, and the
  int32 EDT of its occupancy;
* cfg2: 1024 mixed primitives in 512^3 (seed 2), block 8, and the paired EDT
  of its occupancy;
* cfg3: the 256 uneven recipes (seed 3; 4..2048 primitives, 64..384^3) at
  each block edge 4, 8, 16 (whichever the autotuner picks is covered);
* cfg5: the first served requests (seed 5; the large-field distribution).

The crowded recipes put hundreds of small spheres into one superblock, so
that its culled list is longer than the 512-entry list capacity (the
superblock then stores no list and its blocks refine from the full recipe);
with more than 2048 primitives the tile pass also skips the per-block
refinement and folds the whole recipe.
"""
import numpy as np

CFG3_BLOCKS = (4, 8, 16)
SERVED_COUNT = 48


def grid(n, block):
    return (0.0, 0.0, 0.0, 1.0, n, n, n, block)


def crowded_recipe(R, seed, n_prims, n_cluster, n=96, lo=40.0, hi=56.0):
    """n_cluster spheres of radius U[1, 3] with centres in [lo, hi)^3 (inside
    one superblock at every block edge for n = 96), unions, followed by
    n_prims - n_cluster mixed primitives over the domain (the last eighth of
    them subtractions)."""
    rng = np.random.default_rng(seed)
    recs = [(R.SPHERE, R.UNION, tuple(rng.uniform(lo, hi, 3)), (float(rng.uniform(1.0, 3.0)),))
            for _ in range(n_cluster)]
    cluster = R.make(recs)
    rest = R.mixed_recipe(rng, n_prims - n_cluster, n, spread=0.8)
    return np.concatenate([cluster, rest])


def crowded_cases(R):
    """(name, n, block, recipe): superblock lists over capacity (1024
    primitives) and recipes beyond the refined tile pass (2500 primitives)."""
    out = []
    for B in (4, 8, 16):
        out.append((f"crowded1024_B{B}", 96, B, crowded_recipe(R, 60 + B, 1024, 700)))
        out.append((f"crowded2500_B{B}", 96, B, crowded_recipe(R, 70 + B, 2500, 2200)))
    return out
