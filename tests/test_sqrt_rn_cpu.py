"""The residual selection of sqrt_rn (csrc/fe_common.hpp), restated in numpy:
starting from either float within one ulp of the exact root (v_sqrt_f32's
bound), the branch-free sequence (rescale below 2^-96, two fma residuals
rounded to fp32, descale) must give the correctly rounded float32 root for
normal, subnormal and special inputs.  The device code itself is pinned by
the bit-exact field tests against the oracle."""
import numpy as np

T = np.float32(2.0 ** -96)


def _select(xs, s):
    u = s.view(np.uint32)
    s_dn = (u - np.uint32(1)).view(np.float32)
    s_up = (u + np.uint32(1)).view(np.float32)
    with np.errstate(all="ignore"):
        # s_dn*s and s_up*s are exact in float64 and x - product cancels
        # exactly; the fp32 fma result is that value rounded once
        r_dn = (xs.astype(np.float64) - s_dn.astype(np.float64) * s.astype(np.float64)).astype(np.float32)
        r_up = (xs.astype(np.float64) - s_up.astype(np.float64) * s.astype(np.float64)).astype(np.float32)
    r = np.where(r_dn <= 0, s_dn, s)
    return np.where(r_up > 0, s_up, r)


def _mismatches(x):
    x = x.astype(np.float32)
    scale = x < T
    with np.errstate(over="ignore"):
        xs = np.where(scale, x * np.float32(2.0 ** 32), x).astype(np.float32)
        exact = np.sqrt(xs.astype(np.float64))
    rn = np.sqrt(xs)
    below = np.where(rn.astype(np.float64) <= exact, rn, np.nextafter(rn, np.float32(0)))
    above = np.where(rn.astype(np.float64) >= exact, rn, np.nextafter(rn, np.float32(np.inf)))
    want = np.sqrt(x).view(np.uint32)
    bad = 0
    for s in (below, above):
        r = _select(xs, s.astype(np.float32))
        r = np.where(scale, r * np.float32(2.0 ** -16), r).astype(np.float32)
        bad += int(np.count_nonzero(r.view(np.uint32) != want))
    return bad


def test_sqrt_rn_selection_is_correctly_rounded():
    rng = np.random.default_rng(3)
    x = np.exp2(rng.uniform(-149, 127, 400_000)).astype(np.float32)
    assert _mismatches(x) == 0
    sub = np.arange(1, 200_000, dtype=np.uint32).view(np.float32)  # subnormals
    assert _mismatches(sub) == 0
    f = rng.uniform(2.0 ** -70, 2.0 ** 60, 200_000).astype(np.float32)
    sq = (f.astype(np.float64) ** 2).astype(np.float32)
    for d in (-1, 0, 1):  # exact squares and their neighbours
        assert _mismatches((sq.view(np.int32) + d).view(np.float32)) == 0
    special = np.array([0.0, np.inf, 2.0 ** -96, 2.0 ** -126, 1.0, 4.0, 2.0], np.float32)
    assert _mismatches(special) == 0
