"""The flocking terms' cone test decides `dot3(fwd, off / d) > cos_max` from an
estimate, fma-dot(fwd, off) times a hardware reciprocal (of d, GPU rule) or
reciprocal square root (of d2, OpenSteer rule), and computes the exact form
only when the estimate lies within E = 2^-18 (|fx| + |fy| + |fz|) + 2^-100 of
a threshold (gs_ops.hpp, flock_begin / flock_accumulate).  That is exact only
if |estimate - exact| < E for every input the fast path accepts (d2 in
[2^-80, 2^80], forward sum <= 2^40).  This test checks the claimed bound,
12 u |fwd| with u = 2^-24, on float32 arithmetic emulated with numpy, the
reciprocals taken one ulp off their correctly rounded value in a random
direction (worse than the hardware's: rcp and rsq are within 1 ulp of the
exact value).  The fused multiply-adds are emulated in float64 and rounded
once more to float32, which can add one ulp of the partial sum per fma; the
asserted bound allows 2 u for that."""
import numpy as np

U = 2.0 ** -24
N = 400_000


def _f32(x):
    return np.asarray(x, dtype=np.float64).astype(np.float32)


def _fma(a, b, c):
    return _f32(a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64))


def _one_ulp_off(x, g):
    up = g.random(x.shape) < 0.5
    return np.where(up, np.nextafter(x, np.float32(np.inf)), np.nextafter(x, np.float32(-np.inf))).astype(np.float32)


def _inputs(seed):
    g = np.random.default_rng(seed)
    off = g.standard_normal((N, 3)) * (2.0 ** g.uniform(-35.0, 35.0, (N, 1)))
    # some components tiny or zero, some pairs along an axis
    pick = g.random((N, 3))
    off = np.where(pick < 0.05, off * 2.0 ** -30, off)
    off = np.where(pick > 0.97, 0.0, off)
    off = _f32(off)
    fwd = g.standard_normal((N, 3))
    fwd /= np.linalg.norm(fwd, axis=1, keepdims=True)
    fwd *= 2.0 ** g.uniform(-10.0, 10.0, (N, 1))
    fwd = np.where(g.random((N, 3)) < 0.03, 0.0, fwd)
    return g, off, _f32(fwd)


def _check(rule, seed):
    g, off, fwd = _inputs(seed)
    ox, oy, oz = off[:, 0], off[:, 1], off[:, 2]
    fx, fy, fz = fwd[:, 0], fwd[:, 1], fwd[:, 2]
    with np.errstate(over="ignore", under="ignore", invalid="ignore", divide="ignore"):
        d2 = (ox * ox + oy * oy) + oz * oz  # float32, the kernel's order
        keep = (d2 >= np.float32(2.0 ** -80)) & (d2 <= np.float32(2.0 ** 80))
        d = np.sqrt(d2)  # correctly rounded float32
        exact = ((fx * (ox / d)) + (fy * (oy / d))) + (fz * (oz / d))
        dotf = _fma(fz, oz, _fma(fy, oy, fx * ox))
        if rule == "gpuref":
            r = _one_ulp_off(_f32(1.0 / d.astype(np.float64)), g)
        else:
            r = _one_ulp_off(_f32(1.0 / np.sqrt(d2.astype(np.float64))), g)
        est = dotf * r
    fsum = (np.abs(fx) + np.abs(fy)) + np.abs(fz)
    fnorm = np.sqrt(fx.astype(np.float64) ** 2 + fy.astype(np.float64) ** 2 + fz.astype(np.float64) ** 2)
    err = np.abs(est.astype(np.float64) - exact.astype(np.float64))
    assert keep.sum() > N // 2
    err, fsum, fnorm = err[keep], fsum[keep].astype(np.float64), fnorm[keep]
    assert np.all(np.isfinite(err))
    # the claimed bound (+2 u of emulation slack) and the margin the kernel uses
    assert np.all(err <= 14.0 * U * fnorm + 2.0 ** -120), np.max(err / (U * fnorm + 1e-300))
    assert np.all(err < 2.0 ** -18 * fsum + 2.0 ** -100)
    return float(np.max(err / (U * np.maximum(fnorm, 1e-300))))


def test_cone_estimate_bound_opensteer_rsq():
    worst = max(_check("opensteer", s) for s in range(3))
    print(f"rsq(d2) estimate: worst error {worst:.2f} u |fwd|")


def test_cone_estimate_bound_gpuref_rcp():
    worst = max(_check("gpuref", s) for s in range(3))
    print(f"rcp(d) estimate: worst error {worst:.2f} u |fwd|")


def test_bound_catches_a_wrong_estimate():
    """The check is sharp enough to see a plausible mistake: the reciprocal
    square root of d instead of d2 (a dropped square) fails it."""
    g, off, fwd = _inputs(7)
    ox, oy, oz = off[:, 0], off[:, 1], off[:, 2]
    fx, fy, fz = fwd[:, 0], fwd[:, 1], fwd[:, 2]
    with np.errstate(over="ignore", under="ignore", invalid="ignore", divide="ignore"):
        d2 = (ox * ox + oy * oy) + oz * oz
        keep = (d2 >= np.float32(2.0 ** -80)) & (d2 <= np.float32(2.0 ** 80))
        d = np.sqrt(d2)
        exact = ((fx * (ox / d)) + (fy * (oy / d))) + (fz * (oz / d))
        wrong = _fma(fz, oz, _fma(fy, oy, fx * ox)) * _f32(1.0 / np.sqrt(d.astype(np.float64)))
    fsum = ((np.abs(fx) + np.abs(fy)) + np.abs(fz)).astype(np.float64)
    err = np.abs(wrong.astype(np.float64) - exact.astype(np.float64))[keep]
    assert not np.all(err < 2.0 ** -18 * fsum[keep] + 2.0 ** -100)
