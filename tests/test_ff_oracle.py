"""Pins of the oracle's double-word FP32 arithmetic (oracle_ff.hpp) and of the final-hit re-run
of Prhi in it (F_ACCURATE):

- add, sub, mul and div of random double-word values against exact rational arithmetic
  (fractions.Fraction): relative error below 2^-43; sqrt through its square;
- the re-run against the FP64 evaluation of the same intersector on the C0 hair at three ray
  lengths: the normal within 1e-6 rad and t within 2e-7 relative for every hit, where the
  FP32 intersector's normal error reaches 1e-3 rad (so a re-run that fell back to FP32 fails);
- trace with F_ACCURATE keeps every FP32 hit/miss and primitive decision."""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib as O


def _dw(rng, n, scale):
    hi = (rng.standard_normal(n) * scale).astype(np.float32)
    lo = (hi.astype(np.float64) * rng.uniform(-2 ** -25, 2 ** -25, n)).astype(np.float32)
    return np.stack([hi, lo], 1).astype(np.float32)


def _exact(p):
    return Fraction(float(p[0])) + Fraction(float(p[1]))


@pytest.mark.parametrize("op", [0, 1, 2, 3])
def test_double_word_ops_against_exact_rationals(op):
    rng = np.random.default_rng(op)
    n = 4000
    a, b = _dw(rng, n, 100.0), _dw(rng, n, 0.3)
    out = np.zeros_like(a)
    O.lib().orc_ff_eval(O.ptr(a), O.ptr(b), n, op, O.ptr(out))
    worst = 0.0
    for i in range(n):
        x, y = _exact(a[i]), _exact(b[i])
        want = x + y if op == 0 else x - y if op == 1 else x * y if op == 2 else x / y
        if want == 0:
            continue
        err = abs((_exact(out[i]) - want) / want)
        worst = max(worst, float(err))
        assert abs(float(out[i][1])) <= abs(float(out[i][0])) * 2 ** -23  # normalized pair
    print(f"op {op}: worst relative error {worst:.3g}")
    assert worst < 2 ** -43


def test_double_word_sqrt():
    rng = np.random.default_rng(9)
    n = 4000
    a = np.abs(_dw(rng, n, 50.0))
    a[:, 1] = np.where(a[:, 0] > 0, np.abs(a[:, 1]), 0)
    out = np.zeros_like(a)
    O.lib().orc_ff_eval(O.ptr(a), O.ptr(a), n, 4, O.ptr(out))
    worst = 0.0
    for i in range(n):
        x = _exact(a[i])
        if x == 0:
            continue
        z = _exact(out[i])
        worst = max(worst, float(abs(z * z - x) / x))
    print(f"sqrt: worst relative error of the square {worst:.3g}")
    assert worst < 2 ** -42


@pytest.mark.parametrize("dist", [(5.0, 10.0), (50.0, 60.0), (390.0, 400.0)])
def test_accurate_rerun_matches_fp64(dist):
    import scenes
    curves = scenes.c0_curves()
    rays = scenes.targeted_rays(curves, 3000, seed=2, dist=dist)
    S = O.Scene()
    S.add(O.MODE_PHANTOM, curves, per_point=True)
    h32 = S.trace(rays)
    hac = S.trace(rays, flags=O.F_ACCURATE)
    assert np.array_equal(h32["prim"], hac["prim"]) and np.array_equal(h32["inst"], hac["inst"])
    rel, ang, ang32 = [], [], []
    for i in np.nonzero(h32["prim"] != 0xFFFFFFFF)[0][:1200]:
        r64 = O.prhi_f64(rays[i], curves[h32["prim"][i]], per_point=True)
        if r64[0] <= 0:
            continue
        rel.append(abs(float(hac["t"][i]) - r64[0]) / r64[0])
        n64 = r64[5:8]
        for h, acc in ((hac, ang), (h32, ang32)):
            n = h["n"][i].astype(np.float64)
            acc.append(math.acos(min(1.0, abs(n @ n64) / np.linalg.norm(n) / np.linalg.norm(n64))))
    rel, ang, ang32 = np.array(rel), np.array(ang), np.array(ang32)
    print(f"ray length {dist}: accurate t rel max {rel.max():.2g}, normal max {ang.max():.2g} rad "
          f"(FP32 normal p99 {np.percentile(ang32, 99):.2g})")
    assert len(rel) > 600
    assert rel.max() < 2e-7 and ang.max() < 1e-6
    assert np.percentile(ang32, 99) > 1e-4
