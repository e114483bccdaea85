"""Pins of the oracle functions on the headline path that no reference output
covers: the shared FP32 sin/cos, sphere-light sampling and pdf, the
cosine-weighted bounce estimator with MIS and Russian roulette, the bilinear
texture lookup, instanced FP64 rebasing and Kahan accumulation.

Each pin is a closed form, an independent library routine or an invariant
chosen so that a plausible slip (a dropped cosine, a wrong MIS weight or pdf,
a transposed operand, an off-by-one texel) fails it."""
import numpy as np

import pyoracle as O


def _ulp_err(got, ref):
    """|got - ref| in units of the FP32 spacing at ref (ref FP64)."""
    sp = np.spacing(np.abs(ref.astype(np.float32))).astype(np.float64)
    return np.abs(got.astype(np.float64) - ref) / np.maximum(sp, np.spacing(np.float32(0)))


def test_sincos_f32_against_fp64():
    """fp_sincos.hpp: Cody-Waite reduction + fitted polynomials, at most 2 ulp from
    FP64 sin/cos over the arguments the path takes (phi in [0, 2 pi), the
    concentric-disk angle in [-pi/4, 3 pi/4]) and well beyond; next to a zero of
    the function (results below 1e-6) the error is bounded absolutely by 2e-14,
    the FP32 tail of pi/2 times the quadrant count."""
    x = np.concatenate([np.linspace(-50.0, 50.0, 1_000_001), np.linspace(0.0, 2 * np.pi, 100_001),
                        np.array([0.0, -0.0, 1e-30, np.pi / 4, np.pi / 2, np.pi, 1e-4])]).astype(np.float32)
    s, c = O.sincos_f32(x)
    xd = x.astype(np.float64)
    for got, ref in ((s, np.sin(xd)), (c, np.cos(xd))):
        big = np.abs(ref) >= 1e-6
        assert _ulp_err(got[big], ref[big]).max() <= 2.0
        assert np.abs(got[~big] - ref[~big]).max() <= 2e-14
    # identities a quadrant or sign slip would break
    assert np.abs(s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2 - 1).max() < 3e-7
    s2, c2 = O.sincos_f32(-x)
    assert np.array_equal(s2, -s) and np.array_equal(c2, c)
    assert O.sincos_f32(np.float32([0.0]))[0][0] == 0.0 and O.sincos_f32(np.float32([0.0]))[1][0] == 1.0
