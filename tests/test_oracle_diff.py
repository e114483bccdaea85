"""CPU checks of the FP64 differentiable direct-illumination oracle
(oracle/oracle_diff.hpp), which the GPU adjoint is compared with.

What pins it:
* forward: the FP64 model on hit records equals the oracle's FP32 path tracer
  at bounces = 0 (independent code: BVH traversal + processIntersection
  restatement) to FP32 rounding;
* adjoint (heights): central finite differences of the L2 loss, h = 1e-3 m,
  over every pixel a node can influence;
* adjoint (texels): the image is linear in the texels, so central differences
  are exact up to rounding;
* flat and smooth shading, and a sphere-draped DEM (radial d position / d height).
"""
import numpy as np
import pytest

import pyoracle as O
from vira_amd import scene as S


def _scene(smooth, kind=0, n=33, img=64, tex=32, seed=5, near_light=False):
    h = O.fbm_heights(n, n, seed, 500.0, 6, 8.0)
    extent = (n - 1) * 30.0
    if kind == 0:
        proj = O.make_projection(0, 30.0)
        cam = S.Camera(img, img, 40.0, (0.0, 0.0, 0.55 * extent / np.tan(np.radians(20))))
        cam.look_in_direction((0.0, 0.0, -1.0))
        sun = S.directional_sun(30.0, 20.0, 1.0)
    else:
        R = 50e3
        alt = 0.55 * extent / np.tan(np.radians(20))
        proj = O.make_projection(1, 30.0, R, (0.0, 0.0, 0.0), (R + alt, 0.0, 0.0))
        cam = S.Camera(img, img, 40.0, (0.0, 0.0, 0.0))
        cam.look_in_direction((-1.0, 0.0, 0.0), up=(0.0, 0.0, 1.0))
        sun = S.directional_sun(30.0, 20.0, 1.0, origin=np.array([-alt, 0.0, 0.0]), up=(1.0, 0.0, 0.0),
                                east=(0.0, 1.0, 0.0))
    if near_light:  # a point light close enough that its direction and distance vary over the DEM
        sun = S.PointLight((0.3 * extent, 0.2 * extent, 1.5 * extent), np.full(8, 1e7))
    om = O.dem_mesh(h, proj, None)
    rng = np.random.default_rng(seed)
    T = (0.11 * np.exp(0.25 * rng.standard_normal((tex, tex)))).astype(np.float32)
    mat = O.material(1.0, T, smooth)
    lights = [O.light_from(sun.to_c())]
    out = om.render(O.camera_from(cam.to_c()), lights, mat, O.options(1, 0, False, 1, 0), precision=32,
                    records=True)
    return h, proj, T, mat, lights, out


@pytest.mark.parametrize("smooth,kind,near", [(False, 0, False), (True, 0, False), (False, 1, False),
                                              (False, 0, True), (True, 0, True)])
def test_fp64_model_forward_and_adjoint(smooth, kind, near):
    h, proj, T, mat, lights, out = _scene(smooth, kind, near_light=near)
    dd = O.DemDiff(h, proj)
    rec = out["records"]
    assert (rec[0, :, 0] != 0xFFFFFFFF).sum() > 2000 and (rec[0, :, 3] & 1).sum() > 1000
    img64 = dd.forward(h, T, rec, lights, mat)
    r32 = out["radiance"].astype(np.float64)
    assert np.abs(img64 - r32).max() <= 1e-5 * np.abs(r32).max()
    rng = np.random.default_rng(1)
    target = img64 * (1 + 0.05 * rng.standard_normal(img64.shape))
    dh, dt = dd.adjoint(h, T, rec, lights, mat, 2 * (img64 - target))
    nodes = rng.choice(h.size, 300, replace=False)
    fd = dd.fd_heights(h, T, rec, lights, mat, target, nodes, 1e-3)
    a = dh[nodes]
    assert np.count_nonzero(a) > 150
    assert np.abs(fd - a).max() <= 1e-5 * np.abs(a).max()
    k = np.argsort(-np.abs(dt))[:10]
    for ti in k:
        tp = T.astype(np.float64).reshape(-1).copy()
        tm = tp.copy()
        tp[ti] += 1e-4
        tm[ti] -= 1e-4
        Lp = ((dd.forward(h, tp, rec, lights, mat) - target) ** 2).sum()
        Lm = ((dd.forward(h, tm, rec, lights, mat) - target) ** 2).sum()
        assert abs((Lp - Lm) / 2e-4 - dt[ti]) <= 1e-7 * np.abs(dt).max()
