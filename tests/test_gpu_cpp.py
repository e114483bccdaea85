"""The C++ host path: vira_amd/examples/render_dem.cpp drives the library
through include/vira_gpu.hpp (RAII context, exceptions, PathTracer with
options / renderPasses / render(camera)); the same scene rebuilt through the
Python binding must give the same passes bit for bit."""
import json
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from vira_amd import _capi as capi
from vira_amd import scene as S

pytestmark = pytest.mark.gpu
EXE = Path(__file__).resolve().parent.parent / "vira_amd" / "lib" / "render_dem"


class _RawCamera:
    def __init__(self, c):
        self.c = c

    def to_c(self):
        return self.c


def test_cpp_wrapper_render_matches_python(gpu_ctx):
    import torch
    assert EXE.exists(), "build the C++ example first (__graft_entry__.build())"
    r = subprocess.run([str(EXE), "65", "64", "2", "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    j = json.loads(r.stdout.strip().splitlines()[-1])
    assert j["bad_mesh_status"] == 1  # invalid mesh id -> VIRA_GPU_ERR_INVALID as an exception
    h = gpu_ctx.synth_fbm(65, 65, 0, 500.0, 6, 8.0)
    mid = gpu_ctx.dem_mesh(h, S.projection(0, 30.0))
    stats = gpu_ctx.blas_build(mid)
    assert stats["num_nodes"] == j["bvh_nodes"]
    gpu_ctx.tlas_build([mid])
    d, el = 1.495978707e11, math.radians(30.0)
    gpu_ctx.lights_set([S.PointLight((d * math.cos(el), 0.0, d * math.sin(el)),
                                     np.full(8, d * d, dtype=np.float32))])
    gpu_ctx.material_set(mid, albedo=0.15)
    c = capi.Camera()
    c.width = c.height = 64
    for k in range(6):
        c.kinv[k] = j["kinv"][k]
    for k, v in enumerate([1, 0, 0, 0, -1, 0, 0, 0, -1]):
        c.rot[k] = v
    c.position[0], c.position[1], c.position[2] = 0.0, 0.0, j["position_z"]
    c.z_dir = 1.0
    passes = S.RenderPasses(64, 64)
    st = gpu_ctx.render(_RawCamera(c), S.PathTracerOptions(samples=2, bounces=1, russian_roulette=True), passes)
    torch.cuda.synchronize()
    tri = passes.triangle_id.cpu().numpy().reshape(-1).view(np.uint32)
    hit = tri != 0xFFFFFFFF
    assert int(hit.sum()) == j["hits"] and int(tri[hit].astype(np.uint64).sum()) == j["tri_sum"]
    h64 = 1469598103934665603
    for b in passes.total_radiance.cpu().numpy().reshape(-1).view(np.uint32):
        h64 = ((h64 ^ int(b)) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    assert f"{h64:016x}" == j["radiance_fnv"]
    assert st["primary_rays"] == j["primary"] and st["shadow_rays"] == j["shadow"]


QEXE = Path(__file__).resolve().parent.parent / "vira_amd" / "lib" / "quipu_scene"


class _Cam:
    """A raw camera struct with the sensor the C++ example uses."""

    def __init__(self, c, area):
        self.c, self.width, self.height = c, c.width, c.height
        self.area = area

    def to_c(self):
        return self.c

    def sensor(self):
        s = capi.Sensor()
        for k in range(8):
            s.optical_efficiency[k] = 1.0
        s.aperture_area = self.area
        return s


class _Sim:
    def __init__(self, c):
        self.c = c

    def to_c(self):
        return self.c


def _fnv(a: np.ndarray) -> str:
    h = 1469598103934665603
    for b in np.ascontiguousarray(a).reshape(-1).view(np.uint32):
        h = ((h ^ int(b)) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return f"{h:016x}"


def test_cpp_quipu_scene_matches_python(gpu_ctx, tmp_path):
    """quipu_scene.cpp (QuipuDEM::readBuffers, PathTracer with options.denoise,
    receivedPower, UnresolvedRenderer, simulateSensor) against the same steps
    through the Python binding: identical radiance, power and sensor images."""
    import torch
    from vira_amd import quipu as Q
    assert QEXE.exists(), "build the C++ examples first (__graft_entry__.build())"
    h0 = gpu_ctx.synth_fbm(65, 65, 7, 400.0).cpu().numpy()
    rng = np.random.default_rng(7)
    levels = []
    for k in range(3):
        h = np.ascontiguousarray(h0[::1 << k, ::1 << k])
        p = Q.DEMProjection(resolution=(h.shape[1], h.shape[0]), pixel_scale=(25.0, 25.0), dx=float(1 << k),
                            dy=float(1 << k))
        levels.append(Q.DEMLevel(h, p, (0.05 + 0.1 * rng.random(h.shape)).astype(np.float32)))
    path = tmp_path / "scene.qld"
    Q.write_pyramid(str(path), levels, compress=True)
    r = subprocess.run([str(QEXE), str(path), "96", "60"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    j = json.loads(r.stdout.strip().splitlines()[-1])
    assert j["level"] == 1 and j["levels"] == 3 and j["width"] == 33  # GSDs 25, 50, 100: the coarsest <= 60
    d = Q.QuipuDEM(str(path))
    mid = d.read_buffers(gpu_ctx, 60.0)
    gpu_ctx.blas_build(mid)
    gpu_ctx.tlas_build([mid])
    dd, el = 1.495978707e11, math.radians(35.0)
    gpu_ctx.lights_set([S.PointLight((dd * math.cos(el), 0.0, dd * math.sin(el)),
                                     np.full(8, dd * dd, dtype=np.float32))])
    gpu_ctx.material_set(mid, albedo=1.0)
    c = capi.Camera()
    c.width = c.height = 96
    for k in range(6):
        c.kinv[k] = j["kinv"][k]
    for k, v in enumerate([1, 0, 0, 0, -1, 0, 0, 0, -1]):
        c.rot[k] = v
    c.position[2] = j["position_z"]
    c.z_dir = 1.0
    cam = _Cam(c, float(np.float32(2.5e-4)))
    passes = S.RenderPasses(96, 96, which=("radiance", "direct", "indirect", "albedo", "normal", "depth"))
    st = gpu_ctx.render(cam, S.PathTracerOptions(samples=4, bounces=1, denoise=True), passes)
    torch.cuda.synchronize()
    assert st["primary_rays"] == j["primary"]
    assert _fnv(passes.total_radiance.cpu().numpy()) == j["radiance_fnv"]
    power, _ = gpu_ctx.received_power(cam, passes.total_radiance)
    sp = capi.StarPSF()
    sp.count = 0
    f = j["f"]
    for k, v in enumerate([f, 0.0, 0.0, f, 48.0, 48.0]):
        sp.intrinsic[k] = float(np.float32(v))
    sp.cull_behind = 1
    dirs = torch.tensor([[0.01, 0.02, -1.0], [-0.05, 0.03, -1.0], [0.3, 0.2, 1.0]], dtype=torch.float32).cuda()
    irr = torch.full((3, 8), 1e-3, dtype=torch.float32).cuda()
    gpu_ctx.render_stars(cam, sp, dirs, irr, out=power)
    torch.cuda.synchronize()
    assert _fnv(power.cpu().numpy()) == j["power_fnv"]
    sim = capi.SensorSim()
    sim.exposure_time = 0.01
    for k in range(8):
        lam = (400.0 + 75.0 * (k + 0.5)) * 1e-9
        sim.photon_energy[k] = 6.62607015e-34 * 299792458.0 / lam
        sim.quantum_efficiency[k] = 0.5
    sim.gain, sim.max_adu = 0.5, 4095.0
    img, _ = gpu_ctx.simulate_sensor(_Sim(sim), power)
    torch.cuda.synchronize()
    assert _fnv(img.cpu().numpy()) == j["image_fnv"]
    assert j["image_sum"] > 0
    d.close()
