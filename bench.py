"""Benchmark: Mrays/s (primary + shadow) on the 33.5M-triangle DEM at 2048^2.

Workload (one "step" = one full frame): 4097^2 value-noise fBm DEM (seed 1)
draped on a 1737.4 km sphere with FP64 camera-relative rebasing, camera at
100 km, 2048^2 pixels, 16 spp, 3 bounces with Russian roulette, 8 spectral
bins, newSun at 1 AU and 30 deg elevation, 4096^2 log-normal albedo texture.
Inputs (mesh, BVH, texture) are resident in HBM before the timed region.

Prints one JSON line (see the field notes at the bottom of this file).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# star catalogue epoch: seconds past J2000 (about 2020)
STAR_EPOCH_S = 6.3e8
COUNTS_FIXTURE = ROOT / "tests" / "golden" / "dem4097_2048px_traversal_counts.json"
# HBM bytes per launch of the fused renderer from rocprofv3 FETCH_SIZE / WRITE_SIZE
# runs, scaled by the same counters on a copy of known size
TRAFFIC_PROFILE = ROOT / "profiles" / "k_render_fused_dem4097_2048px_traffic.json"
METRIC = "Mrays/s, 33M-tri DEM @2048^2: batched vs plain GPU; adjoint/forward time ratio"


def algorithmic_bytes(v_int: float, v_tri: float) -> float:
    """Bytes per traced ray: ray in (32) + hit out (16) + 64 B per BVH2 internal
    node visited + 36 B per triangle tested, with the visit counts of the
    reference binned-SAH BVH2 traversal on the same rays."""
    return 48.0 + 64.0 * v_int + 36.0 * v_tri


# Per shaded surface point: three vertices' albedo (32 B), normal (12 B) and uv
# (8 B), plus the 4 texels of one bilinear fetch of a one-channel FP32 albedo
# texture when the scene has one.
BYTES_PER_HIT = 156.0
BYTES_PER_HIT_TEXEL = 4 * 4.0
# Per camera sample: the pixel's global Kahan accumulators (value and
# compensation, read and write)
BYTES_PER_SAMPLE_ACC = 128.0


def traffic_bytes(kernel: str, workload: str):
    """Measured HBM bytes per launch (read + write) of this kernel on this workload, or None."""
    if not TRAFFIC_PROFILE.exists():
        return None
    t = json.loads(TRAFFIC_PROFILE.read_text())
    if not t["kernel"].startswith(kernel) or t["workload"] != workload:
        return None
    return int(t["traffic_bytes"])


# Grid build bytes per triangle: index read and slot written (12 + 4), flag pass (4 + 4), scan
# (4 + 4), order pass (slot and rank read, id written: 12), leaf triangle emitted (id 4 + index 12
# + three 12 B positions + 48 B written), refit of the bottom level (48 B read) and the Node4s
# written (128 B per four 4-triangle leaves, x 4/3 for the levels above: 11); a square
# power-of-two grid has one slot per triangle
GRID_BUILD_BYTES_PER_TRI = 12 + 4 + 8 + 8 + 12 + 100 + 48 + 11


def bvh_build_row(stats: dict, triangles: int) -> dict:
    """The BVH build against HBM over the build's device time.  Binned SAH: ~36
    B per triangle (its box and id) read per level of the binary tree.  Grid
    build: GRID_BUILD_BYTES_PER_TRI per triangle, one pass each."""
    grid = stats.get("method") == 2
    levels = int(stats["max_depth"])
    nbytes = (GRID_BUILD_BYTES_PER_TRI if grid else 36.0 * levels) * triangles
    a = nbytes / (stats["build_ms"] * 1e-3) / 1e9
    return {"ms": round(stats["build_ms"], 2), "method": "grid" if grid else "binned_sah", "triangles": triangles,
            "levels": levels,
            "roofline": {"bound": "hbm", "achieved": round(a, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(a / HBM_PEAK_GBS, 4), "algorithmic_bytes": nbytes},
            "sah_cost": round(stats["sah_cost"], 3), "wide_sah_cost": round(stats["wide_sah_cost"], 3),
            "wide_nodes": stats["num_wide_nodes"]}


def cpu_baseline(built, ctx, stride: int, threads: int):
    """Times the oracle (FP32 CPU restatement of the reference algorithm) on a
    strided pixel sample of the same frame; returns (dict, traversal counts)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as O  # noqa: E402  (checker only, never the measured GPU path)
    from vira_amd import workloads as W

    sc = built.scene
    p = W.make_projection(sc)
    heights = built.heights.cpu().numpy()
    t0 = time.time()
    om = O.dem_mesh(heights, O.make_projection(p.kind, p.spacing, p.radius, tuple(p.center), tuple(p.offset)),
                    None if built.vertex_albedo is None else built.vertex_albedo.cpu().numpy())
    build_s = time.time() - t0
    cam = O.camera_from(built.camera.to_c())
    lights = [O.light_from(L.to_c()) for L in built.lights]
    tex = None if built.texture is None else built.texture.cpu().numpy()
    mat = O.material(1.0, tex, sc.smooth_shading)
    opt = O.options(sc.spp, sc.bounces, sc.russian_roulette, 1, sc.render_seed)
    ii, jj = np.meshgrid(np.arange(0, sc.image, stride), np.arange(0, sc.image, stride))
    pix = (ii + jj * sc.image).astype(np.int32).reshape(-1)
    t0 = time.time()
    out = om.render(cam, lights, mat, opt, precision=32, pixels=pix, threads=threads)
    dt = time.time() - t0
    c = out["counts"]
    ps = c["primary"] + c["shadow"]
    res = {
        "value": ps / dt / 1e6,
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{pix.size} pixels (every {stride}th row and column) of" if stride > 1 else "all pixels of")
                  + f" the same frame x {sc.spp} spp x {sc.bounces} bounces, {ps} primary+shadow rays in {dt:.2f} s "
                  f"(reference-algorithm FP32 CPU restatement, binned-SAH BVH2, std::thread); mesh + BVH2 build "
                  f"{build_s:.1f} s not timed",
        "bvh_build_ms": om.build_ms,
        "bvh_build_threads": O.nthreads(),
    }
    counts = {
        "closest_rays": c["primary"] + c["bounce"],
        "closest_internal": c["closest_internal"],
        "closest_tris": c["closest_tris"],
        "shadow_rays": c["shadow"],
        "any_internal": c["any_internal"],
        "any_tris": c["any_tris"],
        "sample_pixels": int(pix.size),
    }
    return res, counts


def cpu_baseline_rows(built, passes) -> dict:
    """The oracle's CPU restatements of the sensor tail, denoiser and star
    renderer timed on bounded samples of the same inputs (one host thread each)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as O  # noqa: E402  (checker only, never the measured GPU path)
    from vira_amd import sensor as SN
    from vira_amd import workloads as W

    out = {}
    # pixel solid angles (FP64 Girard) of the full 2048^2 camera
    cam = built.camera
    t0 = time.perf_counter()
    O.pixel_solid_angles(O.camera_from(cam.to_c()))
    dt = time.perf_counter() - t0
    out["received_power"] = {"value": round(cam.width * cam.height / dt / 1e6, 2), "unit": "Mpixel/s", "cores": 1,
                             "kind": "port", "sample": f"{cam.width}x{cam.height} pixel solid angles in {dt:.2f} s"}
    # 81x81 convolution of a 128x128 crop of the frame's radiance
    psf = SN.GaussianPSF(1.5)
    k81 = O.psf_kernel(O.psf_from(psf.to_c()), 81)
    crop = passes.total_radiance[:128, :128].cpu().numpy()
    t0 = time.perf_counter()
    O.convolve(crop, k81)
    dt = time.perf_counter() - t0
    out["psf81_convolution"] = {"value": round(128 * 128 / dt / 1e6, 4), "unit": "Mpixel/s", "cores": 1,
                                "kind": "port", "sample": f"128x128 crop x 8 bands, 81x81 kernel in {dt:.2f} s"}
    # EATWT denoiser (5 + 5 levels) on a 512x512 crop of the passes
    sl = (slice(0, 512), slice(0, 512))
    arr = [t[sl].cpu().numpy() for t in (passes.direct_radiance, passes.indirect_radiance, passes.albedo,
                                         passes.depth, passes.normal_global)]
    t0 = time.perf_counter()
    O.denoise(*arr)
    dt = time.perf_counter() - t0
    out["denoise_eatwt"] = {"value": round(512 * 512 / dt / 1e6, 3), "unit": "Mpixel/s", "cores": 1, "kind": "port",
                            "sample": f"512x512 crop of the frame's passes, 5 + 5 levels in {dt:.2f} s"}
    # the 100,000-star field (the whole field)
    scam, spsf = W.star_field_camera()
    ks = [O.psf_kernel(O.psf_from(spsf.to_c()), K) for K in (3, 9, 27, 81)]
    cat = W.make_star_catalogue()
    t0 = time.perf_counter()
    dirs, irr = O.star_light(cat, STAR_EPOCH_S)
    dt = time.perf_counter() - t0
    out["star_light"] = {"value": round(cat.shape[0] / dt / 1e6, 3), "unit": "Mstars/s", "cores": 1, "kind": "port",
                         "sample": f"the {cat.shape[0]}-star catalogue (numpy, vectorised over stars) in {dt:.2f} s"}
    pmin = SN.minimum_detectable_power(scam.aperture_area(), scam.exposure_time)
    R = np.asarray(scam.rotation, np.float64).astype(np.float32)
    t0 = time.perf_counter()
    O.render_stars(R, scam.width, scam.height, 1.0, scam.aperture_area(), ks, pmin, SN.intrinsic_matrix(scam),
                   dirs, irr)
    dt = time.perf_counter() - t0
    out["stars"] = {"value": round(dirs.shape[0] / dt / 1e6, 3), "unit": "Msources/s", "cores": 1, "kind": "port",
                    "sample": f"{dirs.shape[0]} sources into 4096^2 in {dt:.2f} s (kernels computed beforehand)"}
    return out


def visits_leg(ctx, built, passes) -> dict:
    """One untimed render of the same frame by the visit-counting kernel: BVH
    nodes, leaves and triangles fetched per ray on the GPU's own tree, pushes
    past the 8 LDS stack entries, and the deepest stack."""
    from dataclasses import replace
    st = ctx.render(built.camera, replace(built.options, count_visits=True), passes)
    n_closest = st["primary_rays"] + st["bounce_rays"]
    n_shadow = max(1, st["shadow_rays"])
    out = {"kernel_ms": round(st["fused_ms"], 2)}
    for kind, v, n in (("closest", st["closest_visits"], n_closest), ("shadow", st["shadow_visits"], n_shadow)):
        out[kind] = {"rays": n, "nodes_per_ray": round(v["nodes"] / n, 3), "leaves_per_ray": round(v["leaves"] / n, 3),
                     "triangles_per_ray": round(v["triangles"] / n, 3),
                     "spills_per_ray": round(v["stack_spills"] / n, 4), "max_stack": v["max_stack"]}
    # camera, bounce and shadow rays share one traversal call: the node loop's SIMD use over both kinds
    cv, sv = st["closest_visits"], st["shadow_visits"]
    steps = cv["wave_node_steps"] + sv["wave_node_steps"]
    out["node_loop_simd_use"] = round((cv["nodes"] + sv["nodes"]) / max(1, 64 * steps), 3)
    return out


def compressed_leg(ctx, built, passes, steps: int) -> dict:
    """The DEM frame with the 64 B quantised nodes (blas_build compress=True)
    beside the 128 B nodes it is made from: kernel time per frame of each,
    passes compared bit for bit."""
    import torch

    def frames(n):
        ms = []
        for _ in range(n + 1):
            ms.append(ctx.render(built.camera, built.options, passes)["fused_ms"])
        torch.cuda.synchronize()
        return float(np.mean(ms[1:]))

    plain_ms = frames(steps)
    ref = passes.total_radiance.clone()
    st = ctx.blas_build(built.mesh, compress=True)
    q_ms = frames(steps)
    same = bool(torch.equal(ref, passes.total_radiance))
    ctx.blas_build(built.mesh)
    out = {"node4_ms": round(plain_ms, 2), "node4q_ms": round(q_ms, 2), "identical_passes": same,
           "node4_bytes": int(st["num_wide_nodes"]) * 128, "node4q_bytes": int(st["compressed_bytes"])}
    print(f"# compressed nodes (DEM frame): Node4 {plain_ms:.2f} ms, Node4Q {q_ms:.2f} ms, identical passes {same}",
          file=sys.stderr, flush=True)
    return out


def adjoint_leg(ctx, steps: int, warmup: int) -> dict:
    """Differentiable DI: 1025^2 DEM (2,097,152 triangles), 1024^2 image, fixed-visibility
    direct illumination.  Times the forward render (keeping hit records) and the
    adjoint (dL/dimage of the L2 loss -> dL/dheights for 1,050,625 nodes and
    dL/dtexels for 1024^2 texels), each a synchronous call, and the adjoint that
    re-traces the records instead (record budget 0)."""
    import torch
    from vira_amd import scene as S
    from vira_amd import workloads as W

    sc = W.DEM1025_DI
    d = W.build_diff(ctx, sc)
    b = d.built
    gh, gt = torch.empty_like(b.heights), torch.empty_like(b.texture)

    def timed(fn, n):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n * 1e3

    res = {}
    for label, budget in (("kept", 1 << 30), ("retraced", 0)):
        opts = S.DIOptions(samples=1, seed=0, record_budget=budget)
        img = ctx.di_render(b.camera, opts)
        grad = 2.0 * (img - d.target)
        for _ in range(warmup):
            ctx.di_render(b.camera, opts)
            ctx.adjoint(b.camera, opts, grad, gh, gt)
        fwd = timed(lambda: ctx.di_render(b.camera, opts), steps)
        st = {}

        def adj():
            st.update(ctx.adjoint(b.camera, opts, grad, gh, gt))
        adj_ms = timed(adj, steps)
        assert st["used_records"] == (1 if budget else 0)
        res[label] = (fwd, adj_ms)
    update_ms = timed(lambda: ctx.dem_update(b.mesh, b.heights), max(1, steps))
    refit_ms = timed(lambda: ctx.dem_update(b.mesh, b.heights, refit=True), max(1, steps))
    fwd, adj_ms = res["kept"]
    print(f"# DI adjoint ({sc.name}): forward {fwd:.3f} ms, adjoint {adj_ms:.3f} ms (records kept), "
          f"adjoint {res['retraced'][1]:.3f} ms (re-traced); mesh update + BVH rebuild {update_ms:.1f} ms, "
          f"+ BVH refit {refit_ms:.2f} ms", file=sys.stderr, flush=True)
    return {"workload": sc.name, "forward_ms": round(fwd, 3), "adjoint_ms": round(adj_ms, 3),
            "ratio": round(adj_ms / fwd, 3),
            "adjoint_retraced_ms": round(res["retraced"][1], 3),
            "ratio_retraced": round(res["retraced"][1] / res["retraced"][0], 3),
            "mesh_update_bvh_rebuild_ms": round(update_ms, 2), "mesh_update_bvh_refit_ms": round(refit_ms, 2)}


def batch_leg(ctx, dem_built, steps: int, warmup: int) -> dict:
    """Mixed batch: 64 requests (image side uniform over {128, 256, 512, 1024,
    2048}, spp uniform over {1, 4, 16, 64}, 3 bounces + RR, one of four meshes:
    the 257^2, 4097^2 and 1025^2 DEMs and a 200k-triangle small body; random camera and
    sun; seed 3).  Times the per-request form (one render call each) and the
    batched form (one persistent graph launch, configuration from the
    autotuner), after checking they write identical passes."""
    import torch
    from vira_amd import workloads as W

    meshes = W.build_batch_meshes(ctx, dem_built)
    reqs = W.make_batch(ctx, meshes, n=64, seed=3)
    t0 = time.perf_counter()
    cfg, tune_ms = ctx.batch_autotune(reqs, sample_stride=16)
    tune_s = time.perf_counter() - t0
    plain_reqs = [r.__class__(r.mesh, r.camera, r.options, r.light,
                              r.passes.__class__(r.passes.width, r.passes.height, which=("radiance", "depth", "tri_id")))
                  for r in reqs]

    def timed(fn, n):
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / n

    for _ in range(warmup):
        W.render_plain(ctx, plain_reqs)
        ctx.batch_submit(reqs)
    identical = all(torch.equal(a.passes.total_radiance, b.passes.total_radiance) and
                    torch.equal(a.passes.triangle_id, b.passes.triangle_id) for a, b in zip(reqs, plain_reqs))
    plain_s = timed(lambda: W.render_plain(ctx, plain_reqs), steps)
    st = {}
    batch_s = timed(lambda: st.update(ctx.batch_submit(reqs)), steps)
    rays = st["primary_rays"] + st["shadow_rays"]
    ctx.blas_build(dem_built.mesh)  # the autotuner rebuilt the meshes at its width: back to the default tree
    print(f"# mixed batch: 64 requests, {st['tiles']} tiles, {rays} primary+shadow rays; per-request {plain_s * 1e3:.1f} ms, "
          f"batched {batch_s * 1e3:.1f} ms (kernel {st['ms']:.1f} ms, graph reused {st['graph_reused']}); "
          f"autotune {tune_s:.1f} s -> {cfg}; identical passes: {identical}", file=sys.stderr, flush=True)
    return {"workload": "batch64_dem_smallbody_seed3", "plain_ms": round(plain_s * 1e3, 2),
            "batched_ms": round(batch_s * 1e3, 2), "speedup": round(plain_s / batch_s, 3),
            "batched_mrays_s": round(rays / batch_s / 1e6, 1), "identical_passes": bool(identical),
            "config": {"bvh_width": cfg.bvh_width, "waves_per_block": cfg.waves_per_block,
                       "waves_per_cu": cfg.waves_per_cu, "tiles_per_grab": cfg.tiles_per_grab},
            "autotune_s": round(tune_s, 2)}


def instanced_leg(ctx, dem_built, steps: int, warmup: int) -> dict:
    """Instanced scene: two-level scene (the 4097^2 DEM tile + 64k sphere for the rest of
    the Moon, 200k small body, 50k spacecraft) with FP64 model matrices at
    ~1.5e8 km, camera on the spacecraft; 4096^2, 8 spp, 2 bounces, sun as a
    sphere light at the origin.  Rebasing into the camera frame runs on the GPU
    in FP64 at every render."""
    import torch
    from vira_amd import scene as S
    from vira_amd import workloads as W

    isc = W.build_instanced(ctx, dem_built)
    passes = S.RenderPasses(4096, 4096, which=("radiance", "depth", "tri_id", "instance_id"))
    for _ in range(warmup):
        ctx.render(isc.camera, isc.options, passes)
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(ctx.render(isc.camera, isc.options, passes))
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    rays = stats[0]["primary_rays"] + stats[0]["shadow_rays"]
    inst = passes.instance_id.reshape(-1)
    seen = sorted(int(v) for v in torch.unique(inst[inst >= 0]).cpu())
    print(f"# instanced scene: {dt * 1e3:.1f} ms/frame, {rays / dt / 1e6:.1f} Mrays/s (primary+shadow), instances in view "
          f"{seen}", file=sys.stderr, flush=True)
    # roofline of the per-pixel kernel: bytes the traversal fetches per ray on this layout (the visit
    # counts of a counting render of the same frame), shaded hits and the per-sample accumulators,
    # over the kernel's own HIP-event time (mean of the timed frames)
    from dataclasses import replace
    vst = ctx.render(isc.camera, replace(isc.options, count_visits=True), passes)
    st0 = stats[0]
    n_closest = st0["primary_rays"] + st0["bounce_rays"]
    per_ray = lambda v, n: 48.0 + (128.0 * v["nodes"] + 48.0 * v["triangles"] + 32.0 * v["tlas_nodes"]) / max(1, n)
    a_c, a_s = per_ray(vst["closest_visits"], n_closest), per_ray(vst["shadow_visits"], st0["shadow_rays"])
    nbytes = n_closest * a_c + st0["shadow_rays"] * a_s + st0["closest_hits"] * BYTES_PER_HIT + \
        st0["primary_rays"] * BYTES_PER_SAMPLE_ACC
    kms = sum(s["fused_ms"] for s in stats) / len(stats)
    achieved = nbytes / (kms / 1e3) / 1e9
    visits = {k: {"nodes_per_ray": round(vst[f"{k}_visits"]["nodes"] / max(1, n), 3),
                  "tlas_nodes_per_ray": round(vst[f"{k}_visits"]["tlas_nodes"] / max(1, n), 3),
                  "triangles_per_ray": round(vst[f"{k}_visits"]["triangles"] / max(1, n), 3),
                  "max_stack": vst[f"{k}_visits"]["max_stack"]}
              for k, n in (("closest", n_closest), ("shadow", st0["shadow_rays"]))}
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "kernel": "k_render_fused<4, true, false, false>",
            "kernel_ms": round(kms, 2), "bytes_per_closest_ray": round(a_c, 1), "bytes_per_shadow_ray": round(a_s, 1),
            "bytes_per_hit": BYTES_PER_HIT, "bytes_per_sample_acc": BYTES_PER_SAMPLE_ACC, "visits": visits}
    print(f"# instanced roofline: {json.dumps(roof)}", file=sys.stderr, flush=True)
    # the same frame with every BLAS's quantised node copy
    for m in isc.meshes:
        ctx.blas_build(m, compress=True)
    ctx.tlas_build(isc.meshes, isc.transforms)
    ref = passes.total_radiance.clone()
    qms = [ctx.render(isc.camera, isc.options, passes)["fused_ms"] for _ in range(steps + 1)][1:]
    torch.cuda.synchronize()
    q_same = bool(torch.equal(ref, passes.total_radiance))
    roof["node4q_kernel_ms"] = round(float(np.mean(qms)), 2)
    roof["node4q_identical_passes"] = q_same
    print(f"# instanced scene with Node4Q: {np.mean(qms):.2f} ms (Node4 {kms:.2f} ms), identical passes {q_same}",
          file=sys.stderr, flush=True)
    for m in isc.meshes:
        ctx.blas_build(m)
    ctx.tlas_build([dem_built.mesh])  # back to the 4097^2 DEM scene
    ctx.lights_set(dem_built.lights)
    stars = star_leg(ctx, steps)
    return {"workload": "instanced_moon_smallbody_spacecraft_4096px_8spp_2b", "ms_per_frame": round(dt * 1e3, 2),
            "mrays_s": round(rays / dt / 1e6, 1), "bounce_rays": stats[0]["bounce_rays"], "instances_in_view": seen,
            "roofline": roof, "stars": stars}


def star_leg(ctx, steps: int) -> dict:
    """The star field: 100,000 stars (V magnitudes exponential on 0-10,
    uniform on the sky) through the unresolved renderer into the 4096^2 camera
    with its Airy-disk PSF (kernels 3/9/27/81 chosen per star)."""
    import torch
    from vira_amd import workloads as W

    cam, psf = W.star_field_camera()
    sp = ctx.star_psf(psf, cam)
    cat = W.make_star_catalogue()
    records = torch.from_numpy(cat.view(np.uint8).reshape(-1, 64).copy()).cuda()
    d, e = ctx.star_light(records, STAR_EPOCH_S)
    out = torch.zeros((cam.height, cam.width, 8), dtype=torch.float32, device="cuda")
    ctx.render_stars(cam, sp, d, e, out=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        d, e = ctx.star_light(records, STAR_EPOCH_S)
    torch.cuda.synchronize()
    dt_light = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.render_stars(cam, sp, d, e, out=out)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    lit = int((out[:, :, 0] > 0).sum().item())
    print(f"# stars: 100,000-star catalogue lit in {dt_light * 1e3:.2f} ms, rendered in {dt * 1e3:.2f} ms "
          f"({lit} pixels lit)", file=sys.stderr, flush=True)
    return {"sources": int(cat.shape[0]), "star_light_ms": round(dt_light * 1e3, 3), "ms": round(dt * 1e3, 3),
            "pixels_lit": lit}


def fft_convolution_bytes(W: int, H: int, K: int) -> int:
    """HBM bytes of the three FFT passes of one convolution (the kernel's
    spectrum kept from an earlier call): the row pass reads the 8-band image
    and writes the H rows' half spectra (Lx/2 + 1 columns padded to strips of
    4, 8 bands, complex FP32); the column pass reads and writes those and
    reads the kernel's Ly-row spectrum; the inverse pass reads the half
    spectra and writes the image."""
    h = K // 2
    Lx = 1 << max(1, (max(W + h, K) - 1).bit_length())
    Ly = 1 << max(1, (max(H + h, K) - 1).bit_length())
    cw = 4 if Ly <= 4096 else 2
    cols = -(-(Lx // 2 + 1) // cw) * cw
    image = W * H * 32
    half = 8 * cols * H * 8
    return (image + half) + (half + 8 * cols * Ly * 8 + half) + (8 * (Lx // 2 + 1) * H * 8 + image)


def sensor_leg(ctx, built, passes, steps: int) -> dict:
    """The sensor tail after a 4097^2 DEM frame: received power (pixel solid
    angles by Girard's theorem in FP64), convolution with an 81x81 Gaussian PSF
    (the largest of PointSpreadFunction's kernels, used for resolved renders)
    and the photosite/noise simulation, on the 2048^2 x 8-band image."""
    import torch
    from vira_amd import sensor as SN

    cam = built.camera
    psf = ctx.psf_kernel(SN.GaussianPSF(1.5), 81)
    cfg = SN.SensorConfig(exposure_time=0.01, photon_noise=True, noise_model=True, dark_current=5.0,
                          readout_noise_std=2.0, bit_depth=12)

    def timed(fn):
        fn()  # warm-up: first-call allocations and the PSF spectrum are set-up, not per-frame work
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps * 1e3, out

    rp_ms, (power, _) = timed(lambda: ctx.received_power(cam, passes.total_radiance))
    cv_ms, blurred = timed(lambda: ctx.convolve(power, psf))
    ss_ms, _ = timed(lambda: ctx.simulate_sensor(cfg, blurred))
    # EATWT denoiser (5 + 5 levels) on copies of the frame's passes
    from vira_amd import scene as S
    work = S.RenderPasses(passes.width, passes.height)
    for name in ("total_radiance", "direct_radiance", "indirect_radiance", "albedo", "normal_global", "depth"):
        getattr(work, name).copy_(getattr(passes, name))
    dn_ms, _ = timed(lambda: ctx.denoise(work))
    print(f"# sensor tail (2048^2 x 8): received power {rp_ms:.2f} ms, 81x81 PSF convolution {cv_ms:.2f} ms, "
          f"photosite + noise {ss_ms:.2f} ms; EATWT denoiser {dn_ms:.2f} ms", file=sys.stderr, flush=True)
    npx = passes.width * passes.height

    def hbm(nbytes, ms):
        a = nbytes / (ms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(a, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(a / HBM_PEAK_GBS, 4)}

    conv_fma = npx * 8 * 81 * 81
    conv_rate = conv_fma / (cv_ms * 1e-3)
    conv = hbm(fft_convolution_bytes(passes.width, passes.height, 81), cv_ms)
    conv["direct_sum_equivalent_tfma_s"] = round(conv_rate / 1e12, 1)
    return {"received_power_ms": round(rp_ms, 3), "psf81_convolution_ms": round(cv_ms, 3),
            "sensor_ms": round(ss_ms, 3), "denoise_eatwt_ms": round(dn_ms, 3),
            # algorithmic bytes / FMA per call over the synchronous call's wall time
            "roofline": {
                "received_power": hbm(npx * (32 + 32 + 4), rp_ms),
                "psf81_convolution": conv,
                "sensor": hbm(npx * (32 + 4), ss_ms),
                "denoise_eatwt": hbm(npx * 144 * 10, dn_ms),
            }}


def quipu_leg(ctx, built) -> dict:
    """The 4097^2 DEM as a quipu pyramid (levels 4097^2, 2049^2,
    1025^2; 16-bit heights in LZ4 blocks) written to a temporary file, then read
    back: file parse + LZ4 decode on the host, dequantisation and mesh on the
    GPU, and the BVH build of the finest level."""
    import tempfile
    import torch
    from vira_amd import quipu as Q

    h = built.heights.cpu().numpy()
    sp = float(built.scene.spacing)
    levels = []
    for k in range(3):
        hk = np.ascontiguousarray(h[::1 << k, ::1 << k])
        p = Q.DEMProjection(resolution=(hk.shape[1], hk.shape[0]), pixel_scale=(sp, sp), dx=float(1 << k),
                            dy=float(1 << k))
        levels.append(Q.DEMLevel(hk, p))
    with tempfile.TemporaryDirectory() as tmp:
        path = f"{tmp}/dem4097.qld"
        t0 = time.perf_counter()
        Q.write_pyramid(path, levels, compress=True)
        write_s = time.perf_counter() - t0
        size = Path(path).stat().st_size
        d = Q.QuipuDEM(path)
        mid, _, _, info = d.read_level(ctx, 0)          # warm (file cache, allocations)
        ctx.mesh_release(mid)
        t0 = time.perf_counter()
        mid, _, _, info = d.read_level(ctx, 0)
        torch.cuda.synchronize()
        read_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        ctx.blas_build(mid)
        torch.cuda.synchronize()
        bvh_s = time.perf_counter() - t0
        ctx.mesh_release(mid)
        d.close()
    print(f"# quipu: 4097^2 level ({size / 2**20:.1f} MiB file, 16-bit LZ4) read + mesh {read_s * 1e3:.1f} ms "
          f"(parse {info['read_ms']:.1f} ms of which LZ4 {info['decode_ms']:.1f} ms, device {info['build_ms']:.1f} ms), "
          f"BVH {bvh_s * 1e3:.1f} ms; write {write_s:.1f} s", file=sys.stderr, flush=True)
    return {"file_mib": round(size / 2**20, 1), "read_mesh_ms": round(read_s * 1e3, 2),
            "parse_ms": round(info["read_ms"], 2), "lz4_ms": round(info["decode_ms"], 2),
            "device_ms": round(info["build_ms"], 2), "bvh_ms": round(bvh_s * 1e3, 2)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scene", default="dem4097", choices=["dem257", "dem4097"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-stride", type=int, default=1, help="time every n-th pixel row/column on the CPU")
    ap.add_argument("--counts-out", default=None, help="write the oracle traversal counts (fixture) to this path")
    ap.add_argument("--integrator", default="auto", choices=["auto", "fused", "wavefront"])
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable DI adjoint/forward leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the mixed-batch batched/per-request leg")
    ap.add_argument("--no-instanced", action="store_true", help="skip the instanced-scene leg")
    ap.add_argument("--no-quipu", action="store_true", help="skip the quipu pyramid read leg")
    ap.add_argument("--no-visits", action="store_true", help="skip the BVH visit-counting render")
    ap.add_argument("--no-compressed", action="store_true", help="skip the quantised-node comparison")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a known-size kernel for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        return 2

    import torch
    from vira_amd import scene as S
    from vira_amd import workloads as W

    sc = W.DEM4097 if args.scene == "dem4097" else W.DEM257
    if args.copy_calibration_mib > 0:
        n = args.copy_calibration_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"# calibration copy: {args.copy_calibration_mib} MiB read + {args.copy_calibration_mib} MiB written",
              file=sys.stderr, flush=True)
        del src, dst
    ctx = S.Context(0)
    t0 = time.time()
    built = W.build(ctx, sc)
    torch.cuda.synchronize()
    setup_s = time.time() - t0
    print(f"# scene {sc.name}: {sc.triangles} triangles, BVH build {built.bvh['build_ms']:.1f} ms (device), "
          f"setup {setup_s:.1f} s", file=sys.stderr, flush=True)
    passes = S.RenderPasses(sc.image, sc.image)
    built.options.integrator = args.integrator
    for _ in range(args.warmup):
        ctx.render(built.camera, built.options, passes)
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(ctx.render(built.camera, built.options, passes))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prim = sum(s["primary_rays"] for s in stats)
    shad = sum(s["shadow_rays"] for s in stats)
    bnc = sum(s["bounce_rays"] for s in stats)
    value = (prim + shad) / elapsed / 1e6
    fused = stats[0]["integrator"] == 1
    print(f"# {args.steps} frames in {elapsed:.3f} s ({'fused' if fused else 'wavefront'}): primary {prim}, "
          f"shadow {shad}, bounce {bnc}; fused kernel {sum(s['fused_ms'] for s in stats):.1f} ms, closest-hit "
          f"kernels {sum(s['trace_closest_ms'] for s in stats):.1f} ms, any-hit "
          f"{sum(s['trace_shadow_ms'] for s in stats):.1f} ms, shading {sum(s['shade_ms'] for s in stats):.1f} ms",
          file=sys.stderr, flush=True)

    # BVH build times of the same mesh, as frames are timed: after one untimed build of that kind
    # (first-call allocations and code loading), the mean of two builds; the default build the
    # frames used, and the binned SAH for comparison
    def timed_build(method: str) -> dict:
        ctx.blas_build(built.mesh, method=method)
        runs = [ctx.blas_build(built.mesh, method=method) for _ in range(2)]
        st = dict(runs[-1])
        st["build_ms"] = sum(r["build_ms"] for r in runs) / len(runs)
        return st

    bvh_default = dict(built.bvh)
    bvh_sah = None
    if built.bvh.get("method") == 2:
        bvh_sah = bvh_build_row(timed_build("sah"), sc.triangles)
        bvh_default = timed_build("auto")  # leaves the default tree in place for the legs below
        print(f"# BVH: grid build {bvh_default['build_ms']:.2f} ms, binned-SAH build {bvh_sah['ms']} ms",
              file=sys.stderr, flush=True)
    visits = visits_leg(ctx, built, passes) if (fused and not args.no_visits) else None
    compressed = compressed_leg(ctx, built, passes, max(2, min(5, args.steps))) if (fused and not args.no_compressed) \
        else None
    if visits is not None:
        print(f"# GPU visits (BVH{built.bvh.get('width', 4)}): {json.dumps(visits)}", file=sys.stderr, flush=True)
    sensor = sensor_leg(ctx, built, passes, max(1, min(3, args.steps))) if args.scene == "dem4097" else None
    quipu = quipu_leg(ctx, built) if (args.scene == "dem4097" and not args.no_quipu) else None
    adjoint = None if args.no_adjoint else adjoint_leg(ctx, max(3, args.steps), max(1, args.warmup))
    batch = None if (args.no_batch or args.scene != "dem4097") else batch_leg(ctx, built, max(1, min(3, args.steps)), 1)
    instanced = None if (args.no_instanced or args.scene != "dem4097") else instanced_leg(ctx, built, max(1, min(3, args.steps)), 1)

    cpu = None
    counts = None
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
        cpu, counts = cpu_baseline(built, ctx, args.cpu_stride, threads)
        cpu_build_ms, cpu_build_threads = cpu.pop("bvh_build_ms"), cpu.pop("bvh_build_threads")
        if args.scene == "dem4097":
            cpu["rows"] = cpu_baseline_rows(built, passes)
            cpu["rows"]["bvh_build"] = {
                "value": round(sc.triangles / (cpu_build_ms * 1e-3) / 1e6, 2), "unit": "Mtriangles/s",
                "cores": cpu_build_threads, "kind": "port",
                "sample": f"the oracle's binned-SAH BVH2 build (8 bins, leaves of <= 2) of the same "
                          f"{sc.triangles}-triangle mesh in {cpu_build_ms / 1e3:.2f} s, subtrees below the top "
                          f"levels on worker threads"}
        if args.counts_out:
            Path(args.counts_out).parent.mkdir(parents=True, exist_ok=True)
            Path(args.counts_out).write_text(json.dumps({"scene": sc.name, **counts}, indent=1) + "\n")
    if counts is None and COUNTS_FIXTURE.exists() and args.scene == "dem4097":
        counts = json.loads(COUNTS_FIXTURE.read_text())
    roofline = None
    if counts is not None and counts.get("closest_rays", 0) > 0:
        vi = counts["closest_internal"] / counts["closest_rays"]
        vt = counts["closest_tris"] / counts["closest_rays"]
        si = counts["any_internal"] / max(1, counts["shadow_rays"])
        stt = counts["any_tris"] / max(1, counts["shadow_rays"])
        a_closest = algorithmic_bytes(vi, vt)
        a_shadow = algorithmic_bytes(si, stt)
        closest_rays = prim + bnc
        hits, a_hit = None, None
        if fused:
            # one launch traces camera, bounce and shadow rays
            kernel, launches = "k_render_fused", sum(s["fused_launches"] for s in stats)
            ms = sum(s["fused_ms"] for s in stats)
            hits = sum(s["closest_hits"] for s in stats)
            a_hit = BYTES_PER_HIT + (BYTES_PER_HIT_TEXEL if sc.albedo == "texture" else 0.0)
            nbytes = closest_rays * a_closest + shad * a_shadow + hits * a_hit + prim * BYTES_PER_SAMPLE_ACC
        else:
            kernel, launches = "k_trace_queue (closest-hit launches)", sum(s["trace_closest_launches"] for s in stats)
            ms = sum(s["trace_closest_ms"] for s in stats)
            nbytes = closest_rays * a_closest
        achieved = nbytes / (ms / 1e3) / 1e9
        roofline = {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_bytes(kernel, sc.name), "kernel": kernel,
            "bytes_per_closest_ray": round(a_closest, 1), "bytes_per_shadow_ray": round(a_shadow, 1),
            "v_int_closest": round(vi, 2), "v_tri_closest": round(vt, 2),
            "v_int_shadow": round(si, 2), "v_tri_shadow": round(stt, 2),
            "bytes_per_hit": None if a_hit is None else round(a_hit, 1), "closest_hits": hits,
            "bytes_per_sample_acc": BYTES_PER_SAMPLE_ACC if fused else None,
            "launches": launches, "avg_launch_ms": round(ms / max(1, launches), 4),
        }
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded fBm DEM, log-normal albedo texture)",
        "config": {"workload": sc.name, "integrator": "fused" if fused else "wavefront",
                   "dem": sc.dem, "triangles": sc.triangles, "image": sc.image,
                   "spp": sc.spp, "bounces": sc.bounces, "russian_roulette": sc.russian_roulette,
                   "bins": 8, "bounce_rays_per_frame": bnc // args.steps,
                   "bvh_build_ms": round(bvh_default["build_ms"], 1)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched_vs_plain": None if batch is None else batch["speedup"],
        "adjoint_forward_ratio": None if adjoint is None else adjoint["ratio"],
        "batch": batch,
        "adjoint": adjoint,
        "instanced": instanced,
        "gpu_visits": visits,
        "bvh_compressed": compressed,
        "bvh_build": bvh_build_row(bvh_default, sc.triangles),
        "bvh_build_binned_sah": bvh_sah,
        "sensor_tail": sensor,
        "quipu": quipu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
