"""The binned-SAH build's phases, traced against brute force (the oracle).

The build splits nodes of more than 256 triangles level by level (4096-slot
chunks, decisions on the GPU) and finishes smaller subtrees one wave each
(whole-wave splits above 32 triangles, one lane per node below).  These cases
drive each path: sizes around those boundaries, roots of many chunks,
coincident triangles (object-median splits in both phases), a chain of
triangles at exponentially growing distances (one-sided SAH splits, a deep
tree), and repeated builds (the same tree every time).

Bar: closest-hit triangle ids and distances bit-exact against the oracle's
brute-force test of every triangle (ties go to the lower id on both sides).
"""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu


def _rays(rng, n, lo, hi):
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays6 = np.concatenate([o, d], 1).astype(np.float32)
    rays8 = np.zeros((n, 8), np.float32)
    rays8[:, :3], rays8[:, 4:7], rays8[:, 7] = rays6[:, :3], rays6[:, 3:], np.inf
    return rays6, rays8


def _check(ctx, pos, idx, rays6, rays8, width, label):
    import torch
    mid = ctx.mesh_upload(pos, idx)
    st = ctx.blas_build(mid, width=width, method="sah")
    t, ids, _, _, _ = ctx.trace_rays(mid, torch.from_numpy(rays8).cuda())
    ot, oid, _, _, _ = O.Mesh(pos, idx).trace(rays6, brute=True)
    gid = ids.cpu().numpy().view(np.uint32)
    hits = int((oid != 0xFFFFFFFF).sum())
    print(f"{label} width {width}: {len(idx)} triangles, {hits} of {len(oid)} rays hit, "
          f"depth {st['max_depth']}, {st['num_leaves']} leaves", flush=True)
    assert np.array_equal(gid, oid), label
    assert np.array_equal(t.cpu().numpy().view(np.uint32), ot.view(np.uint32)), label
    ctx.mesh_release(mid)
    return st


def _soup(rng, ntri, spread=10.0, size=0.7):
    c = rng.uniform(-spread, spread, (ntri, 1, 3))
    pos = (c + rng.normal(0, size, (ntri, 3, 3))).reshape(-1, 3).astype(np.float32)
    return pos, np.arange(3 * ntri, dtype=np.uint32).reshape(-1, 3)


@pytest.mark.parametrize("width", [2, 4])
@pytest.mark.parametrize("ntri", [255, 256, 257, 4097, 100_000])
def test_binned_sah_sizes(gpu_ctx, ntri, width):
    """Roots below, at and above the subtree size, and roots of many chunks."""
    rng = np.random.default_rng(ntri)
    pos, idx = _soup(rng, ntri, spread=10.0 * (ntri / 4000) ** (1 / 3) + 5.0)
    rays6, rays8 = _rays(rng, 2000 if ntri > 10_000 else 4000, -15, 15)
    _check(gpu_ctx, pos, idx, rays6, rays8, width, f"soup {ntri}")


@pytest.mark.parametrize("copies", [40, 300, 5000])
def test_binned_sah_coincident_triangles(gpu_ctx, copies):
    """Many copies of one triangle (their centroids coincide, so the build splits them at the
    object median, in a small subtree or a large level) among ordinary ones; a ray through the
    copies hits the lowest id."""
    rng = np.random.default_rng(copies)
    tri = np.array([[-1.0, -1.0, 0.0], [1.0, -1.0, 0.0], [0.0, 1.0, 0.0]], np.float32)
    rep = np.tile(tri, (copies, 1))
    pos_s, _ = _soup(rng, 600, spread=8.0)
    pos = np.concatenate([pos_s[: 3 * 300], rep, pos_s[3 * 300:]]).astype(np.float32)
    idx = np.arange(len(pos), dtype=np.uint32).reshape(-1, 3)
    rays6, rays8 = _rays(rng, 3000, -10, 10)
    # rays aimed at the copies from both sides
    k = 500
    o = np.stack([rng.uniform(-0.4, 0.4, k), rng.uniform(-0.4, 0.4, k), np.where(rng.random(k) < 0.5, 5.0, -5.0)], 1)
    d = np.zeros((k, 3))
    d[:, 2] = -np.sign(o[:, 2])
    extra6 = np.concatenate([o, d], 1).astype(np.float32)
    extra8 = np.zeros((k, 8), np.float32)
    extra8[:, :3], extra8[:, 4:7], extra8[:, 7] = extra6[:, :3], extra6[:, 3:], np.inf
    rays6, rays8 = np.concatenate([rays6, extra6]), np.concatenate([rays8, extra8])
    for width in (2, 4):
        _check(gpu_ctx, pos, idx, rays6, rays8, width, f"{copies} copies")


def test_binned_sah_exponential_chain(gpu_ctx):
    """Triangles at distances 1.6^i along x beside a dense cluster: the SAH peels the far ones off
    a few at a time, so the large levels and the subtrees meet far down a one-sided tree."""
    rng = np.random.default_rng(7)
    pos_c, _ = _soup(rng, 3000, spread=2.0, size=0.2)
    chain = []
    for i in range(30):
        x = 3.0 + 1.6 ** i
        chain.append([[x, -0.5, -0.5], [x, 0.5, -0.5], [x, 0.0, 0.5]])
    pos = np.concatenate([pos_c, np.asarray(chain, np.float32).reshape(-1, 3)]).astype(np.float32)
    idx = np.arange(len(pos), dtype=np.uint32).reshape(-1, 3)
    rays6, rays8 = _rays(rng, 3000, -4, 4)
    # rays along +x through the chain
    k = 400
    o = np.stack([np.full(k, -5.0), rng.uniform(-0.3, 0.3, k), rng.uniform(-0.3, 0.3, k)], 1)
    d = np.tile([1.0, 0.0, 0.0], (k, 1))
    e6 = np.concatenate([o, d], 1).astype(np.float32)
    e8 = np.zeros((k, 8), np.float32)
    e8[:, :3], e8[:, 4:7], e8[:, 7] = e6[:, :3], e6[:, 3:], np.inf
    st = _check(gpu_ctx, pos, idx, np.concatenate([rays6, e6]), np.concatenate([rays8, e8]), 4, "chain")
    assert st["max_depth"] > 20


def test_binned_sah_repeatable(gpu_ctx):
    """Repeated builds of one mesh give the same tree (statistics and traversal results)."""
    import torch
    rng = np.random.default_rng(11)
    pos, idx = _soup(rng, 60_000, spread=30.0)
    mid = gpu_ctx.mesh_upload(pos, idx)
    rays6, rays8 = _rays(rng, 4000, -35, 35)
    ref = None
    for _ in range(3):
        st = gpu_ctx.blas_build(mid, width=4, method="sah")
        t, ids, _, _, _ = gpu_ctx.trace_rays(mid, torch.from_numpy(rays8).cuda())
        key = (st["sah_cost"], st["wide_sah_cost"], st["num_nodes"], st["num_wide_nodes"], st["num_leaves"],
               st["max_depth"], st["max_leaf_size"])
        res = (ids.cpu().numpy().tobytes(), t.cpu().numpy().tobytes())
        if ref is None:
            ref = (key, res)
        assert key == ref[0]
        assert res == ref[1]
    gpu_ctx.mesh_release(mid)
