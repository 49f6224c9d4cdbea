"""Luminosity kernel, RGB loader and the HIP-solver test runner on the GPU."""
import io
import os

import numpy as np
import pytest
from PIL import Image as PILImage

import ghcm_amd as g
import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("w,h", [(1, 1), (7, 3), (64, 64), (1001, 517)])
def test_luminosity_kernel_vs_fp64(w, h):
    print(f"case luminosity {w}x{h}", flush=True)
    rng = np.random.default_rng(w * h)
    rgb = rng.random((h, w, 3), dtype=np.float32)
    with g.Context(0) as ctx:
        d_in = ctx.to_device(rgb)
        d_out = ctx.alloc((h, w), np.float32)
        g._check(g.lib().ghcm_rgb_luminosity_device(ctx._h, d_in.ptr, w, h, d_out.ptr))
        got = d_out.download()
        # an unaligned input (one float in) takes the per-pixel path
        if w * h > 1:
            flat = np.concatenate([np.zeros(1, np.float32), rgb.ravel()])
            d_in2 = ctx.to_device(flat)
            g._check(g.lib().ghcm_rgb_luminosity_device(ctx._h, d_in2.ptr + 4, w, h, d_out.ptr))
            assert np.array_equal(d_out.download(), got)
    ref = rgb.astype(np.float64) @ np.array([0.299, 0.587, 0.114])
    assert np.max(np.abs(got - ref)) <= 4 * np.finfo(np.float32).eps


def test_rgb_luminosity_loader(tmp_path):
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (40, 56, 3)).astype(np.uint8)
    p = tmp_path / "c.png"
    PILImage.fromarray(img, "RGB").save(p)
    with g.Context(0) as ctx:
        lum = g.field_load_rgb_luminosity(ctx, p)
    rgbf = g.image_load_f32(p, 3).astype(np.float64)
    ref = rgbf @ np.array([0.299, 0.587, 0.114])
    assert lum.shape == (40, 56)
    assert np.max(np.abs(lum - ref)) <= 4 * np.finfo(np.float32).eps


def test_reference_ring_image_solves(tmp_path):
    """field_cost_from_img on the reference's 256^2 ring image, solved on the GPU and by the
    oracle FMM (the reference's 7-ring job, src/test_runner.cpp:304-313, at 256^2)."""
    path = os.path.join(os.path.dirname(__file__), "golden", "7_ring_256_2d.png")
    speed = g.field_load_mono(path)
    cost = (1.0 / speed).astype(np.float32)
    seeds = [[0, 0]]
    ref = oracle.fmm_f32(cost, seeds)
    with g.Context(0, tile_w=64, tile_h=64, kernel=g.KERNEL_SWEEP) as ctx:
        t, st = ctx.solve(cost, seeds)
        g.test_run("hcm", cost, seeds, tmp_path, solver_name="ghcm_sweep", job_name="7 rings",
                   ctx=ctx, trials=2)
    fin = np.isfinite(ref)
    assert np.array_equal(np.isfinite(t), fin)
    assert np.max(np.abs(t[fin].astype(np.float64) - ref[fin])) <= 1e-4 * ref[fin].max()
    blob = g.blob_read(tmp_path / "ghcm_sweep_7 rings.blob")
    assert np.max(np.abs(blob[fin].astype(np.float64) - ref[fin])) <= 1e-4 * ref[fin].max()
    csv = (tmp_path / "ghcm_sweep_7 rings.csv").read_text().splitlines()
    assert csv[0].startswith("ghcm_sweep,7 rings, device,") and len(csv[0].split(",")) == 5
