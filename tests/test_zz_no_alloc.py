"""solhip_render_noalloc (the megakernel into a device image, its scratch in a static device array)
against solhip_render with the megakernel into a device image: bit-identical images on two calls in
a row, the second with a smaller image, each adding into an image that already holds values, with
both node layouts; bad arguments refused."""
import numpy as np
import pytest
import torch

import sol_amd
from sol_amd import _ffi as F

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return sol_amd.Context(0)


@pytest.mark.parametrize("width", [4, 2])
def test_noalloc_render_equals_render(ctx, width):
    scene = sol_amd.Scene.sphere_field(num_spheres=40, subdivision=2, seed=5, extent=4.0, radius_max=1.0)
    g = ctx.upload(scene)
    cfg = sol_amd.PathTracerConfig(max_path_len=12)
    rng = np.random.default_rng(width)
    ctx.set_tuning(bvh_width=width)
    try:
        for w, h, si, sc in [(160, 90, 0, 8), (48, 30, 77, 3)]:     # the second call smaller
            print(f"bvh width {width}: {w}x{h}, samples {si}..{si + sc - 1}", flush=True)
            start = rng.random((3, h, w), dtype=np.float32)
            a = torch.from_numpy(start).cuda()
            b = a.clone()
            torch.cuda.synchronize()
            g.render_device(a.data_ptr(), w, h, si, sc, cfg, renderer=F.RENDERER_MEGAKERNEL)
            ctx.reset_stats()
            g.render_noalloc(b.data_ptr(), w, h, si, sc, cfg)
            st = ctx.stats()
            torch.cuda.synchronize()
            want, got = a.cpu().numpy(), b.cpu().numpy()
            assert (want != start).any()
            np.testing.assert_array_equal(got, want)
            assert st.samples == w * h * sc and st.extension_rays >= st.samples
    finally:
        ctx.set_tuning(bvh_width=0)


def test_noalloc_rejects_bad_arguments(ctx):
    g = ctx.upload(sol_amd.Scene.cornell_box())
    img = torch.zeros((3, 8, 8), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(F.SolError):
        g.render_noalloc(0, 8, 8, 0, 1)
    with pytest.raises(F.SolError):
        g.render_noalloc(img.data_ptr(), 0, 8, 0, 1)
    with pytest.raises(F.SolError):
        g.render_noalloc(img.data_ptr(), 8, 8, 0, 1,
                         sol_amd.PathTracerConfig(min_survival_prob=0.9, max_survival_prob=0.1))
    g.render_noalloc(img.data_ptr(), 8, 8, 0, 0)       # zero samples: nothing added
    torch.cuda.synchronize()
    assert not img.any()
