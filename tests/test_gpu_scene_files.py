"""sol's own test scenes (test/data/cornell_box*.toml: OBJ / MTL Cornell boxes with diffuse, mixed
diffuse-Phong, mirror and glass materials) loaded by the product and rendered on the GPU, against
the oracle rendering the tables of the oracle's own loader. Bit-exact: both sides run the same fp32
operations per sample, so the images agree to the last bit. The adjoint of the mixed BSDF is
compared with the oracle's adjoint (1e-3 relative, atomics sum in another order).
"""
import os

import numpy as np
import pytest

import sol_amd
from sol_amd import _ffi as F
from oracle import scene_files as SF
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu

FIXTURES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cornell")
NAMES = ["cornell_box", "cornell_box_glossy", "cornell_box_specular", "cornell_box_water"]
W, H = 64, 48


@pytest.fixture(scope="module")
def ctx():
    return sol_amd.Context(0)


@pytest.fixture(scope="module")
def loaded():
    out = {}
    for name in NAMES:
        path = os.path.join(FIXTURES, name + ".toml")
        out[name] = (sol_amd.Scene.load(path, width=W, height=H), SF.load_scene(path, aspect=W / H))
    return out


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("kind", [F.RENDERER_WAVEFRONT, F.RENDERER_MEGAKERNEL])
def test_sol_test_scene_renders_match_oracle(ctx, loaded, name, kind):
    scene, tables = loaded[name]
    cfg = sol_amd.PathTracerConfig(max_path_len=16)
    print(f"scene {name} renderer {kind} {W}x{H} 8 spp", flush=True)
    img = sol_amd.new_image(W, H)
    ctx.upload(scene).render(img, 3, 8, cfg, renderer=kind)
    want = np.zeros((3, H, W), np.float32)
    O.OracleScene.from_tables(tables).render(want, 3, 8, cfg)
    assert want.any() and np.isfinite(want).all()
    np.testing.assert_array_equal(img, want)


def test_glossy_scene_adjoint_matches_oracle(ctx, loaded):
    scene, tables = loaded["cornell_box_glossy"]
    mats = scene.materials()
    assert any(m.type == F.BSDF_INTERP for m in mats)
    cfg = sol_amd.PathTracerConfig(max_path_len=8)
    d_image = np.random.default_rng(4).standard_normal((3, H, W)).astype(np.float32)
    print("adjoint cornell_box_glossy", flush=True)
    dm, dl = ctx.upload(scene).backward(d_image, 0, 4, cfg)
    om, ol = O.backward(O.OracleScene.from_tables(tables), d_image, 0, 4, cfg, len(mats), len(scene.lights()))
    interp = [i for i, m in enumerate(mats) if m.type == F.BSDF_INTERP]
    assert np.abs(om[interp]).max() > 0          # the mixed materials receive gradients
    for a, b in ((dm, om), (dl, ol)):
        scale = max(np.abs(b).max(), 1e-12)
        np.testing.assert_allclose(a, b, rtol=1e-3, atol=1e-3 * scale)
