"""The C++ path end to end: sol_render (sol::RenderJob driving sol::GpuPathTracer over the C++
scene model, the flow of sol's driver, on a procedural scene or a TOML scene file) against the
same samples rendered through the C ABI from Python, and the PFM it writes."""
import os
import subprocess

import numpy as np
import pytest

import sol_amd
from tests.test_image_files import read_exr

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "sol_amd", "sol_render")


def read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF"
        w, h = (int(v) for v in f.readline().split())
        assert float(f.readline()) < 0          # little endian
        data = np.frombuffer(f.read(), dtype="<f4")
    rows = data.reshape(h, w, 3)[::-1]          # stored bottom to top
    return np.ascontiguousarray(rows.transpose(2, 0, 1))


GLOSSY = os.path.join(ROOT, "tests", "golden", "cornell", "cornell_box_glossy.toml")


@pytest.mark.parametrize("renderer,scene_arg", [("wavefront", "cornell"), ("megakernel", "cornell"),
                                                ("wavefront", GLOSSY)])
def test_render_job_driver_matches_c_abi_render(tmp_path, renderer, scene_arg):
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} is not built (make -C sol_amd)")
    w, h, spp, spf = 40, 30, 7, 3
    out = tmp_path / "out.pfm"
    r = subprocess.run([DRIVER, "--scene", scene_arg, "--width", str(w), "--height", str(h), "--spp", str(spp),
                        "--spf", str(spf), "--max-path-len", "8", "--renderer", renderer, "-o", str(out)],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "3 frames, 7 samples per pixel" in r.stdout
    got = read_pfm(out)
    # RenderJob renders frames of samples [0, 3), [3, 6), [6, 7) into one image, then scales it
    scene = sol_amd.Scene.cornell_box(aspect=w / h) if scene_arg == "cornell" else sol_amd.Scene.load(GLOSSY, w, h)
    pt = sol_amd.GpuPathTracer(scene, sol_amd.PathTracerConfig(max_path_len=8), kind=renderer)
    want = sol_amd.new_image(w, h)
    for first in range(0, spp, spf):
        pt.render(want, first, min(spf, spp - first))
    want *= np.float32(1.0) / np.float32(spp)   # Image::scale(1.0f / spp)
    np.testing.assert_array_equal(got, want)
    assert got.mean() > 0.01


def test_driver_takes_sol_options_and_writes_exr_by_default(tmp_path):
    """sol's driver command line (positional scene file, -w, --height, -spp, the Russian-roulette
    and ray-offset options): the image goes to render.exr in channels B, G, R."""
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} is not built (make -C sol_amd)")
    w, h, spp = 36, 24, 5
    r = subprocess.run([DRIVER, GLOSSY, "-w", str(w), "--height", str(h), "-spp", str(spp), "--max-path-len", "6",
                        "--min-rr-path-len", "2", "--min-survival-prob", "0.1", "--max-survival-prob", "0.9",
                        "--ray-offset", "2e-5", "-a", "path_tracer"],
                       capture_output=True, text=True, timeout=120, cwd=tmp_path)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "Image was saved to 'render.exr'" in r.stdout
    assert "Scene summary:" in r.stdout
    names, ch = read_exr(tmp_path / "render.exr")
    assert names == ["B", "G", "R"]
    got = np.stack([ch["R"], ch["G"], ch["B"]])
    cfg = sol_amd.PathTracerConfig(max_path_len=6, min_rr_path_len=2, min_survival_prob=0.1, max_survival_prob=0.9,
                                   ray_offset=2e-5)
    pt = sol_amd.GpuPathTracer(sol_amd.Scene.load(GLOSSY, w, h), cfg)
    want = sol_amd.new_image(w, h)
    pt.render(want, 0, spp)
    want *= np.float32(1.0) / np.float32(spp)
    np.testing.assert_array_equal(got, want)
    # unknown options and formats are refused as sol's driver refuses them
    bad = subprocess.run([DRIVER, GLOSSY, "-f", "bmp"], capture_output=True, text=True, timeout=60, cwd=tmp_path)
    assert bad.returncode == 1 and "Unknown image format 'bmp'" in bad.stderr
