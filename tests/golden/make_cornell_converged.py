"""Writes cornell_64_4096spp.npy: the oracle's Cornell box at 64x64, 4096 samples per pixel
(samples 0..4095, max path length 8, Russian roulette from 3), divided by 4096: the converged image
that the equal-RMSE test measures 16-spp renders against.

Run from the repository root: python tests/golden/make_cornell_converged.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import sol_amd  # noqa: E402
from tests import oracle_lib  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cornell_64_4096spp.npy")


def main():
    scene = sol_amd.Scene.cornell_box()
    osc = oracle_lib.OracleScene(scene)
    img = sol_amd.new_image(64, 64)
    osc.render(img, 0, 4096, sol_amd.PathTracerConfig(max_path_len=8))
    np.save(OUT, (img / np.float32(4096)).astype(np.float32))
    print(OUT, img.mean() / 4096)


if __name__ == "__main__":
    main()
