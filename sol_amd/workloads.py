"""Synthetic workloads of the benchmarks and tests, seeded.

batch_specs: the batched-rendering workload. Independent sphere-field scenes (the 1080p
generator's scene family) with a sphere count log-uniform in [4, 2000] at icosphere subdivision 2
(320 triangles per sphere: about 1.3k to 640k triangles per scene), square images with a side
drawn from {64, 128, 256, 512, 1024}, samples per pixel drawn from {4, 16, 64}, max path length 12.
"""
from dataclasses import dataclass

import numpy as np

from .scene import Scene

SIDES = (64, 128, 256, 512, 1024)
SPPS = (4, 16, 64)
BATCH_DEPTH = 12


@dataclass
class BatchSpec:
    num_spheres: int
    seed: int
    width: int
    height: int
    spp: int

    def scene(self, subdivision=2):
        return Scene.sphere_field(num_spheres=self.num_spheres, subdivision=subdivision, seed=self.seed, aspect=1.0)


def batch_specs(count=64, seed=2024, min_spheres=4, max_spheres=2000, sides=SIDES, spps=SPPS):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(round(np.exp(rng.uniform(np.log(min_spheres), np.log(max_spheres)))))
        side = int(rng.choice(sides))
        spp = int(rng.choice(spps))
        out.append(BatchSpec(num_spheres=n, seed=1000 + i, width=side, height=side, spp=spp))
    return out
