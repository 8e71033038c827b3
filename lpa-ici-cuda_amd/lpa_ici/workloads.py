"""The synthetic workloads the throughput and parity figures are quoted on.

Every image comes from lpa_ici.synth_image (polygons, disks, ramps and
oriented stripes) plus Gaussian noise from lpa_ici.make_noisy, both
deterministic in their seeds, so the benchmark and the tests build the same
pixels.  Image i of a set uses scene seed `scene_base + i` and noise seed
20250910 + i; per-image sigma and gamma are drawn with numpy's PCG64 from
the set's own seed.

  single(size)       one size x size image, scene seed 42, sigma 20/255, gamma 2
                     (512: the CPU-runnable case; 4096: the plain form at scale)
  uneven(256)        256 images, each side uniform in 128..2048, sigma from
                     {10,20,30,40}/255 and gamma from {1.5,2.0,2.5} per image
  batch_512(64)      64 images of 512 x 512, sigma 20/255, gamma 2 (the
                     differentiation workload)
  large_uneven(32)   32 images, each side uniform in 1024..8192, sigma and
                     gamma per image as uneven()

Generation runs the C library's generator on a pool of host threads (the
calls release the GIL).
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import make_noisy, synth_image

NOISE_SEED = 20250910
SIGMA, GAMMA = 20.0 / 255.0, 2.0


class ImageSet:
    """Shapes (width, height), per-image sigma and gamma, scene and noise
    seeds of a workload; images are built on demand."""

    def __init__(self, name, sides, sigmas, gammas, scene_base):
        self.name = name
        self.sides = [(int(w), int(h)) for w, h in sides]
        self.sigmas = [float(s) for s in sigmas]
        self.gammas = [float(g) for g in gammas]
        self.scene_seeds = [scene_base + i for i in range(len(self.sides))]
        self.noise_seeds = [NOISE_SEED + i for i in range(len(self.sides))]

    def __len__(self):
        return len(self.sides)

    @property
    def pixels(self):
        return sum(w * h for w, h in self.sides)

    def clean(self, i):
        w, h = self.sides[i]
        return synth_image(w, h, self.scene_seeds[i])

    def noisy(self, i):
        return make_noisy(self.clean(i), self.sigmas[i], self.noise_seeds[i])

    def noisy_all(self, threads=16):
        with ThreadPoolExecutor(max_workers=threads) as pool:
            return list(pool.map(self.noisy, range(len(self))))


def _uneven(name, n, lo, hi, seed, scene_base):
    rng = np.random.default_rng(seed)
    sides = rng.integers(lo, hi + 1, size=(n, 2))
    sig = [float(rng.choice([10, 20, 30, 40])) / 255.0 for _ in range(n)]
    gam = [float(rng.choice([1.5, 2.0, 2.5])) for _ in range(n)]
    return ImageSet(name, sides, sig, gam, scene_base)


def single(size=4096):
    s = ImageSet(f"single {size}x{size}", [(size, size)], [SIGMA], [GAMMA], 42)
    s.noise_seeds = [NOISE_SEED]
    return s


def uneven(n=256):
    return _uneven(f"{n} uneven images, sides 128..2048", n, 128, 2048, seed=3, scene_base=1000)


def batch_512(n=64):
    return ImageSet(f"{n} x 512^2", [(512, 512)] * n, [SIGMA] * n, [GAMMA] * n, 2000)


def large_uneven(n=32):
    return _uneven(f"{n} uneven images, sides 1024..8192", n, 1024, 8192, seed=5, scene_base=3000)
