"""Workload definitions shared by tests and the benchmark (sizes, seeds, POI
grids) -- the configurations of the speckle cases named in each helper."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from tw_amd import api


@dataclass
class Case:
    name: str
    img: int
    S: int
    grid: int
    N: int
    roi: tuple            # (start_x, start_y, w, h)
    margin: int
    ref: np.ndarray
    tar: np.ndarray
    deform: api.Deform | None

    def config(self, interp=api.INTERP_FUSED, N=None, max_iter=20, delta_p=0.001) -> api.Config:
        x, y, w, h = self.roi
        return api.make_config(self.img, self.img, x, y, w, h, self.S, self.S, self.grid, self.grid,
                               self.margin, self.margin, N or self.N, max_iter, delta_p, interp)

    def poi(self) -> np.ndarray:
        import math
        x, y, w, h = self.roi
        nx = int(math.floor((w - 2 * self.S - 2 * self.margin) / float(self.grid))) + 1
        ny = int(math.floor((h - 2 * self.S - 2 * self.margin) / float(self.grid))) + 1
        iy, ix = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
        py = y + self.margin + self.S + iy * self.grid
        px = x + self.margin + self.S + ix * self.grid
        return np.stack([py.ravel(), px.ravel()], 1).astype(np.int32)


def speckle_512(seed=1001) -> Case:
    """512x512, 4000 spots, target = Catmull-Rom resample under u=2.35, v=-1.72,
    noise sigma 2; POIs on a 10 px grid with a 20 px border, 31x31 subsets,
    FFT-CC window 32."""
    ref, tar = api.synth_pair(512, 512, 4000, seed, api.resample(2.35, -1.72), noise_sigma=2.0)
    # POI = start + margin + S = 1 + 4 + 15 = 20
    return Case("speckle512_S15_g10", 512, 15, 10, 32, (1, 1, 510, 510), 4, ref, tar, api.resample(2.35, -1.72))


def speckle_2048(seed=1002) -> Case:
    """2048x2048, 64000 spots, affine exx=0.5%, eyy=-0.2%, exy=0.1%, u=5.6,
    v=-3.3 about the centre, noise sigma 2; POIs on a 5 px grid from a 34 px
    border (max(N/2, S + ceil(max|u|) + 2)), 41x41 subsets, FFT-CC window 64."""
    d = api.affine(5.6, -3.3, 0.005, -0.002, 0.001, 1023.5, 1023.5)
    ref, tar = api.synth_pair(2048, 2048, 64000, seed, d, noise_sigma=2.0)
    # POI = 1 + 13 + 20 = 34; 397 x 397 = 157,609 POIs
    return Case("speckle2048_S20_g5", 2048, 20, 5, 64, (1, 1, 2046, 2046), 13, ref, tar, d)


def precision_4096(seed=1004) -> tuple[Case, np.ndarray]:
    """4096x4096, 0.0153 spots/px^2, sinusoidal field u = 0.5 sin(2 pi x/256),
    v = 0.3 cos(2 pi y/192), noise sigma 3, spots removed from 10 random
    200x200 boxes (low texture); 41x41 subsets on a 4 px grid from a 32 px
    border (1009 x 1009 = 1,018,081 POIs), FFT-CC window 64.  Returns the case
    and the boxes (int32[10][4] = x0, y0, x1, y1)."""
    rng = np.random.default_rng(seed)
    x0 = rng.integers(100, 4096 - 300, 10)
    y0 = rng.integers(100, 4096 - 300, 10)
    boxes = np.stack([x0, y0, x0 + 200, y0 + 200], 1).astype(np.int32)
    d = api.sinusoid(0.5, 256.0, 0.3, 192.0)
    ref, tar = api.synth_pair(4096, 4096, int(0.0153 * 4096 * 4096), seed, d, noise_sigma=3.0, boxes=boxes)
    # POI = 1 + 11 + 20 = 32; (4094 - 40 - 22) / 4 + 1 = 1009 per axis
    return Case("precision4096_S20_g4", 4096, 20, 4, 64, (1, 1, 4094, 4094), 11, ref, tar, d), boxes


def truth(case: Case, poi: np.ndarray) -> np.ndarray:
    """Analytic (u, ux, uy, v, vx, vy) at each POI of an affine case."""
    out = np.zeros((len(poi), 6))
    for i, (y, x) in enumerate(poi):
        out[i] = api.displacement(case.deform, float(x), float(y))
    return out


@dataclass
class BatchPair:
    """One image pair of the batched configuration: a size x size speckle pair
    (0.015 spots/px^2, noise sigma 2) under a random affine field, subset
    radius S, FFT-CC window N (32 for S <= 15, else 64), POIs on a `grid` px
    grid from a border of max(N/2, S + 2 + ceil(8 + size/100)) px."""
    size: int
    S: int
    grid: int
    N: int
    ref: np.ndarray
    tar: np.ndarray
    poi: np.ndarray
    deform: api.Deform

    def config(self, max_iter=20, delta_p=0.001) -> api.Config:
        s = self.size
        return api.make_config(s, s, 1, 1, s - 2, s - 2, self.S, self.S, 1, 1, 0, 0, self.N, max_iter, delta_p)

    @property
    def roi(self) -> tuple:
        return (1, 1, self.size - 2, self.size - 2)


def batch_layout(count=64, seed=1003) -> list[dict]:
    """The batched configuration without its images: per pair the size, S,
    grid step, N, affine field and POI array that batch_pairs draws (the same
    random stream), for tests that need only the work lists."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        size = int(rng.choice([512, 1024, 2048, 4096]))
        S = int(rng.choice([10, 15, 20, 30]))
        g = int(rng.choice([4, 5, 10]))
        N = 32 if S <= 15 else 64
        d = api.affine(*rng.uniform(-8, 8, 2), *rng.uniform(-0.01, 0.01, 3), size / 2, size / 2)
        m = max(N // 2, S + 2 + int(np.ceil(8 + 0.01 * size)))
        ys = np.arange(m, size - m, g)
        yy, xx = np.meshgrid(ys, ys, indexing="ij")
        poi = np.stack([yy.ravel(), xx.ravel()], 1).astype(np.int32)
        out.append(dict(size=size, S=S, grid=g, N=N, deform=d, poi=poi))
    return out


def batch_pairs(count=64, seed=1003, on_pair=None) -> list[BatchPair]:
    """The batched configuration: `count` speckle pairs with sizes drawn from
    {512, 1024, 2048, 4096}^2, subset radius from {10, 15, 20, 30}, grid step
    from {4, 5, 10}, strains uniform in [-1 %, 1 %] and translations uniform in
    [-8, 8] px about the image centre; pair i's images from seed 1003000 + i.
    on_pair(i, pair) is called after each pair is made (progress)."""
    out = []
    for i, L in enumerate(batch_layout(count, seed)):
        size = L["size"]
        ref, tar = api.synth_pair(size, size, int(0.015 * size * size), 1003000 + i, L["deform"], noise_sigma=2.0)
        out.append(BatchPair(size, L["S"], L["grid"], L["N"], ref, tar, L["poi"], L["deform"]))
        if on_pair:
            on_pair(i, out[-1])
    return out
