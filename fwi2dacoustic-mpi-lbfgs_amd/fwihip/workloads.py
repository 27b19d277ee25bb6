"""Acquisition geometry, wavelets and seeded synthetic models.

Conventions follow the reference's setup code:
  * physical ("b") grid nx x nz, row-major with x fastest (alloc.c:3-12);
  * the propagation ("f") grid adds npml cells left, right and below, none
    on top under a free surface (fwi_2d.c:148-165);
  * coordinates -> cells: ix = round(x/dx), iz = round(z/dz - 0.5), where
    round() is C's round-half-away-from-zero (fwi_2d.c:363-371).
"""
from dataclasses import dataclass, field
import math

import numpy as np


@dataclass
class Geometry:
    nx: int
    nz: int
    npml: int
    dx: float
    dz: float
    dt: float

    @property
    def pnx(self):
        return self.nx + 2 * self.npml

    @property
    def pnz(self):
        return self.nz + self.npml


@dataclass
class Shot:
    nt: int
    isx: int
    isz: int
    wav: np.ndarray
    irx: np.ndarray
    irz: np.ndarray
    f0: float = 0.0

    @property
    def nr(self):
        return int(len(self.irx))


def c_round(v):
    """C99 round(): halves away from zero."""
    return int(math.floor(v + 0.5)) if v >= 0 else -int(math.floor(-v + 0.5))


def cell_x(x, dx):
    return c_round(x / dx)


def cell_z(z, dz):
    return c_round(z / dz - 0.5)


def ricker(nt, dt, f0, t0=None):
    """Ricker wavelet (1 - 2 (pi f0 tau)^2) exp(-(pi f0 tau)^2), tau = t - t0.
    The reference declares ricker_wavelet(nt, dt, t0, fm, out) (header.h:360)
    without a body; this is the textbook definition, t0 = 1.2/f0 by default."""
    if t0 is None:
        t0 = 1.2 / f0
    t = np.arange(nt, dtype=np.float64) * dt - t0
    a = (math.pi * f0 * t) ** 2
    return ((1.0 - 2.0 * a) * np.exp(-a)).astype(np.float32)


def triangle_smooth(m, rect, repeat=1):
    """Separable triangle smoothing with folded edges: the smoothing the
    reference applies to start models and gradients (triangle_smooth.c:36-110,
    box filter of length rect applied twice per axis)."""
    out = np.asarray(m, np.float64).copy()

    def tri1d(a, nr, axis):
        if nr <= 1:
            return a
        a = np.moveaxis(a, axis, -1)
        n = a.shape[-1]
        box = np.ones(nr) / nr
        tri = np.convolve(box, box)  # length 2nr-1, centred at nr-1
        # fold the edges: reflect about the first and last sample
        pad = nr - 1
        ap = np.concatenate([a[..., 1:pad + 1][..., ::-1], a, a[..., -pad - 1:-1][..., ::-1]], axis=-1) \
            if pad < n else np.pad(a, [(0, 0)] * (a.ndim - 1) + [(pad, pad)], mode="reflect")
        res = np.apply_along_axis(lambda r: np.convolve(r, tri, mode="valid"), -1, ap)
        return np.moveaxis(res, -1, axis)

    for _ in range(repeat):
        out = tri1d(out, rect, 1)
        out = tri1d(out, rect, 0)
    return out.astype(np.float32)


# ---------------------------------------------------------------------------
# Workloads

def two_layer(geom, v_top=1500.0, v_bot=2500.0, z_if=500.0):
    z = (np.arange(geom.nz) + 0.5) * geom.dz
    vel = np.where(z[:, None] < z_if, v_top, v_bot) * np.ones((1, geom.nx))
    return vel.astype(np.float32)


def layered_faulted(geom, seed=2026, vmin=1500.0, vmax=4500.0):
    """Seeded layered model: 8-12 dipping layers with random thickness, two
    normal faults with 50-150 m throw, three Gaussian high-velocity lenses of
    radius 100-300 m, clipped to [vmin, vmax]."""
    rng = np.random.default_rng(seed)
    nx, nz, dx, dz = geom.nx, geom.nz, geom.dx, geom.dz
    X = np.arange(nx) * dx
    Z = (np.arange(nz) + 0.5) * dz
    xx, zz = np.meshgrid(X, Z)
    depth = nz * dz
    nl = int(rng.integers(8, 13))
    th = rng.uniform(0.5, 1.5, nl)
    tops = np.concatenate([[0.0], np.cumsum(th / th.sum() * depth)[:-1]])
    vels = np.sort(rng.uniform(vmin + 100.0, vmax - 400.0, nl))
    dip = math.tan(math.radians(rng.uniform(-8.0, 8.0)))
    # two normal faults: x position, dip direction, throw
    zs = zz - dip * (xx - X[-1] / 2)
    for _ in range(2):
        xf = rng.uniform(0.2, 0.8) * X[-1]
        throw = rng.uniform(50.0, 150.0)
        slope = math.tan(math.radians(rng.uniform(55.0, 75.0)))
        hanging = xx > xf + zz / slope
        zs = np.where(hanging, zs - throw, zs)
    vel = np.full((nz, nx), vels[0])
    for k in range(1, nl):
        vel = np.where(zs >= tops[k], vels[k], vel)
    for _ in range(3):
        cx = rng.uniform(0.15, 0.85) * X[-1]
        cz = rng.uniform(0.3, 0.8) * depth
        r = rng.uniform(100.0, 300.0)
        vel += 600.0 * np.exp(-((xx - cx) ** 2 + (zz - cz) ** 2) / (2 * r * r))
    return np.clip(vel, vmin, vmax).astype(np.float32)


def surface_shots(geom, nshots, nt, seed=2027, fmin=8.0, fmax=12.0, z_src=20.0, z_rec=20.0,
                  rec_step=1, aperture=None):
    """Shots spread evenly along the surface, receivers at every rec_step-th
    column (all columns, or an aperture centred on the source)."""
    rng = np.random.default_rng(seed)
    nx = geom.nx
    shots = []
    for j in range(nshots):
        isx = int(round((j + 0.5) * nx / nshots))
        isx = min(max(isx, 0), nx - 1)
        f0 = float(rng.uniform(fmin, fmax))
        if aperture is None:
            irx = np.arange(0, nx, rec_step, dtype=np.int32)
        else:
            lo = max(0, isx - aperture // 2)
            hi = min(nx, lo + aperture)
            lo = max(0, hi - aperture)
            irx = np.arange(lo, hi, rec_step, dtype=np.int32)
        irz = np.full(len(irx), cell_z(z_rec, geom.dz), np.int32)
        shots.append(Shot(nt=nt, isx=isx, isz=cell_z(z_src, geom.dz), wav=ricker(nt, geom.dt, f0),
                          irx=irx, irz=irz, f0=f0))
    return shots


def config1():
    """Plain reference method, single shot: 200x100, dx=dz=10 m, dt=1 ms,
    nt=1000, npml=20, free surface; two layers 1500/2500 m/s at z=500 m;
    10 Hz Ricker at (1000 m, 20 m); 100 receivers at 20 m depth."""
    g = Geometry(nx=200, nz=100, npml=20, dx=10.0, dz=10.0, dt=1e-3)
    vel = two_layer(g)
    nt = 1000
    irx = np.array([cell_x(10.0 + 20.0 * i, g.dx) for i in range(100)], np.int32)
    irz = np.full(100, cell_z(20.0, g.dz), np.int32)
    shot = Shot(nt=nt, isx=cell_x(1000.0, g.dx), isz=cell_z(20.0, g.dz), wav=ricker(nt, g.dt, 10.0),
                irx=irx, irz=irz, f0=10.0)
    start = triangle_smooth(vel, 15)
    return g, vel, start, [shot]


def config2(nshots=64, nt=4000, seed=2026):
    """Batched forward+adjoint: 1024x320, dx=dz=10 m, dt=0.8 ms, nt=4000,
    npml=20, 64 shots, 8-12 Hz Ricker, 1024 receivers at every column."""
    g = Geometry(nx=1024, nz=320, npml=20, dx=10.0, dz=10.0, dt=0.8e-3)
    vel = layered_faulted(g, seed)
    shots = surface_shots(g, nshots, nt, seed=seed + 1)
    return g, vel, shots


def config4(nshots=16, nt=10000, seed=2026):
    """Large grid: 4096x1024, dx=dz=5 m, dt=0.4 ms, nt=10000, npml=20, 16
    shots, 8-12 Hz Ricker, 4096 receivers at every column, on the seeded
    layered/faulted model (4,317,984 propagation cells per shot)."""
    g = Geometry(nx=4096, nz=1024, npml=20, dx=5.0, dz=5.0, dt=0.4e-3)
    vel = layered_faulted(g, seed)
    shots = surface_shots(g, nshots, nt, seed=seed + 1)
    return g, vel, shots


def config3(nshots=96, nt_range=(2000, 6000), seed=2026):
    """Uneven batch: 1536x400, dx=dz=10 m, dt=0.8 ms, npml=20, 96 shots with
    per-shot nt drawn uniformly from nt_range, receiver apertures of 256-1536
    contiguous columns centred on each source, Ricker peak frequency drawn
    uniformly from 5-15 Hz, on the seeded layered/faulted model."""
    g = Geometry(nx=1536, nz=400, npml=20, dx=10.0, dz=10.0, dt=0.8e-3)
    vel = layered_faulted(g, seed)
    rng = np.random.default_rng(seed + 2)
    shots = []
    for j in range(nshots):
        nt = int(rng.integers(nt_range[0], nt_range[1] + 1))
        aperture = int(rng.integers(256, 1537))
        s = surface_shots(g, nshots, nt, seed=seed + 1 + j, fmin=5.0, fmax=15.0, aperture=aperture)[j]
        shots.append(s)
    return g, vel, shots
