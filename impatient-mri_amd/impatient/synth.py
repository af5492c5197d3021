"""Seeded synthetic reconstruction problems (no datasets ship for the large
configurations: the reference's 256x256x32 set is absent and its 384^2 / 512^2
sets lack k-space data).

Geometry only: trajectory, readout times, voxel coordinates, phantom, field
map and coil sensitivities.  k-space data is the forward NUDFT of the phantom
(computed by the caller: the GPU operator, or an FP64 reference for small
tests) plus complex Gaussian noise at a stated SNR.

Layout: image index = (y*Nx + x)*Nz + z (x fastest in 2D, z fastest in 3D),
matching the Q grid that main_mri.cu:461-500 builds, so the direct Toeplitz
path is self-consistent.  Coordinates: x = (i - N/2)/N in FOV units; k in
cycles/FOV; field map in rad/s; t in seconds.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Problem:
    name: str
    Nx: int
    Ny: int
    Nz: int
    ncoils: int
    ntime_segs: int
    kx: np.ndarray
    ky: np.ndarray
    kz: np.ndarray | None
    t: np.ndarray
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray | None
    fm: np.ndarray
    rho: np.ndarray            # phantom, complex64 [Nvox]
    sens: np.ndarray | None    # complex64 [C, Nvox] or None (single coil, S = 1)
    lam: float = 1e-3
    cg_iters: int = 10
    meta: dict = field(default_factory=dict)

    @property
    def nvox(self) -> int:
        return self.Nx * self.Ny * self.Nz

    @property
    def nk(self) -> int:
        return len(self.kx)

    @property
    def shape(self):
        """Image shape as (N1, N2, N3) = (Ny, Nx, Nz)."""
        return (self.Ny, self.Nx, self.Nz)


def spiral(N: int, J: int, S: int, t_read: float, t_start: float = 0.0):
    """Variable-density spiral: J interleaves x S samples, k in [-N/2, N/2)."""
    tau = np.arange(S) / S
    kmax = N / 2 * (1 - 1e-6)
    r = kmax * tau ** 1.5
    n_turns = N / (2 * J)
    kx, ky = [], []
    for j in range(J):
        th = 2 * np.pi * n_turns * tau + 2 * np.pi * j / J
        kx.append(r * np.cos(th))
        ky.append(r * np.sin(th))
    t = np.tile(t_start + tau * t_read, J)
    return np.concatenate(kx), np.concatenate(ky), t


_SHEPP = [  # (A, a, b, x0, y0, phi_deg): modified Shepp-Logan
    (1.0, .69, .92, 0, 0, 0), (-.8, .6624, .8740, 0, -.0184, 0),
    (-.2, .1100, .3100, .22, 0, -18), (-.2, .1600, .4100, -.22, 0, 18),
    (.1, .2100, .2500, 0, .35, 0), (.1, .0460, .0460, 0, .1, 0),
    (.1, .0460, .0460, 0, -.1, 0), (.1, .0460, .0230, -.08, -.605, 0),
    (.1, .0230, .0230, 0, -.606, 0), (.1, .0230, .0460, .06, -.605, 0)]


def shepp_logan(X, Y, Z=None):
    """Phantom at coordinates in FOV units (ellipses on [-1, 1] scaled to FOV/2)."""
    u, v = 2 * X, 2 * Y
    img = np.zeros(X.shape)
    for A, a, b, x0, y0, ph in _SHEPP:
        c, s = np.cos(np.radians(ph)), np.sin(np.radians(ph))
        xr = (u - x0) * c + (v - y0) * s
        yr = -(u - x0) * s + (v - y0) * c
        q = (xr / a) ** 2 + (yr / b) ** 2
        if Z is not None:
            q = q + (2 * Z / 0.9) ** 2
        img += A * (q <= 1.0)
    return img


def field_map(rng, X, Y, nblobs: int, max_hz: float, z_ramp_hz: float = 0.0, Z=None):
    f = np.zeros(X.shape)
    for _ in range(nblobs):
        cx, cy = rng.uniform(-0.35, 0.35, 2)
        sg = rng.uniform(0.08, 0.25)
        amp = rng.choice([-1.0, 1.0]) * rng.uniform(0.5, 1.0)
        f += amp * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * sg ** 2))
    f *= max_hz / max(1e-12, np.abs(f).max())
    if Z is not None and z_ramp_hz:
        f += z_ramp_hz * (2 * Z)
    return 2 * np.pi * f        # rad/s


def sensitivities(rng, X, Y, ncoils: int, Z=None):
    S = []
    for c in range(ncoils):
        ang = 2 * np.pi * c / ncoils
        cx, cy = 0.45 * np.cos(ang), 0.45 * np.sin(ang)
        d2 = (X - cx) ** 2 + (Y - cy) ** 2
        if Z is not None:
            d2 = d2 + Z ** 2
        S.append(np.exp(-d2 / (2 * 0.4 ** 2)) * np.exp(1j * rng.uniform(0, 2 * np.pi)))
    S = np.array(S)
    S /= np.sqrt(np.sum(np.abs(S) ** 2, axis=0, keepdims=True))
    return S.astype(np.complex64)


def make_problem(name, N, Nz, ncoils, J, S, t_read, max_hz, nblobs, ntime_segs, seed,
                 z_ramp_hz=0.0, lam=1e-3, cg_iters=10, Nx=None) -> Problem:
    rng = np.random.default_rng(seed)
    Nx = N if Nx is None else Nx
    Ny = N
    kx2, ky2, t2 = spiral(N, J, S, t_read)
    if Nz > 1:
        kzs = np.arange(-Nz // 2, Nz // 2)
        kx = np.tile(kx2, Nz)
        ky = np.tile(ky2, Nz)
        kz = np.repeat(kzs.astype(np.float64), len(kx2))
        t = np.tile(t2, Nz)
    else:
        kx, ky, kz, t = kx2, ky2, None, t2
    iy, ix, iz = np.meshgrid(np.arange(Ny), np.arange(Nx), np.arange(Nz), indexing="ij")
    X = (ix - Nx // 2) / Nx
    Y = (iy - Ny // 2) / Ny
    Z = (iz - Nz // 2) / Nz if Nz > 1 else None
    rho = shepp_logan(X, Y, Z).astype(np.complex64)
    fm = field_map(rng, X, Y, nblobs, max_hz, z_ramp_hz, Z)
    sens = sensitivities(rng, X, Y, ncoils, Z).reshape(ncoils, -1) if ncoils > 1 else None
    f32 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float32).reshape(-1)
    return Problem(name=name, Nx=Nx, Ny=Ny, Nz=Nz, ncoils=ncoils, ntime_segs=ntime_segs,
                   kx=f32(kx), ky=f32(ky), kz=f32(kz), t=f32(t), x=f32(X), y=f32(Y),
                   z=f32(Z), fm=f32(fm), rho=rho.reshape(-1), sens=sens, lam=lam,
                   cg_iters=cg_iters,
                   meta=dict(seed=seed, J=J, S=S, t_read=t_read, max_hz=max_hz))


# Workloads (seed = 20260100 + index).  Configurations 0..3 of the build
# plan; `tiny` variants keep the same structure at oracle-friendly sizes.
def config(idx: int) -> Problem:
    seed = 20260100 + idx
    if idx == 0:
        return make_problem("cfg0_64x64x1_1coil", 64, 1, 1, 8, 1024, 5e-3, 100.0, 3, 12, seed,
                            lam=1e-3, cg_iters=10)
    if idx == 1:
        return make_problem("cfg1_64x64x16_4coil", 64, 16, 4, 8, 1024, 5e-3, 100.0, 3, 6, seed,
                            z_ramp_hz=50.0, lam=1e-3, cg_iters=20)
    if idx == 2:
        return make_problem("cfg2_256x256x32_4coil", 256, 32, 4, 32, 4096, 8e-3, 150.0, 5, 8,
                            seed, lam=1e-3, cg_iters=200)
    if idx == 3:
        return make_problem("cfg3_512x512x1_32coil", 512, 1, 32, 48, 8192, 10e-3, 200.0, 4, 10,
                            seed, lam=1e-3, cg_iters=200)
    raise ValueError(f"unknown configuration {idx}")


def tiny(N=16, Nz=1, ncoils=2, J=4, S=128, ntime_segs=3, max_hz=60.0, seed=7,
         t_read=4e-3) -> Problem:
    return make_problem(f"tiny_{N}x{N}x{Nz}_{ncoils}c", N, Nz, ncoils, J, S, t_read, max_hz, 2,
                        ntime_segs, seed)


def add_noise(d: np.ndarray, snr_db: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed + 7919)
    rms = np.sqrt(np.mean(np.abs(d) ** 2))
    sigma = rms / (10 ** (snr_db / 20))
    n = (rng.standard_normal(d.shape) + 1j * rng.standard_normal(d.shape)) * sigma / np.sqrt(2)
    return (d + n).astype(np.complex64)
