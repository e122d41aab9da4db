"""Seeded initial conditions of the benchmark configurations (host side).

The reference sets initial states through its problem routines
(problems.F90:242-2047) and an xorshift RNG (random.F90:307-458).  The
configurations measured here instead use seeded states produced by this
module with a counter-based SplitMix64 generator (Steele, Lea & Flood 2014),
so that every consumer (C oracle, reference harness, HIP path, Fortran host)
starts from bit-identical conserved arrays ``U[nv][nz][ny][nx]`` (C order).
"""
from __future__ import annotations

import numpy as np

from .config import Params, EQSYS

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

DEFAULT_SEED = 20261202


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser applied to a uint64 array (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        z = (x + _GOLDEN).astype(np.uint64)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """n doubles in [0,1): the 53 high bits of SplitMix64(key + i)."""
    key = splitmix64(np.array([np.uint64(seed) ^ (np.uint64(stream) << np.uint64(40))],
                              dtype=np.uint64))[0]
    with np.errstate(over="ignore"):
        ctr = np.arange(n, dtype=np.uint64) + key
    return (splitmix64(ctr) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def cell_centres(p: Params):
    nz, ny, nx = p.shape
    dx = p.dx()
    x = p.xmin + (np.arange(nx) + 0.5) * dx[0]
    y = p.ymin + (np.arange(ny) + 0.5) * dx[1]
    z = (p.zmin + (np.arange(nz) + 0.5) * dx[2]) if p.ndims == 3 else np.zeros(1)
    return z, y, x


def prim_to_cons(p: Params, q: np.ndarray) -> np.ndarray:
    """prim2cons of equations.F90:1762 (HD) / 2815 (MHD) / 3478 (SRHD) / 4590 (SRMHD) /
    1360, 2220 (isothermal HD / MHD, 4 or 8 variables), array form."""
    g = p.gamma
    u = np.empty_like(q)
    if EQSYS[p.equation_system.lower()] == 3:
        # prim2cons_srhd_adi (equations.F90:3478-3532)
        gammaxi = (g - 1.0) / g
        vv = (q[1] * q[1] + q[2] * q[2]) + q[3] * q[3]
        vm = 1.0 - vv
        vs = np.sqrt(vm)
        ww = (q[0] + q[4] / gammaxi) / vm
        u[0] = q[0] / vs
        u[1:4] = ww * q[1:4]
        u[4] = (ww - q[4]) - u[0]
        return u
    if EQSYS[p.equation_system.lower()] == 4:
        gammaxi = (g - 1.0) / g
        vv = (q[1] * q[1] + q[2] * q[2]) + q[3] * q[3]
        bb = (q[4] * q[4] + q[5] * q[5]) + q[6] * q[6]
        vb = (q[1] * q[4] + q[2] * q[5]) + q[3] * q[6]
        vm = 1.0 - vv
        vs = np.sqrt(vm)
        wt = (q[0] + q[8] / gammaxi) / vm + bb
        u[0] = q[0] / vs
        u[1:4] = wt * q[1:4] - vb * q[4:7]
        u[4:8] = q[4:8]
        u[8] = ((wt - q[8]) - u[0]) - 0.5 * (vm * bb + vb * vb)
        return u
    u[0] = q[0]
    u[1:4] = q[0] * q[1:4]
    if q.shape[0] in (4, 8):
        # the isothermal systems (prim2cons_hd_iso / _mhd_iso, equations.F90:
        # 1360, 2220): no energy
        u[4:] = q[4:]
        return u
    ek = 0.5 * (u[1] * q[1] + u[2] * q[2] + u[3] * q[3])
    if EQSYS[p.equation_system.lower()] == 1:
        u[4] = q[4] / (g - 1.0) + ek
    else:
        u[4:8] = q[4:8]
        em = 0.5 * (q[4] * q[4] + q[5] * q[5] + q[6] * q[6])
        u[8] = (q[8] / (g - 1.0) + ek) + em
    return u


def blast_hd_2d(p: Params, seed: int = DEFAULT_SEED) -> np.ndarray:
    """rho=1, v=0, p=0.1 with p=10 where the cell centre has r<0.1, and a
    seeded uniform 1% density perturbation rho *= 1 + 0.01 (2 xi - 1)."""
    nz, ny, nx = p.shape
    z, y, x = cell_centres(p)
    r = np.sqrt(x[None, None, :] ** 2 + y[None, :, None] ** 2) * np.ones((nz, 1, 1))
    q = np.zeros((p.nv, nz, ny, nx))
    xi = uniform(seed, nz * ny * nx).reshape(nz, ny, nx)
    q[0] = 1.0 * (1.0 + 0.01 * (2.0 * xi - 1.0))
    q[4] = np.where(r < 0.1, 10.0, 0.1)
    return prim_to_cons(p, q)


def _random_modes(seed: int, stream: int, kmin: float, kmax: float, solenoidal: bool):
    """Random-phase modes with kmin <= |k| <= kmax over integer wave vectors in
    a half space (k and -k are one real mode): list of (k, complex 3-vector)."""
    ks = []
    r = int(np.ceil(kmax))
    for kz in range(-r, r + 1):
        for ky in range(-r, r + 1):
            for kx in range(-r, r + 1):
                k2 = kx * kx + ky * ky + kz * kz
                if k2 < kmin * kmin or k2 > kmax * kmax:
                    continue
                if (kz, ky, kx) <= (0, 0, 0):      # keep one of +-k
                    continue
                ks.append((kx, ky, kz))
    ks = np.array(ks, dtype=np.float64)
    n = len(ks)
    u = uniform(seed, 4 * 3 * n, stream).reshape(n, 3, 4)
    # Box-Muller Gaussian amplitudes for the real and imaginary parts
    rad = np.sqrt(-2.0 * np.log(1.0 - u[..., 0]))
    amp = rad * np.cos(2.0 * np.pi * u[..., 1]) + 1j * rad * np.sin(2.0 * np.pi * u[..., 1])
    ph = np.exp(2j * np.pi * u[..., 2])
    a = amp * ph
    if solenoidal:
        kk = (ks * ks).sum(1)
        kd = (a * ks).sum(1)
        a = a - ks * (kd / kk)[:, None]
    return ks, a


def _synth(p: Params, ks: np.ndarray, a: np.ndarray) -> np.ndarray:
    """Real field sum_k Re(a_k exp(2 pi i k.x)) at the cell centres, by FFT."""
    nz, ny, nx = p.shape
    L = np.array([p.xmax - p.xmin, p.ymax - p.ymin, p.zmax - p.zmin])
    spec = np.zeros((3, nz, ny, nx), dtype=np.complex128)
    x0 = np.array([p.xmin + 0.5 * (L[0] / nx), p.ymin + 0.5 * (L[1] / ny),
                   p.zmin + 0.5 * (L[2] / nz)])
    for (kx, ky, kz), av in zip(ks.astype(int), a):
        shift = np.exp(2j * np.pi * (kx * x0[0] / L[0] + ky * x0[1] / L[1] + kz * x0[2] / L[2]))
        c = 0.5 * av * shift
        spec[:, kz % nz, ky % ny, kx % nx] += c
        spec[:, (-kz) % nz, (-ky) % ny, (-kx) % nx] += np.conj(c)
    out = np.empty((3, nz, ny, nx))
    try:
        import scipy.fft as sfft
        for d in range(3):
            out[d] = sfft.ifftn(spec[d], workers=-1).real * (nx * ny * nz)
    except ImportError:
        for d in range(3):
            out[d] = np.fft.ifftn(spec[d]).real * (nx * ny * nz)
    return out


def turbulent_mhd_3d(p: Params, seed: int = DEFAULT_SEED, amplitude: float = 0.1) -> np.ndarray:
    """rho=1, p=0.6, psi=0; v and B from seeded random-phase Fourier modes
    with 1 <= |k| <= 4 (integer k in units of the box), each field scaled to
    an rms magnitude of ``amplitude``; the B modes are projected normal to k,
    so div B = 0 for the continuous field."""
    nz, ny, nx = p.shape
    q = np.zeros((p.nv, nz, ny, nx))
    q[0] = 1.0
    q[8] = 0.6
    ks, av = _random_modes(seed, 1, 1.0, 4.0, solenoidal=False)
    v = _synth(p, ks, av)
    ks, ab = _random_modes(seed, 2, 1.0, 4.0, solenoidal=True)
    b = _synth(p, ks, ab)
    v *= amplitude / np.sqrt((v * v).sum(0).mean())
    b *= amplitude / np.sqrt((b * b).sum(0).mean())
    q[1:4] = v[[0, 1, 2]]
    q[4:7] = b[[0, 1, 2]]
    return prim_to_cons(p, q)


def srmhd_blast_2d(p: Params, seed: int = DEFAULT_SEED, noise: float = 0.0,
                   vnoise: float = 0.0) -> np.ndarray:
    """Relativistic magnetised cylindrical blast: rho = 1, v = 0, p = 1 where
    the cell centre has r < 0.1 and 0.01 elsewhere, uniform Bx = 0.5, psi = 0;
    optionally a seeded uniform density perturbation rho *= 1 + noise (2 xi - 1)
    and seeded velocities vnoise (2 xi - 1) with transverse fields
    0.5 vnoise (2 xi - 1) (a state off the exact zeros of the gas at rest)."""
    nz, ny, nx = p.shape
    z, y, x = cell_centres(p)
    r = np.sqrt(x[None, None, :] ** 2 + y[None, :, None] ** 2) * np.ones((nz, 1, 1))
    q = np.zeros((p.nv, nz, ny, nx))
    q[0] = 1.0
    if noise:
        xi = uniform(seed, nz * ny * nx, stream=7).reshape(nz, ny, nx)
        q[0] = 1.0 + noise * (2.0 * xi - 1.0)
    q[4] = 0.5
    if vnoise:
        for k, v in enumerate((1, 2, 3, 5, 6)):
            xi = uniform(seed, nz * ny * nx, stream=30 + k).reshape(nz, ny, nx)
            q[v] = (0.5 if v > 4 else 1.0) * vnoise * (2.0 * xi - 1.0)
    q[8] = np.where(r < 0.1, 1.0, 0.01)
    return prim_to_cons(p, q)


def initial_state(p: Params, seed: int = DEFAULT_SEED) -> np.ndarray:
    """Seeded initial conserved state for the configuration ``p``."""
    eq = EQSYS[p.equation_system.lower()]
    if eq == 1:
        return blast_hd_2d(p, seed)
    if eq == 4:
        return srmhd_blast_2d(p, seed)
    return turbulent_mhd_3d(p, seed)


def orszag_tang_2d(p: Params, seed: int = DEFAULT_SEED, noise: float = 0.01) -> np.ndarray:
    """Orszag-Tang-type vortex on [0,1]^2: rho = 25/(36 pi), p = 5/(12 pi),
    v = (-sin 2 pi y, sin 2 pi x, 0), B = (-sin 2 pi y, sin 4 pi x, 0)/sqrt(4 pi),
    psi = 0, with a seeded uniform density perturbation rho *= 1 + noise (2 xi - 1)."""
    nz, ny, nx = p.shape
    z, y, x = cell_centres(p)
    Y, X = np.meshgrid(y, x, indexing="ij")
    q = np.zeros((p.nv, nz, ny, nx))
    xi = uniform(seed, nz * ny * nx, stream=5).reshape(nz, ny, nx)
    q[0] = 25.0 / (36.0 * np.pi) * (1.0 + noise * (2.0 * xi - 1.0))
    q[1] = -np.sin(2.0 * np.pi * Y)
    q[2] = np.sin(2.0 * np.pi * X)
    b0 = 1.0 / np.sqrt(4.0 * np.pi)
    q[4] = -b0 * np.sin(2.0 * np.pi * Y)
    q[5] = b0 * np.sin(4.0 * np.pi * X)
    q[8] = 5.0 / (12.0 * np.pi)
    return prim_to_cons(p, q)


def kelvin_helmholtz_2d(p: Params, seed: int = DEFAULT_SEED) -> np.ndarray:
    """MHD Kelvin-Helmholtz shear layer: vx = 0.5 tanh(y/0.02), seeded vy noise
    1e-3 U[-1,1], rho = 1 (y > 0) / 2 (y < 0), p = 2.5, uniform Bx = 0.1."""
    nz, ny, nx = p.shape
    z, y, x = cell_centres(p)
    Y = np.broadcast_to(y[None, :, None], (nz, ny, nx))
    q = np.zeros((p.nv, nz, ny, nx))
    xi = uniform(seed, nz * ny * nx, stream=3).reshape(nz, ny, nx)
    q[0] = np.where(Y > 0.0, 1.0, 2.0)
    q[1] = 0.5 * np.tanh(Y / 0.02)
    q[2] = 1.0e-3 * (2.0 * xi - 1.0)
    q[4] = 0.1
    q[8] = 2.5
    return prim_to_cons(p, q)
