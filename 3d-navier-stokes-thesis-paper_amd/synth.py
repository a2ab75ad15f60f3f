"""Seeded synthetic initial fields for the channel step (host-side input
preparation, numpy).  Fields are returned in the data1 layout
``[3, ny+1, nx/2+1, nz]`` (bytes of the reference's data1(nz, nx/2+1, ny+1, 3))
and are meant to be passed through ``chan_initialize_flow``, which removes
their divergence exactly as ``remove_initial_divergence`` does
(PFC/init_vel.f90:518-536) before the initial pressure is computed.

The kx = 0 plane is made Hermitian in kz so that the c2r transform (which
ignores the imaginary part of the DC x-bin) and the Fourier-space solve see
the same field; Nyquist modes are left empty.
"""
from __future__ import annotations

import math

import numpy as np


def _kz_index(nz):
    """Signed kz index of storage slot z (PFC/setup.f90:1451-1458)."""
    z = np.arange(nz)
    return np.where(z <= nz // 2, z, z - nz)


def make_hermitian(U):
    """Enforce U(kx=0, -kz) = conj U(kx=0, kz) and real kz = 0 / Nyquist entries."""
    nz = U.shape[-1]
    U = U.copy()
    for z in range(1, nz // 2):
        U[..., 0, nz - z] = np.conj(U[..., 0, z])
    U[..., 0, 0] = U[..., 0, 0].real
    U[..., 0, nz // 2] = U[..., 0, nz // 2].real
    return U


def _zero_walls(U):
    ny = U.shape[1] - 1
    U[:, 0] = 0.0
    U[0, ny] = 0.0
    U[2, ny] = 0.0
    U[1, ny - 1:] = 0.0
    return U


def laminar_perturbed(nx, ny, nz, yh, y, seed=0, amp=1.0e-3, kmax=4.0, lx=None, lz=None):
    """Laminar u = 1 - yh^2 plus a seeded random perturbation on the Fourier
    modes with sqrt(m^2 + n^2) <= kmax (integer mode numbers), uniform random
    phases, wall-vanishing y profiles and amplitude ``amp``."""
    rng = np.random.default_rng(seed)
    nxh = nx // 2 + 1
    U = np.zeros((3, ny + 1, nxh, nz), np.complex128)
    U[0, :, 0, 0] = 1.0 - np.asarray(yh) ** 2
    shape_u = 1.0 - np.asarray(yh) ** 2
    shape_v = (1.0 - np.asarray(y) ** 2) ** 2
    kzi = _kz_index(nz)
    for x in range(nxh - 1):
        for z in range(nz):
            m, n = x, int(kzi[z])
            if (m == 0 and n == 0) or abs(n) == nz // 2 or math.hypot(m, n) > kmax:
                continue
            ph = np.exp(2j * np.pi * rng.random((3, ny + 1)))
            mag = rng.random((3, ny + 1))
            U[0, :, x, z] = amp * shape_u * mag[0] * ph[0]
            U[2, :, x, z] = amp * shape_u * mag[2] * ph[2]
            U[1, :ny, x, z] = amp * shape_v * mag[1, :ny] * ph[1, :ny]
    U = _zero_walls(U)
    U[0, :, 0, 0] = 1.0 - np.asarray(yh) ** 2
    U[0, 0, 0, 0] = 0.0
    U[0, ny, 0, 0] = 0.0
    return make_hermitian(U)


def reichardt_profile(yh, re_tau=180.0):
    """Reichardt law of the wall, normalised to unit centreline velocity."""
    kappa, cst = 0.41, 7.8

    def uplus(yp):
        return (np.log1p(kappa * yp) / kappa
                + cst * (1.0 - np.exp(-yp / 11.0) - (yp / 11.0) * np.exp(-yp / 3.0)))
    yp = (1.0 - np.abs(np.asarray(yh))) * re_tau
    return uplus(yp) / uplus(re_tau)


def turbulent_like(nx, ny, nz, yh, y, seed=0, rms=0.1, re_tau=180.0):
    """Reichardt mean profile plus seeded random fluctuations with a k^-5/3
    energy spectrum over the resolved (kx, kz) modes, scaled so that the
    fluctuation rms over the box is ``rms`` (before projection)."""
    rng = np.random.default_rng(seed)
    nxh = nx // 2 + 1
    yh = np.asarray(yh)
    y = np.asarray(y)
    U = np.zeros((3, ny + 1, nxh, nz), np.complex128)
    kzi = _kz_index(nz)
    shape_u = 1.0 - yh ** 2
    shape_v = (1.0 - y ** 2) ** 2
    m = np.arange(nxh)[:, None].astype(float)
    n = kzi[None, :].astype(float)
    k = np.hypot(m, n)
    with np.errstate(divide="ignore"):
        amp = np.where(k > 0, k ** (-5.0 / 6.0) / np.sqrt(np.maximum(k, 1.0)), 0.0)
    amp[nxh - 1, :] = 0.0
    amp[:, nz // 2] = 0.0
    ph = np.exp(2j * np.pi * rng.random((3, ny + 1, nxh, nz)))
    mag = rng.random((3, ny + 1, nxh, nz))
    U[0] = amp[None] * shape_u[:, None, None] * mag[0] * ph[0]
    U[2] = amp[None] * shape_u[:, None, None] * mag[2] * ph[2]
    U[1, :ny] = amp[None] * shape_v[:, None, None] * mag[1, :ny] * ph[1, :ny]
    U = _zero_walls(make_hermitian(U))
    # Parseval: kx > 0 modes count twice in a real field
    wx = np.full(nxh, 2.0)
    wx[0] = 1.0
    e = (np.abs(U) ** 2 * wx[None, None, :, None]).sum() / (3.0 * (ny + 1))
    U *= rms / math.sqrt(max(e, 1e-300))
    U[:, :, 0, 0] = 0.0
    U[0, :, 0, 0] = reichardt_profile(yh, re_tau)
    U[0, 0, 0, 0] = 0.0
    U[0, ny, 0, 0] = 0.0
    return make_hermitian(U)


def ensemble_draw(n=48, seed=3):
    """Members of a seeded ensemble of channels of uneven size and Reynolds
    number: Nx, Nz in {32, 48, 64, 96}, Ny in {49, 65, 97}, Re_c uniform in
    [2000, 3500], 3/2 dealiasing, box 2 pi x 2 x pi.  Returns a list of dicts
    (nx, ny, nz, nx2, nz2, re, seed) with a per-member field seed."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nx = int(rng.choice([32, 48, 64, 96]))
        nz = int(rng.choice([32, 48, 64, 96]))
        ny = int(rng.choice([49, 65, 97]))
        re = float(rng.uniform(2000.0, 3500.0))
        out.append(dict(nx=nx, ny=ny, nz=nz, nx2=3 * nx // 2, nz2=3 * nz // 2, re=re, seed=seed * 1000 + i))
    return out
