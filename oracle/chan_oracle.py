"""CPU oracle for the pseudo-spectral channel-flow RK3 time step.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or the timed CPU baseline).  The product path (the HIP library under
``3d-navier-stokes-thesis-paper_amd/``) never calls it.

This is a numpy restatement of the reference's MPI-Fortran step on one rank
(``nproc == 1`` takes the ``_parallel`` code paths, whose data movement reduces
to the index arithmetic written out below).  Every function cites the
reference file:line it follows (paths relative to the reference root,
``PFC/`` = ``ParallelFortran-Code/``).

Array conventions (all C-ordered numpy arrays):

* Fourier velocity / NL term ``U[c, j, x, z]`` complex128, ``c`` in 0..2
  (u, v, w), ``j`` in 0..ny (plane ``j+1`` of the reference), ``x`` in
  0..nx/2 (kx >= 0), ``z`` in 0..nz-1.  Its bytes are exactly the reference's
  ``data1(nz, nx/2+1, ny+1, 3)`` (``PFC/data.f90:140``, shape from
  ``PFC/setup.f90:447``).
* RK register ``R[c, j, x, z]`` and pressure ``P[j, x, z]``; the reference keeps
  them as ``data2(ny+1,3,nz,nxh)`` / ``data4(ny+1,nz,nxh)``
  (``PFC/setup.f90:448-449``); ``to_fortran_data2`` / ``to_fortran_data4``
  convert.
* Physical planes ``phys[zz, xx]`` real, the reference's ``data(nx2, nz2)``.

Pinning (tests/test_oracle_*.py): the grid against ``PFC/makegrid.m``; the dt
rule against ``PFC/diagnostics.txt:98``; transforms against a brute-force DFT
with FFTW's documented conventions and the c2r/r2c round trip; the divergence
of fluxes against closed forms; the banded solvers against dense
``numpy.linalg.solve``; the projection against its divergence post-condition;
the nonlinear term against an independent brute-force 3-D restatement; the RK3
stepper against its temporal order and mass-flux/divergence invariants; the
linearized step against the Orr-Sommerfeld known answer of
``PFC/setup.f90:1387``.  No function here is left unpinned.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# ---------------------------------------------------------------------------
# RK3 coefficients.  The reference writes them as default-REAL (single
# precision) literal expressions assigned to REAL(8) parameters
# (PFC/pass2.f90:14-18), so each is the float32 quotient widened to double.
# ---------------------------------------------------------------------------


def _f32q(a: float, b: float) -> float:
    return float(np.float32(a) / np.float32(b))


RKA = (0.0, _f32q(-17.0, 60.0), _f32q(-5.0, 12.0))
RKB = (_f32q(8.0, 15.0), _f32q(5.0, 12.0), _f32q(3.0, 4.0))
RKC = (_f32q(29.0, 96.0), _f32q(-3.0, 40.0), _f32q(1.0, 6.0))
RKD = (_f32q(37.0, 160.0), _f32q(5.0, 24.0), _f32q(1.0, 6.0))

CFL_FOURIER = 0.55  # PFC/pass1.f90:207
CFL_FD = 1.73


# ---------------------------------------------------------------------------
# Grid, operators and wavenumbers (PFC/setup.f90:541-754, 1363-1466)
# ---------------------------------------------------------------------------
@dataclass
class Grid:
    nx: int
    ny: int
    nz: int
    nx2: int
    nz2: int
    lx: float
    lz: float
    re: float
    beta: float = 0.0           # radians (read_parms scales by asin(1), PFC/setup.f90:104-105)
    fix_mass: int = 1
    dt_flag: int = 0
    cfl_max: float = 0.95
    stretch: float = 2.75
    uniform_half: bool = False  # which_test==6 grid (PFC/setup.f90:608-614)
    ctrlu: float = 0.0
    ctrlw: float = 0.0
    # derived
    y: np.ndarray = field(init=False, repr=False)
    yh: np.ndarray = field(init=False, repr=False)

    def __post_init__(self):
        build_operators(self)

    @property
    def nxh(self) -> int:
        return self.nx // 2 + 1


def build_operators(g: Grid) -> Grid:
    """define_coordinates (PFC/setup.f90:541-754) and the wavenumber part of
    define_parameters (PFC/setup.f90:1439-1466) for the channel grid
    (crossflow == 0, bl_grid == 0).  Arrays are 0-based: ``y[j-1]`` is the
    reference's ``y(j)``."""
    ny = g.ny
    half, one, two = 0.5, 1.0, 2.0
    j = np.arange(1, ny + 1, dtype=np.float64)
    xi = two * (j - 1.0) / float(ny - 1) - one
    y = np.tanh(g.stretch * xi) / math.tanh(g.stretch)           # :601-606
    if g.uniform_half:                                             # :608-614
        y = np.empty(ny)
        y[0] = -half
        y[1:ny - 1] = y[0] + np.arange(1, ny - 1) / float(ny - 1)
        y[ny - 1] = half
    yh = np.empty(ny + 1)
    yh[0] = y[0]
    yh[1:ny] = half * (y[:ny - 1] + y[1:])                         # :621-623
    yh[ny] = y[ny - 1]

    hh = np.empty(ny + 1)                                           # :638-641
    hh[0] = y[1] - y[0]
    hh[1:ny] = y[1:ny] - y[:ny - 1]
    hh[ny] = y[ny - 1] - y[ny - 2]
    dh = one / hh
    hn = yh[1:] - yh[:-1]                                           # :643-644
    dn = one / hn
    iu = np.empty((2, ny))                                          # :648-649
    iu[0] = half * hh[:ny] / hn
    iu[1] = half * hh[1:] / hn

    dm = np.zeros((3, 4))                                           # :662-679 (dm(i,k) -> dm[i-1,k-1])
    dm[0, 0] = -dn[0] * (two * hn[0] + hn[1]) / (hn[0] + hn[1])
    dm[1, 0] = dn[0] * dn[1] * (hn[0] + hn[1])
    dm[2, 0] = -dn[1] * hn[0] / (hn[0] + hn[1])
    dm[0, 1] = dn[ny - 2] * hn[ny - 1] / (hn[ny - 2] + hn[ny - 1])
    dm[1, 1] = -dn[ny - 1] * dn[ny - 2] * (hn[ny - 1] + hn[ny - 2])
    dm[2, 1] = dn[ny - 1] * (two * hn[ny - 1] + hn[ny - 2]) / (hn[ny - 1] + hn[ny - 2])
    dm[0, 2] = (y[0] - y[1] + y[0] - y[2]) / ((y[0] - y[1]) * (y[0] - y[2]))
    dm[1, 2] = (y[0] - y[2]) / ((y[1] - y[0]) * (y[1] - y[2]))
    dm[2, 2] = (y[0] - y[1]) / ((y[2] - y[0]) * (y[2] - y[1]))
    dm[0, 3] = (y[ny - 1] - y[ny - 2]) / ((y[ny - 3] - y[ny - 1]) * (y[ny - 3] - y[ny - 2]))
    dm[1, 3] = (y[ny - 1] - y[ny - 3]) / ((y[ny - 2] - y[ny - 1]) * (y[ny - 2] - y[ny - 3]))
    dm[2, 3] = (y[ny - 1] - y[ny - 2] + y[ny - 1] - y[ny - 3]) / (
        (y[ny - 1] - y[ny - 2]) * (y[ny - 1] - y[ny - 3]))

    d2n = np.zeros((3, ny))                                         # :704-726
    d2h = np.zeros((3, ny + 1))
    for p in range(2, ny):  # reference plane index p = 2..ny-1 -> array p-1
        a = p - 1
        d2n[0, a] = dn[a] * dh[a]
        d2n[1, a] = -dn[a] * dh[a] - dn[a] * dh[a + 1]
        d2n[2, a] = dn[a] * dh[a + 1]
        d2h[0, a] = dh[a] * dn[a - 1]
        d2h[1, a] = -dh[a] * dn[a - 1] - dh[a] * dn[a]
        d2h[2, a] = dh[a] * dn[a]
    d2h[0, 1] = -dh[1] * dm[0, 0]
    d2h[1, 1] = -dh[1] * dn[1] - dh[1] * dm[1, 0]
    d2h[2, 1] = dh[1] * dn[1] - dh[1] * dm[2, 0]
    d2h[0, ny - 1] = dh[ny - 1] * dn[ny - 2] + dh[ny - 1] * dm[0, 1]
    d2h[1, ny - 1] = -dh[ny - 1] * dn[ny - 2] + dh[ny - 1] * dm[1, 1]
    d2h[2, ny - 1] = dh[ny - 1] * dm[2, 1]
    d2h[:, 0] = d2h[:, 1]
    d2n[:, 0] = d2n[:, 1]
    d2n[:, ny - 1] = d2n[:, ny - 2]
    d2h[:, ny] = d2h[:, ny - 1]

    dyh = np.zeros(ny + 1)                                          # :747-749
    dyh[1:ny] = hh[1:ny]
    dy = np.zeros(ny + 1)                                           # :751-754
    dy[0] = half * hh[1]
    dy[1:ny - 1] = half * (hh[1:ny - 1] + hh[2:ny])
    dy[ny - 1] = half * hh[ny - 1]

    pi = math.atan2(0.0, -1.0)                                      # :1373
    kx = np.arange(g.nx // 2 + 1, dtype=np.float64) * 2.0 * pi / g.lx   # :1447-1450
    kz = np.zeros(g.nz)
    kz[:g.nz // 2 + 1] = np.arange(g.nz // 2 + 1, dtype=np.float64) * 2.0 * pi / g.lz
    for z in range(g.nz, g.nz // 2 + 1, -1):                          # :1455-1458
        kz[z - 1] = -kz[g.nz + 2 - z - 1]
    k2 = kx[:, None] * kx[:, None] + kz[None, :] * kz[None, :]       # :1459-1463, [x, z]

    g.y, g.yh, g.hh, g.dh, g.hn, g.dn, g.iu = y, yh, hh, dh, hn, dn, iu
    g.dm, g.d2n, g.d2h, g.dyh, g.dy = dm, d2n, d2h, dyh, dy
    g.kx, g.kz, g.k2 = kx, kz, k2
    g.ikx = 1j * kx
    g.ikz = 1j * kz
    g.buffer = np.zeros((g.nz2, g.nx2))
    g.target_inflow = np.zeros((ny + 1, 3))
    return g


def default_dt(g: Grid) -> float:
    """Initial dt rule 0.25*lx/nx (PFC/setup.f90:1377)."""
    return 0.25 * g.lx / float(g.nx)


def make_grid(nx, ny, nz, nx2, nz2, lx_over_pi, aspect, re, **kw) -> Grid:
    """read_parms box convention: lx = lx*pi, lz = lx/aspect (PFC/setup.f90:63-69)."""
    pi = math.atan2(0.0, -1.0)
    lx = lx_over_pi * pi
    return Grid(nx, ny, nz, nx2, nz2, lx, lx / aspect, re, **kw)


# ---------------------------------------------------------------------------
# Layout helpers
# ---------------------------------------------------------------------------
def to_fortran_data2(R):
    """R[c,j,x,z] -> bytes-order of data2(ny+1,3,nz,nxh) as C array [x,z,c,j]."""
    return np.ascontiguousarray(R.transpose(2, 3, 0, 1))


def from_fortran_data2(D2):
    return np.ascontiguousarray(D2.transpose(2, 3, 0, 1))


def to_fortran_data4(P):
    """P[j,x,z] -> data4(ny+1,nz,nxh) as C array [x,z,j]."""
    return np.ascontiguousarray(P.transpose(1, 2, 0))


def from_fortran_data4(D4):
    return np.ascontiguousarray(D4.transpose(2, 0, 1))


# ---------------------------------------------------------------------------
# FFTW-convention helpers.  Forward = exp(-i...), backward = exp(+i...), both
# unnormalised (PFC/transform.f90:35).  c2r of length n ignores the imaginary
# parts of the DC and (even n) Nyquist bins, as FFTW's c2r does.
# ---------------------------------------------------------------------------
def _c2r_last(Y, n):
    """FFTW unnormalised c2r along the last axis; Y has n//2+1 bins."""
    Y = np.array(Y, dtype=np.complex128, copy=True)
    Y[..., 0] = Y[..., 0].real
    if n % 2 == 0:
        Y[..., n // 2] = Y[..., n // 2].real
    return np.fft.irfft(Y, n=n, axis=-1) * n


def _bwd(X, axis):
    n = X.shape[axis]
    return np.fft.ifft(X, axis=axis) * n


def _fwd(X, axis):
    return np.fft.fft(X, axis=axis)


# ---------------------------------------------------------------------------
# Dealiased x-z transforms (PFC/transform.f90:195-227, 258-409, 573-690)
# ---------------------------------------------------------------------------
def xz_c2r(g: Grid, C):
    """xz_fft_complex_to_real_parallel (PFC/transform.f90:339-409).
    C[..., x, z] complex (nxh, nz)  ->  real [..., nz2, nx2]."""
    nz, nz2, nx, nx2, nxh = g.nz, g.nz2, g.nx, g.nx2, g.nxh
    C = np.asarray(C)
    lead = C.shape[:-2]
    X = np.zeros(lead + (nxh, nz2), dtype=np.complex128)
    if nz2 > nz:                                                    # :345-357
        X[..., :nz // 2] = C[..., :nz // 2]
        X[..., nz // 2] = 0.5 * C[..., nz // 2]
        X[..., nz2 - nz // 2] = 0.5 * np.conj(C[..., nz // 2])
        X[..., nz2 - nz // 2 + 1:] = C[..., nz // 2 + 1:]
        if nx2 > nx:
            X[..., nxh - 1, :] *= 0.5
    else:                                                           # :358-365
        X[..., :nz // 2 + 1] = C[..., :nz // 2 + 1]
        X[..., nz2 - nz // 2 + 1:] = C[..., nz // 2 + 1:]
    Z = _bwd(X, axis=-1)                                            # :368 z backward c2c
    Zt = np.swapaxes(Z, -1, -2)                                     # [..., zz, x]
    Y = np.zeros(lead + (nz2, nx2 // 2 + 1), dtype=np.complex128)
    Y[..., :nxh] = Zt                                               # :396-400 zero pad in x
    return _c2r_last(Y, nx2)                                        # :404 x c2r


def _x_r2c_kept(g: Grid, f):
    """x r2c (PFC/transform.f90:265-276) keeping kx < nxh: [..., zz, xx] -> [..., zz, x]."""
    return np.fft.rfft(f, axis=-1)[..., :g.nxh]


def _z_fold(g: Grid, Dz):
    """Truncate nz2 -> nz with the Nyquist fold (PFC/transform.f90:322-330).
    Dz[..., x, kz'] (length nz2) -> [..., x, z] (length nz); unscaled."""
    nz, nz2 = g.nz, g.nz2
    off = nz2 - nz
    out = np.empty(Dz.shape[:-1] + (nz,), dtype=np.complex128)
    out[..., :nz // 2] = Dz[..., :nz // 2]
    out[..., nz // 2] = Dz[..., nz // 2] + np.conj(Dz[..., off + nz // 2])
    out[..., nz // 2 + 1:] = Dz[..., off + nz // 2 + 1:nz2]
    return out


def xz_r2c(g: Grid, f):
    """xz_fft_real_to_complex_parallel (PFC/transform.f90:258-335).
    real [..., nz2, nx2] -> complex [..., x, z]."""
    F = _x_r2c_kept(g, np.asarray(f, dtype=np.float64))             # [..., zz, x]
    Dz = _fwd(np.swapaxes(F, -1, -2), axis=-1)                     # [..., x, kz'] z forward
    scale = 1.0 / float(g.nx2 * g.nz2)
    out = scale * _z_fold(g, Dz)
    out[..., g.nxh - 1, :] *= 2.0                                   # :331-333
    return out


def div_fluxes(g: Grid, fx, fy, fz):
    """parallel_divergence_of_fluxes (PFC/transform.f90:573-690):
    ikx*F(fx) + F(fy) + ikz*F(fz) with the reference's Nyquist conventions."""
    Fy = _x_r2c_kept(g, np.asarray(fy, dtype=np.float64))
    Fx = _x_r2c_kept(g, np.asarray(fx, dtype=np.float64))
    A = Fy + g.ikx * Fx                                             # :613 (ikx per kx column)
    Dz = _fwd(np.swapaxes(A, -1, -2), axis=-1)                     # :641
    C = xz_r2c(g, fz)                                               # :646
    scale = 1.0 / float(g.nx2 * g.nz2)
    out = g.ikz * C + scale * _z_fold(g, Dz)                        # :655-687
    out[..., g.nxh - 1, :] += scale * _z_fold(g, Dz)[..., g.nxh - 1, :]  # two*scale on the last column
    return out


def small_c2r(g: Grid, C):
    """small_fft_complex_to_real (PFC/transform.f90:950-1000): no padding, no
    Nyquist split. C[..., x, z] -> real [..., nz, nx]."""
    Z = _bwd(np.asarray(C, dtype=np.complex128), axis=-1)           # z backward, length nz
    Zt = np.swapaxes(Z, -1, -2)                                     # [..., zz, x]
    return _c2r_last(Zt, g.nx)


def small_r2c(g: Grid, f):
    """small_fft_real_to_complex (PFC/transform.f90:1071-1142): scale 1/(nx*nz)."""
    F = np.fft.rfft(np.asarray(f, dtype=np.float64), axis=-1)      # [..., zz, nx/2+1]
    Dz = _fwd(np.swapaxes(F, -1, -2), axis=-1)
    return Dz / float(g.nx * g.nz)


# ---------------------------------------------------------------------------
# y stencils (PFC/diff_int.f90:12-57)
# ---------------------------------------------------------------------------
def d2_dy(f, d):
    """d2_dy (PFC/diff_int.f90:31-46) along axis 0; f[n, ...], d[3, n]."""
    n = f.shape[0]
    sh = (slice(None),) + (None,) * (f.ndim - 1)
    df = np.empty_like(f)
    df[0] = d[0, 0] * f[0] + d[1, 0] * f[1] + d[2, 0] * f[2]
    df[1:n - 1] = d[0, 1:n - 1][sh] * f[:n - 2] + d[1, 1:n - 1][sh] * f[1:n - 1] + d[2, 1:n - 1][sh] * f[2:n]
    df[n - 1] = d[0, n - 1] * f[n - 3] + d[1, n - 1] * f[n - 2] + d[2, n - 1] * f[n - 1]
    return df


# ---------------------------------------------------------------------------
# Banded y solvers (PFC/solve.f90), vectorised over trailing axes
# ---------------------------------------------------------------------------
def _thomas(a, b, c, f):
    """Reciprocal-form Thomas elimination of PFC/solve.f90:246-262."""
    a = a.copy(); b = b.copy(); c = c.copy(); f = f.copy()
    n = f.shape[0]
    for i in range(1, n):
        b[i - 1] = 1.0 / b[i - 1]
        c[i - 1] = c[i - 1] * b[i - 1]
        f[i - 1] = f[i - 1] * b[i - 1]
        b[i] = b[i] - a[i] * c[i - 1]
        f[i] = f[i] - a[i] * f[i - 1]
    b[n - 1] = 1.0 / b[n - 1]
    f[n - 1] = f[n - 1] * b[n - 1]
    for i in range(n - 2, -1, -1):
        f[i] = f[i] - c[i] * f[i + 1]
    return f


def solve_helmholtz(g: Grid, n, f, alpha, diff, d2, topbc, botbc, top_neumann=0, bot_neumann=0):
    """SOLVE_HELMHOLTZ (PFC/solve.f90:154-264) for (alpha - diff*D2) x = f with
    the Dirichlet wall rows of the channel (TOP_NEUMANN = BOT_NEUMANN =
    bl_grid = 0).  The Neumann rows are only reachable with bl_grid != 0
    (boundary-layer grids, out of scope); there the reference's interior
    assignment at :204-206 overwrites the modified second row and the n == ny
    branch reads an uninitialised c(2), so they are not restated.
    f[n, ...] complex; alpha may be an array broadcast over trailing axes."""
    if top_neumann or bot_neumann:
        raise NotImplementedError("Neumann wall rows (bl_grid != 0) are out of scope")
    f = np.array(f, dtype=np.complex128, copy=True)
    tail = f.shape[1:]
    alpha = np.broadcast_to(np.asarray(alpha, dtype=np.float64), tail) if tail else float(alpha)
    sh = (slice(None),) + (None,) * len(tail)
    a = np.zeros((n,) + tail); b = np.zeros((n,) + tail); c = np.zeros((n,) + tail)
    b[0] = 1.0                                                      # :174-178
    f[0] = botbc
    a[1:n - 1] = (-diff * d2[0, 1:n - 1])[sh]                       # :204-206
    b[1:n - 1] = (-diff * d2[1, 1:n - 1])[sh] + alpha
    c[1:n - 1] = (-diff * d2[2, 1:n - 1])[sh]
    b[n - 1] = 1.0                                                  # :209-213
    f[n - 1] = topbc
    return _thomas(a, b, c, f)


def solve_p(g: Grid, p, alpha, topbc=0.0, botbc=0.0):
    """SOLVE_P (PFC/solve.f90:269-338): Neumann Poisson plus the wall
    re-extrapolation with DM.  p[ny+1, ...]."""
    ny = g.ny
    dn, dh, d2h, dm = g.dn, g.dh, g.d2h, g.dm
    p = np.array(p, dtype=np.complex128, copy=True)
    tail = p.shape[1:]
    alpha = np.broadcast_to(np.asarray(alpha, dtype=np.float64), tail) if tail else float(alpha)
    sh = (slice(None),) + (None,) * len(tail)
    n = ny + 1
    a = np.zeros((n,) + tail); b = np.zeros((n,) + tail); c = np.zeros((n,) + tail)
    a[0] = 0.0; b[0] = -dn[0]; c[0] = dn[0]                         # :280-282
    a[1] = dh[1] * dn[0]                                            # :288-290
    b[1] = -dh[1] * dn[0] - dh[1] * dn[1] + alpha
    c[1] = dh[1] * dn[1]
    a[2:ny - 1] = d2h[0, 2:ny - 1][sh]                              # :292-296
    b[2:ny - 1] = d2h[1, 2:ny - 1][sh] + alpha
    c[2:ny - 1] = d2h[2, 2:ny - 1][sh]
    a[ny - 1] = dh[ny - 1] * dn[ny - 2]                             # :298-300
    b[ny - 1] = -dh[ny - 1] * dn[ny - 2] - dh[ny - 1] * dn[ny - 1] + alpha
    c[ny - 1] = dh[ny - 1] * dn[ny - 1]
    a[ny] = -dn[ny - 1]; b[ny] = dn[ny - 1]; c[ny] = 0.0            # :306-308
    p[0] = botbc                                                    # :312-315
    p[ny] = topbc
    p = _thomas(a, b, c, p)
    p[0] = (botbc - dm[1, 0] * p[1] - dm[2, 0] * p[2]) / dm[0, 0]   # :335-336
    p[ny] = (topbc - dm[0, 1] * p[ny - 2] - dm[1, 1] * p[ny - 1]) / dm[2, 1]
    return p


def subtract_pressure_gradient(g: Grid, cu, cp, coef):
    """subtract_pressure_gradient (PFC/solve.f90:54-77).  cu[3, ny+1, x, z],
    cp[ny+1, x, z]; coef scalar.  Returns a new array."""
    ny = g.ny
    cu = cu.copy()
    ikx = g.ikx[:, None]
    ikz = g.ikz[None, :]
    dp = g.dn[:, None, None] * (cp[1:] - cp[:-1])                   # d_dy(ny+1,...,ny,dp,dn)
    cu[0, 1:ny] = cu[0, 1:ny] - coef * ikx * cp[1:ny]
    cu[1, 1:ny - 1] = cu[1, 1:ny - 1] - coef * dp[1:ny - 1]
    cu[2, 1:ny] = cu[2, 1:ny] - coef * ikz * cp[1:ny]
    return cu


def divergence(g: Grid, cu):
    """ikx u + d_dy(v; dh) + ikz w at j = 2..ny (PFC/solve.f90:28-30,
    PFC/pass2.f90:390-394).  Returns [ny-1, x, z]."""
    ny = g.ny
    dv = g.dh[1:ny, None, None] * (cu[1, 1:ny] - cu[1, :ny - 1])
    return g.ikx[:, None] * cu[0, 1:ny] + dv + g.ikz[None, :] * cu[2, 1:ny]


def strip_divergence(g: Grid, cu, coef):
    """strip_divergence (PFC/solve.f90:14-50) for every non-mean mode.
    Returns (projected cu, cp).  The mean mode is returned untouched with cp=0."""
    ny = g.ny
    tmp = 1.0 / coef
    cp = np.zeros((ny + 1,) + cu.shape[2:], dtype=np.complex128)
    cp[1:ny] = tmp * divergence(g, cu)
    alpha = -g.k2
    mean = g.k2 == 0.0
    safe_alpha = np.where(mean, -1.0, alpha)  # mean mode handled separately by callers
    cp = solve_p(g, cp, safe_alpha, 0.0, 0.0)
    cp[:, mean] = 0.0
    out = subtract_pressure_gradient(g, cu, cp, coef)
    out[:, :, mean] = cu[:, :, mean]
    return out, cp


# ---------------------------------------------------------------------------
# Nonlinear term (PFC/pass1.f90:26-252) and its linearized variant
# (PFC/pass1_linearized.f90:135-509)
# ---------------------------------------------------------------------------
class CFLAbort(RuntimeError):
    pass


def _cfl_and_dt(g: Grid, step, rk_step, dt, amax, vmax, wmax):
    """check_cfl (PFC/pass1.f90:202-250) and the dt update (:107-124).
    amax/vmax/wmax indexed by reference plane k=2..ny (array k-2)."""
    ny = g.ny
    cfl = np.zeros(ny + 1)
    for k in range(2, ny + 1):
        tempa = 0.5 * amax[k - 2] * float(g.nx) / (CFL_FOURIER * g.lx)
        tempb = vmax[k - 2] / (CFL_FD * min(g.hn[k - 1], g.hn[k - 2]))
        tempc = 0.5 * wmax[k - 2] * float(g.nz) / (CFL_FOURIER * g.lz)
        cfl[k] = (tempa + tempb + tempc) * dt
        if g.dt_flag == 1:
            if cfl[k] >= 3.0e2:
                raise CFLAbort("DT IS TOO BIG")
        elif cfl[k] >= 1.15:
            raise CFLAbort("CFL IS TOO BIG")
    cfl_temp = cfl[2:ny + 1].max()
    if step >= 0 and g.dt_flag == 1:
        dt_temp = dt * g.cfl_max / cfl_temp
        dt = min(dt_temp, dt + (dt_temp - dt) * dt / (1.0 + dt))
    return dt, cfl_temp


def nonlinear(g: Grid, U, step=1, rk_step=1, dt=0.0, return_cfl=False, tau=None):
    """nonlinear (PFC/pass1.f90:26-252), vectorised over planes.  The
    reference overwrites data1 plane k only after planes k and k+1 have been
    transformed (:65-83), so the result equals a computation from the
    original field, which is what this does.  Returns (NL, dt[, cfl]).

    tau (no counterpart in the reference): a subgrid stress [6, ny+1, nz2,
    nx2] at the cell centres (components 11, 12, 13, 22, 23, 33, zero on the
    wall planes) added to each flux u_i u_j where the flux is formed;
    tau_12 and tau_23 reach the interior v-points y(k) with the iu weights
    and are zero at the walls."""
    ny = g.ny
    ph = xz_c2r(g, U)                       # [3, ny+1, nz2, nx2]
    u, v, w = ph[0], ph[1], ph[2]
    buf = g.buffer[None]
    tgt = g.target_inflow
    half = 0.5
    # y-face fluxes of u and w at y(k), k = 1..ny  (:53,56,138-146)
    Fu = np.empty((ny,) + u.shape[1:]); Fw = np.empty_like(Fu)
    Fu[0] = v[0] * u[0]
    Fw[0] = v[0] * w[0]
    Fu[1:ny - 1] = v[1:ny - 1] * (half * (u[1:ny - 1] + u[2:ny]))
    Fw[1:ny - 1] = v[1:ny - 1] * (half * (w[1:ny - 1] + w[2:ny]))
    Fu[ny - 1] = v[ny - 1] * u[ny]
    Fw[ny - 1] = v[ny - 1] * w[ny]
    # v flux at yh(j), j = 2..ny: G[j-2]  (:54-55,184-185)
    tmp = half * (v[:ny - 1] + v[1:ny])
    G = tmp * tmp
    t12y = t23y = 0.0
    if tau is not None:
        tau = np.asarray(tau, dtype=np.float64)
        jv = np.arange(1, ny - 1)
        a1 = g.iu[0, jv][:, None, None]; a2 = g.iu[1, jv][:, None, None]
        t12y = np.zeros_like(Fu); t23y = np.zeros_like(Fu)
        t12y[jv] = a1 * tau[1, jv] + a2 * tau[1, jv + 1]
        t23y[jv] = a1 * tau[4, jv] + a2 * tau[4, jv + 1]
        Fu = Fu + t12y
        Fw = Fw + t23y
        G = G + tau[3, 1:ny]
    NL = np.zeros_like(U)
    ks = np.arange(2, ny + 1)              # planes 2..ny for u,w
    kk = ks - 1                            # array index of plane k
    dh = g.dh[kk][:, None, None]
    yfu = dh * (Fu[kk] - Fu[kk - 1]) + buf * (u[kk] - tgt[kk, 0][:, None, None])
    yfw = dh * (Fw[kk] - Fw[kk - 1]) + buf * (w[kk] - tgt[kk, 2][:, None, None])
    fuu, fwu, fww = u[kk] * u[kk], w[kk] * u[kk], w[kk] * w[kk]
    if tau is not None:
        fuu = fuu + tau[0, kk]; fwu = fwu + tau[2, kk]; fww = fww + tau[5, kk]
    NL[0, kk] = div_fluxes(g, fuu, yfu, fwu)                         # :158-162
    NL[2, kk] = div_fluxes(g, fwu, yfw, fww)                         # :163-167
    kv = np.arange(2, ny) - 1              # planes 2..ny-1 for v
    dn = g.dn[kv][:, None, None]
    yfv = dn * (G[kv] - G[kv - 1]) + buf * (v[kv] - tgt[kv, 1][:, None, None])   # :188-189
    i1 = g.iu[0, kv][:, None, None]; i2 = g.iu[1, kv][:, None, None]
    fvu = v[kv] * (i1 * u[kv] + i2 * u[kv + 1])
    fvw = v[kv] * (i1 * w[kv] + i2 * w[kv + 1])
    if tau is not None:
        fvu = fvu + t12y[kv]; fvw = fvw + t23y[kv]
    NL[1, kv] = div_fluxes(g, fvu, yfv, fvw)                         # :190-195
    # planes 1, ny (v) and ny+1 are zero (:62,96-99)
    if rk_step == 1:
        amax = np.abs(u[1:ny] + u[2:ny + 1]).max(axis=(1, 2))
        vmax = np.abs(v[1:ny]).max(axis=(1, 2))
        wmax = np.abs(w[1:ny] + w[2:ny + 1]).max(axis=(1, 2))
        dt, cfl = _cfl_and_dt(g, step, rk_step, dt, amax, vmax, wmax)
    else:
        cfl = None
    return (NL, dt, cfl) if return_cfl else (NL, dt)


def nonlinear_linearized(g: Grid, U, base, step=1, rk_step=1, dt=0.0, return_cfl=False):
    """nonlinear_linearized (PFC/pass1_linearized.f90:135-509) about the
    physical base flow base[c, j, zz, xx] (the reference's baseflow array)."""
    ny = g.ny
    ph = xz_c2r(g, U)
    u, v, w = ph[0], ph[1], ph[2]
    Bu, Bv, Bw = base[0], base[1], base[2]
    buf = g.buffer[None]
    tgt = g.target_inflow
    half, two = 0.5, 2.0
    Fu = np.empty((ny,) + u.shape[1:]); Fw = np.empty_like(Fu)
    Fu[0] = Bv[0] * u[0] + v[0] * Bu[0]                              # :160-162
    Fw[0] = Bv[0] * w[0] + v[0] * Bw[0]
    Fu[1:ny - 1] = (Bv[1:ny - 1] * half) * (u[1:ny - 1] + u[2:ny]) \
        + (v[1:ny - 1] * half) * (Bu[1:ny - 1] + Bu[2:ny])          # :265-273
    Fw[1:ny - 1] = (Bv[1:ny - 1] * half) * (w[1:ny - 1] + w[2:ny]) \
        + (v[1:ny - 1] * half) * (Bw[1:ny - 1] + Bw[2:ny])
    Fu[ny - 1] = Bv[ny - 1] * u[ny] + v[ny - 1] * Bu[ny]            # :275-280
    Fw[ny - 1] = Bv[ny - 1] * w[ny] + v[ny - 1] * Bw[ny]
    tmp = half * (v[:ny - 1] + v[1:ny])
    tb = half * (Bv[:ny - 1] + Bv[1:ny])
    G = (two * tmp) * tb                                            # :163-166,321-324
    NL = np.zeros_like(U)
    kk = np.arange(2, ny + 1) - 1
    dh = g.dh[kk][:, None, None]
    yfu = dh * (Fu[kk] - Fu[kk - 1]) + buf * (u[kk] - tgt[kk, 0][:, None, None])
    yfw = dh * (Fw[kk] - Fw[kk - 1]) + buf * (w[kk] - tgt[kk, 2][:, None, None])
    cross = Bw[kk] * u[kk] + w[kk] * Bu[kk]
    NL[0, kk] = div_fluxes(g, (two * Bu[kk]) * u[kk], yfu, cross)   # :293-298
    NL[2, kk] = div_fluxes(g, cross, yfw, (two * Bw[kk]) * w[kk])   # :299-304
    kv = np.arange(2, ny) - 1
    dn = g.dn[kv][:, None, None]
    yfv = dn * (G[kv] - G[kv - 1]) + buf * (v[kv] - tgt[kv, 1][:, None, None])
    i1 = g.iu[0, kv][:, None, None]; i2 = g.iu[1, kv][:, None, None]
    fx = Bv[kv] * (i1 * u[kv] + i2 * u[kv + 1]) + v[kv] * (i1 * Bu[kv] + i2 * Bu[kv + 1])
    fz = Bv[kv] * (i1 * w[kv] + i2 * w[kv + 1]) + v[kv] * (i1 * Bw[kv] + i2 * Bw[kv + 1])
    NL[1, kv] = div_fluxes(g, fx, yfv, fz)                          # :329-339
    if rk_step == 1:
        amax = np.maximum(np.abs(u[1:ny] + u[2:ny + 1]).max(axis=(1, 2)),
                          np.abs(Bu[1:ny] + Bu[2:ny + 1]).max(axis=(1, 2)))
        vmax = np.maximum(np.abs(v[1:ny]).max(axis=(1, 2)), np.abs(Bv[1:ny]).max(axis=(1, 2)))
        wmax = np.maximum(np.abs(w[1:ny] + w[2:ny + 1]).max(axis=(1, 2)),
                          np.abs(Bw[1:ny] + Bw[2:ny + 1]).max(axis=(1, 2)))
        dt, cfl = _cfl_and_dt(g, step, rk_step, dt, amax, vmax, wmax)
    else:
        cfl = None
    return (NL, dt, cfl) if return_cfl else (NL, dt)


def laminar_baseflow(g: Grid):
    """init_baseflow laminar branch (PFC/pass1_linearized.f90:39-46)."""
    ny = g.ny
    base = np.zeros((3, ny + 1, g.nz2, g.nx2))
    prof = 1.0 - g.yh ** 2
    base[0] = (prof * math.cos(g.beta))[:, None, None]
    base[2] = (prof * math.sin(g.beta))[:, None, None]
    return base


# ---------------------------------------------------------------------------
# Fourier-space RK3/CN solve (PFC/pass2.f90:29-375)
# ---------------------------------------------------------------------------
@dataclass
class State:
    U: np.ndarray          # data1 (velocity between calls, NL after nonlinear)
    R: np.ndarray          # data2
    P: np.ndarray          # data4
    dt: float
    p_grad: float
    d_um_dy: np.ndarray    # (ny,)
    t_total: float = 0.0
    cbc_top: np.ndarray = None   # [3, x, z]
    cbc_bot: np.ndarray = None

    def copy(self):
        return State(self.U.copy(), self.R.copy(), self.P.copy(), self.dt, self.p_grad,
                     self.d_um_dy.copy(), self.t_total,
                     None if self.cbc_top is None else self.cbc_top.copy(),
                     None if self.cbc_bot is None else self.cbc_bot.copy())


def _explicit_diffusion(g: Grid, vis_exp, cvel):
    """add_explicit_diffusive_terms increment (PFC/pass2.f90:339-351)."""
    ny = g.ny
    k2 = g.k2[None]
    inc = np.zeros_like(cvel)
    inc[0] = vis_exp * (d2_dy(cvel[0], g.d2h) - k2 * cvel[0])
    inc[1, :ny] = vis_exp * (d2_dy(cvel[1, :ny], g.d2n) - k2 * cvel[1, :ny])
    inc[2] = vis_exp * (d2_dy(cvel[2], g.d2h) - k2 * cvel[2])
    return inc


def d_um_dy_of(g: Grid, ucol):
    """Mean-gradient update (PFC/pass2.f90:84-89, PFC/init_vel.f90:333-342);
    ucol = mean-mode u profile (ny+1,) complex."""
    ny = g.ny
    dm, dn = g.dm, g.dn
    out = np.empty(ny)
    out[0] = (dm[0, 0] * ucol[0] + dm[1, 0] * ucol[1] + dm[2, 0] * ucol[2]).real
    out[1:ny - 1] = (dn[1:ny - 1] * (ucol[2:ny] - ucol[1:ny - 1])).real
    out[ny - 1] = (dm[0, 1] * ucol[ny - 2] + dm[1, 1] * ucol[ny - 1] + dm[2, 1] * ucol[ny]).real
    return out


def rhs_solve(g: Grid, st: State, rk_step: int):
    """rhs_solve + PHASE1 (PFC/pass2.f90:29-352) for every (kx,kz) at once.
    st.U must hold the nonlinear term; on return it holds the new velocity."""
    ny = g.ny
    r = rk_step - 1
    dt, re = st.dt, g.re
    nxh, nz = g.nxh, g.nz
    cbc_top = st.cbc_top if st.cbc_top is not None else np.zeros((3, nxh, nz), np.complex128)
    cbc_bot = st.cbc_bot if st.cbc_bot is not None else np.zeros((3, nxh, nz), np.complex128)
    cu = -st.U                                                      # :67
    cru = st.R
    cp = st.P
    k2 = g.k2
    mean = k2 == 0.0
    vis_imp = RKD[r] * (dt / re)                                    # :198-199
    rk_press = (RKC[r] + RKD[r]) * dt
    coef = RKB[r] * dt                                              # :202-203
    cv = cru + coef * cu
    if rk_step == 1:                                                # :205-214
        vis_exp = RKC[r] * (dt / re)
        cv = cv + _explicit_diffusion(g, vis_exp, cru)
        nm = subtract_pressure_gradient(g, cv, cp, rk_press)
        cv = np.where(mean[None, None], cv, nm)
    # --- mean mode (:217-263)
    cvm = cv[:, :, 0, 0].copy()
    cum = cu[:, :, 0, 0]
    p_grad = st.p_grad
    if g.fix_mass == 1:                                             # compute_p_grad :356-375
        tmpv = 0.0
        for plane in range(1, ny):
            tmpv = tmpv + g.dyh[plane] * cum[0, plane].real
        p_grad = -tmpv + 0.5 * (st.d_um_dy[ny - 1] - st.d_um_dy[0]) / re
    cvm[0] = cvm[0] - math.cos(g.beta) * rk_press * p_grad
    cvm[2] = cvm[2] - math.sin(g.beta) * rk_press * p_grad
    cvm[0] = solve_helmholtz(g, ny + 1, cvm[0], 1.0, vis_imp, g.d2h, cbc_top[0, 0, 0], cbc_bot[0, 0, 0])
    cvm[2] = solve_helmholtz(g, ny + 1, cvm[2], 1.0, vis_imp, g.d2h, cbc_top[2, 0, 0], cbc_bot[2, 0, 0])
    cvm[1, 0] = cbc_bot[1, 0, 0]
    cvm[1, 1:ny] = cvm[1, 0]
    cvm[1, ny] = 0.0
    # --- non-mean modes (:265-310)
    alpha = 1.0 + vis_imp * k2
    alpha_s = np.where(mean, 1.0, alpha)
    new = np.empty_like(cv)
    new[0] = solve_helmholtz(g, ny + 1, cv[0], alpha_s, vis_imp, g.d2h, cbc_top[0], cbc_bot[0])
    new[2] = solve_helmholtz(g, ny + 1, cv[2], alpha_s, vis_imp, g.d2h, cbc_top[2], cbc_bot[2])
    new[1, :ny] = solve_helmholtz(g, ny, cv[1, :ny], alpha_s, vis_imp, g.d2n, cbc_top[1], cbc_bot[1])
    new[1, ny] = 0.0
    new, ch = strip_divergence(g, new, rk_press)
    alpha2 = 1.0 + vis_imp * k2                                     # :307-309
    cp_new = cp + alpha2[None] * ch - vis_imp * d2_dy(ch, g.d2h)
    cp_new[:, mean] = 0.0
    new[:, :, 0, 0] = cvm
    cv = new
    if rk_step < 3:                                                 # :312-328
        coef2 = dt * RKA[r + 1]
        cru_new = cv + coef2 * cu
        vis_exp = (dt / re) * RKC[r + 1]
        cru_new = cru_new + _explicit_diffusion(g, vis_exp, cv)
        rk_press2 = (RKC[r + 1] + RKD[r + 1]) * dt
        nm = subtract_pressure_gradient(g, cru_new, cp_new, rk_press2)
        cru_new = np.where(mean[None, None], cru_new, nm)
    else:
        cru_new = cv.copy()
    st.U = cv
    st.R = cru_new
    st.P = cp_new
    st.p_grad = p_grad
    st.d_um_dy = d_um_dy_of(g, cv[0, :, 0, 0])
    return st


def jet_profile_spec(g: Grid, wall, amplitude, radius, xjet, zjet):
    """Gaussian wall jet (no counterpart in the reference, whose jet_profile,
    PFC/pass2.f90:423-673, is tanh-shaped): v = +-A exp(-r^2/radius^2) about
    (xjet, zjet) on the nx x nz wall grid with periodic images (+ bottom, -
    top: blowing into the channel), small r2c (PFC/transform.f90:1071-1142),
    mean mode removed (uniform suction, no net mass flux).  [nxh, nz]."""
    x = np.arange(g.nx) * g.lx / g.nx - xjet
    z = np.arange(g.nz) * g.lz / g.nz - zjet
    x = x - g.lx * np.floor(x / g.lx + 0.5)
    z = z - g.lz * np.floor(z / g.lz + 0.5)
    sign = 1.0 if wall == 0 else -1.0
    f = sign * amplitude * np.exp(-(z[:, None] ** 2 + x[None, :] ** 2) / (radius * radius))
    C = small_r2c(g, f)
    C[0, 0] = 0.0
    return C


def jet_amplitude(st: State, rk, forcing, freq):
    """1 + forcing sin(2 pi f t) at the end time of substep rk,
    t = t_total + dt sum_{r<=rk}(rkc + rkd) (PFC/pass2.f90:38-39)."""
    s = 0.0
    for r in range(rk):
        s = s + (RKC[r] + RKD[r])
    t = st.t_total + st.dt * s
    return 1.0 + forcing * math.sin(6.283185307179586 * freq * t)


def rk3_step(g: Grid, st: State, step: int, linear_base=None, jet=None):
    """One time step of the timestep loop (PFC/main_mpi.f90:70-87): three
    (nonlinear, rhs_solve) substeps, then t_total += dt.  boundary() with
    zero gains just zeroes cbc (PFC/read_write.f90:1470-1471).  jet = (wall,
    spectral profile, forcing, frequency) sets the wall-normal wall velocity
    of that wall before each rhs_solve, as jet_profile does (:42-43)."""
    boundary(g, st)
    for rk in (1, 2, 3):
        if linear_base is None:
            st.U, st.dt = nonlinear(g, st.U, step=step, rk_step=rk, dt=st.dt)
        else:
            st.U, st.dt = nonlinear_linearized(g, st.U, linear_base, step=step, rk_step=rk, dt=st.dt)
        if jet is not None:
            wall, prof, forcing, freq = jet
            amp = jet_amplitude(st, rk, forcing, freq)
            (st.cbc_bot if wall == 0 else st.cbc_top)[1] = amp * prof
        rhs_solve(g, st, rk)
    st.t_total = st.t_total + st.dt
    return st


def boundary(g: Grid, st: State):
    """boundary (PFC/read_write.f90:1465-1532): wall-velocity feedback law."""
    ny = g.ny
    nxh, nz = g.nxh, g.nz
    st.cbc_top = np.zeros((3, nxh, nz), np.complex128)
    st.cbc_bot = np.zeros((3, nxh, nz), np.complex128)
    dm = g.dm
    if g.ctrlu != 0.0:
        ul2 = xz_c2r(g, st.U[0, 1]); ul3 = xz_c2r(g, st.U[0, 2])
        ubot = (dm[1, 0] * ul2 + dm[2, 0] * ul3 - 2.0) / (g.ctrlu - dm[0, 0])
        st.cbc_bot[0] = xz_r2c(g, ubot)
        ul2 = xz_c2r(g, st.U[0, ny - 1]); ul3 = xz_c2r(g, st.U[0, ny - 2])
        utop = -(dm[1, 1] * ul2 + dm[0, 1] * ul3 + 2.0) / (g.ctrlu + dm[2, 1])
        st.cbc_top[0] = xz_r2c(g, utop)
    if g.ctrlw != 0.0:
        ul2 = xz_c2r(g, st.U[2, 1]); ul3 = xz_c2r(g, st.U[2, 2])
        ubot = (dm[1, 0] * ul2 + dm[2, 0] * ul3) / (g.ctrlw - dm[0, 0])
        st.cbc_bot[2] = xz_r2c(g, ubot)
        ul2 = xz_c2r(g, st.U[2, ny - 1]); ul3 = xz_c2r(g, st.U[2, ny - 2])
        utop = -(dm[1, 1] * ul2 + dm[0, 1] * ul3) / (g.ctrlw + dm[2, 1])
        st.cbc_top[2] = xz_r2c(g, utop)
    return st


# ---------------------------------------------------------------------------
# Initialisation pieces that call the hot-path kernels (PFC/init_vel.f90)
# ---------------------------------------------------------------------------
def remove_initial_divergence(g: Grid, U):
    """remove_initial_divergence (PFC/init_vel.f90:518-536)."""
    ny = g.ny
    out, _ = strip_divergence(g, U, 1.0)
    out[1, 1:ny, 0, 0] = U[1, 0, 0, 0]
    return out


def solve_continuous_poisson(g: Grid, Uvel, NL):
    """solve_continuous_poisson (PFC/init_vel.f90:477-513,
    PFC/arnoldi.f90:569-608): initial pressure from velocity and NL term."""
    ny = g.ny
    cu = Uvel
    cru = -NL + _explicit_diffusion(g, 1.0 / g.re, cu)
    _, cp = strip_divergence(g, cru, 1.0)
    cp[:, g.k2 == 0.0] = 0.0
    return cp


def initial_state(g: Grid, U0, dt=None, p_grad=None, linear_base=None):
    """define_initial_field tail (PFC/init_vel.f90:329-351) and
    compute_initial_pressure (:437-474) for an injected velocity field."""
    U = remove_initial_divergence(g, U0)
    dudy = d_um_dy_of(g, U[0, :, 0, 0])
    R = U.copy()
    if linear_base is None:
        NL, _ = nonlinear(g, U, step=-1, rk_step=3, dt=0.0)
    else:
        NL, _ = nonlinear_linearized(g, U, linear_base, step=-1, rk_step=3, dt=0.0)
    P = solve_continuous_poisson(g, U, NL)
    if dt is None:
        dt = default_dt(g)
    if p_grad is None:
        p_grad = -2.0 / g.re                                        # PFC/setup.f90:1384
    return State(U=U, R=R, P=P, dt=dt, p_grad=p_grad, d_um_dy=dudy)


# ---------------------------------------------------------------------------
# Diagnostics
# ---------------------------------------------------------------------------
def mass_flux(g: Grid, U):
    """Q = int_y Re(u_mean) dyh (PFC/pass2.f90:106-108)."""
    return float(np.sum(U[0, :, 0, 0].real * g.dyh))


def max_divergence(g: Grid, U):
    """max |div| over non-mean modes, j = 2..ny (PFC/pass2.f90:380-401)."""
    d = divergence(g, U)
    return float(np.abs(d).max())


# ---------------------------------------------------------------------------
# Diagnostics: l2_norm (PFC/read_write.f90:1432-1462), per-mode kinetic
# energy (PFC/pass2.f90:113-118), per-plane velocity moments
# (PFC/stats.f90:423-440, one instantaneous field: wgt_old = 0, wgt_new = 1)
# ---------------------------------------------------------------------------
def l2_norm(g: Grid, U):
    """The l2 value l2_norm writes to l2.out.
    u is taken on the dealiased grid, the streamwise component minus the
    laminar profile 1 - y(plane)^2 at the v-point coordinate y(plane), and
    each plane is weighted with dy(plane); l2 = sqrt(sum/(nx2 nz2)/2).
    y has ny entries while the loop reaches plane ny+1 (y(ny+1) is read one
    past the end of the array, PFC/setup.f90:551): the top wall value
    y(ny) is used for that plane here (parity unpinned for that term)."""
    yext = np.concatenate([g.y, g.y[-1:]])
    tot = 0.0
    for j in range(g.ny + 1):
        for c in range(3):
            u = xz_c2r(g, U[c, j][None])[0]
            if c == 0:
                u = u - (1.0 - yext[j] ** 2)
            tot += float(np.sum(u ** 2)) * g.dy[j]
    return math.sqrt(tot / g.nx2 / g.nz2 / 2.0)


def mode_energy(g: Grid, U):
    """KE(x, z) = int_y(|u|^2 + |w|^2 + |v interpolated to yh|^2, dyh)
    (PFC/pass2.f90:113-118) for every stored mode; returns [nx/2+1, nz]."""
    ny = g.ny
    t = np.abs(U[0]) ** 2 + np.abs(U[2]) ** 2
    t[0] += np.abs(U[1, 0]) ** 2
    t[1:ny] += np.abs(0.5 * (U[1, 1:ny] + U[1, 0:ny - 1])) ** 2
    t[ny] += np.abs(U[1, ny - 1]) ** 2
    return np.tensordot(g.dyh, t, axes=(0, 0))


def plane_moments(g: Grid, U):
    """Per plane j on the nx x nz grid (small transform): mean, 2nd, 3rd
    and 4th central moments of u, v, w and the fluxes <u_i' v'>
    (compute_velocity_moments, PFC/stats.f90:423-440).  Returns [15, ny+1]:
    rows 0-2 means, 3-5 second, 6-8 third, 9-11 fourth moments, 12-14 fluxes."""
    out = np.zeros((15, g.ny + 1))
    scale = 1.0 / float(g.nx * g.nz)
    for j in range(g.ny + 1):
        u = small_c2r(g, U[:, j])                       # [3, nz, nx]
        m = np.array([scale * np.sum(u[c]) for c in range(3)])
        d = u - m[:, None, None]
        out[0:3, j] = m
        for c in range(3):
            out[3 + c, j] = scale * np.sum(d[c] ** 2)
            out[6 + c, j] = scale * np.sum(d[c] ** 3)
            out[9 + c, j] = scale * np.sum(d[c] ** 4)
            out[12 + c, j] = scale * np.sum(d[c] * d[1])
    return out


# ---------------------------------------------------------------------------
# Arnoldi matvec (PFC/arnoldi.f90:485-567) and eigenvalue map (:805-806)
# ---------------------------------------------------------------------------
def arnoldi_weights(g: Grid):
    """sqrt(dy)/sqrt(dyh) plane weights used to pack/unpack (PFC/arnoldi.f90:225-234,504-523)."""
    return np.sqrt(g.dy), np.sqrt(g.dyh)


def linearized_dt(g: Grid) -> float:
    """init_baseflow dt = cfl_max*delta1/32 (PFC/pass1_linearized.f90:46)."""
    return g.cfl_max * (g.lx / float(g.nx2)) / 32.0


def linearized_tend(dt: float) -> int:
    """tend = max(int(0.1/dt), 10) (PFC/arnoldi.f90:36)."""
    return max(int(0.1 / dt), 10)


def compute_time_deriv(g: Grid, v, base, dt, tend, eps=1.0e-8):
    """compute_time_deriv (PFC/arnoldi.f90:485-567).  v[c, j, zz, xx] real on
    the nx x nz grid; returns w of the same shape."""
    ny = g.ny
    sdy, sdyh = np.sqrt(g.dy), np.sqrt(g.dyh)
    scale = eps / np.abs(v).max()
    C = small_r2c(g, v)                                             # [3, ny+1, x, z]
    wts = np.empty((3, ny + 1))
    wts[0] = sdyh; wts[1] = sdy; wts[2] = sdyh
    fac = np.where(wts != 0.0, scale / np.where(wts != 0.0, wts, 1.0), 0.0)
    U = C * fac[:, :, None, None]
    st = State(U=U.copy(), R=U.copy(), P=None, dt=dt, p_grad=0.0, d_um_dy=np.zeros(ny))
    NL, _ = nonlinear_linearized(g, U, base, step=-1, rk_step=0, dt=dt)
    st.P = solve_continuous_poisson(g, U, NL)
    st.U = U.copy()
    for step in range(1, tend + 1):
        for rk in (1, 2, 3):
            st.U, st.dt = nonlinear_linearized(g, st.U, base, step=step, rk_step=rk, dt=st.dt)
            rhs_solve(g, st, rk)
        st.t_total += st.dt
    out = st.U * (wts / scale)[:, :, None, None]
    return small_c2r(g, out), st.t_total


def eig_map(lam, t_total):
    """Ritz value -> (growth rate, frequency) (PFC/arnoldi.f90:805-806)."""
    lam = np.asarray(lam, dtype=np.complex128)
    eps = np.finfo(np.float64).eps
    cr = np.log(np.sqrt(lam.real ** 2 + lam.imag ** 2) + eps) / t_total
    ci = np.arctan2(lam.imag, lam.real) / t_total
    return cr, ci
