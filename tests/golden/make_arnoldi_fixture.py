"""Writes tests/golden/arnoldi_64x97x64.npz: the leading eigenvalues of the
Arnoldi operator of the linearized-stability configuration (64x97x64,
96x96 dealiased, box 2pi x 2 x pi, stretching 2.75, laminar base flow at
Re = 2800, dt = init_baseflow's rule, tend = max(int(0.1/dt), 10)), i.e. of
compute_time_deriv's w = W P_T W^-1 v (PFC/arnoldi.f90:485-567), computed
block by block with the oracle.

The base flow depends on y only, so P_T maps each Fourier block onto itself:
the (kx, kz) coefficients of every field in the step (velocity, RK register,
pressure) evolve without touching any other (kx, kz).  Block (p, q) of the
64x64 grid is therefore reproduced exactly by the (1, +-1) block of a 4x4
grid whose box makes those wavenumbers equal (the block's arrays of kx, kz
and k^2 are set to the 64x64 grid's values bit for bit): the nonlinear pass
only ever multiplies the block by functions of y, which the dealiased
transforms of either grid carry out exactly.

Per block the oracle gives:
  S  one RK3 step (PFC/main_mpi.f90:70-87) as a real-linear map of the
     block's velocity and pressure (data1, data4); data2 equals data1
     after every step (PFC/pass2.f90:326) and nothing else carries over
     (fix_mass = 0, zero wall data, fixed dt);
  I  the initial pressure of compute_time_deriv (nonlinear_linearized at
     rk_step 0, solve_continuous_poisson, PFC/arnoldi.f90:525-531, 569-608);
and the block of P_T is E S^tend I, E taking the velocity back.  The wall
rows with zero Arnoldi weight (u, w at the walls, v at the top) are zeroed
on input and output by W (PFC/arnoldi.f90:504-523) and drop out.  Inputs are
real fields: the kx = 0 blocks carry their Hermitian partner (0, -q), and
block (0, 0) is real.  The eigenvalues of the real representation are those
of the real operator on that pair of blocks (a kx > 0 block stands for itself
and its conjugate (-p, -q)).

Blocks: kx index p = 0..PMAX, kz index q = -QMAX..QMAX (q >= 0 for p = 0).
The x- and z-Nyquist blocks (p = 32, q = 32) decay fastest of all and are
left out.  The GPU test checks every Ritz value ARPACK returns against this
list, so a block missing from it would show as an unmatched eigenvalue.

Run: python tests/golden/make_arnoldi_fixture.py   (about 20 minutes, 8 processes)
"""
import math
import os
import sys
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import chan_oracle as O  # noqa: E402

BIG = dict(nx=64, ny=97, nz=64, nx2=96, nz2=96, lx_over_pi=2.0, aspect=2.0, re=2800.0)
STRETCH = 2.75
PMAX = 6
QMAX = 4
AMP = 1.0e-6      # input amplitude of a basis column (keeps the CFL check quiet)


def big_grid():
    c = BIG
    return O.make_grid(c["nx"], c["ny"], c["nz"], c["nx2"], c["nz2"], c["lx_over_pi"], c["aspect"], c["re"],
                       stretch=STRETCH, fix_mass=0)


def block_grid(G, p, q):
    """4x4 grid whose (ix, iz) block has the wavenumbers of block (p, q) of G."""
    pi = math.atan2(0.0, -1.0)
    kxv = G.kx[p]
    kzv = G.kz[q % G.nz]
    lx = 2.0 * pi / kxv if p > 0 else G.lx
    lz = 2.0 * pi / abs(kzv) if q != 0 else G.lz
    g = O.make_grid(4, G.ny, 4, 6, 6, lx / pi, lx / lz, G.re, stretch=STRETCH, fix_mass=0)
    ix = 1 if p > 0 else 0
    iz = 0 if q == 0 else (1 if q > 0 else 3)
    # the block's wavenumbers bit for bit; the other columns are never excited
    g.kx = g.kx.copy(); g.kz = g.kz.copy()
    g.kx[ix] = kxv
    g.kz[iz] = kzv
    if q != 0:
        g.kz[(4 - iz) % 4] = -kzv
    g.k2 = g.kx[:, None] * g.kx[:, None] + g.kz[None, :] * g.kz[None, :]
    g.k2[ix, iz] = G.k2[p, q % G.nz]
    g.ikx = 1j * g.kx
    g.ikz = 1j * g.kz
    return g, ix, iz


def _place(g, ix, iz, herm, u, pr):
    """Fields holding the block values u (3, ny+1), pr (ny+1); with herm, the
    (0, -q) partner holds the conjugates."""
    U = np.zeros((3, g.ny + 1, g.nxh, g.nz), np.complex128)
    P = np.zeros((g.ny + 1, g.nxh, g.nz), np.complex128)
    U[:, :, ix, iz] = u
    P[:, ix, iz] = pr
    if herm:
        izp = (4 - iz) % 4
        U[:, :, ix, izp] = np.conj(u)
        P[:, ix, izp] = np.conj(pr)
    return U, P


def block_matrix(p, q):
    """Real matrix of the block of P_T (W-zeroed rows dropped) and T."""
    G = big_grid()
    dt = O.linearized_dt(G)
    tend = O.linearized_tend(dt)
    g, ix, iz = block_grid(G, p, q)
    base = O.laminar_baseflow(g)
    nyp = g.ny + 1
    herm = (p == 0 and q != 0)
    real_only = (p == 0 and q == 0)
    ns = 4 * nyp                      # complex state per block: u, v, w, p
    nu = 3 * nyp

    def vec(U, P):
        z = np.concatenate([U[:, :, ix, iz].ravel(), P[:, ix, iz]])
        return z.real if real_only else np.concatenate([z.real, z.imag])

    def unvec(x):
        z = x.astype(np.complex128) if real_only else x[:ns] + 1j * x[ns:]
        return z[:nu].reshape(3, nyp), z[nu:]

    nreal = ns if real_only else 2 * ns

    def step(x):
        u, pr = unvec(x * AMP)
        U, P = _place(g, ix, iz, herm, u, pr)
        st = O.State(U=U.copy(), R=U.copy(), P=P, dt=dt, p_grad=0.0, d_um_dy=np.zeros(g.ny))
        O.rk3_step(g, st, 1, linear_base=base)
        assert st.dt == dt
        return vec(st.U, st.P) / AMP

    S = np.empty((nreal, nreal))
    for i in range(nreal):
        e = np.zeros(nreal); e[i] = 1.0
        S[:, i] = step(e)

    # initial state of compute_time_deriv: velocity u -> (u, P0(u))
    nin = nu if real_only else 2 * nu
    Imat = np.empty((nreal, nin))
    for i in range(nin):
        e = np.zeros(nin); e[i] = AMP
        if real_only:
            u = e.reshape(3, nyp).astype(np.complex128)
        else:
            u = (e[:nu] + 1j * e[nu:]).reshape(3, nyp)
        U, P = _place(g, ix, iz, herm, u, np.zeros(nyp))
        NL, _ = O.nonlinear_linearized(g, U, base, step=-1, rk_step=0, dt=dt)
        P0 = O.solve_continuous_poisson(g, U, NL)
        Imat[:, i] = vec(U, P0) / AMP

    A = Imat
    for _ in range(tend):
        A = S @ A
    # E: velocity part of the state
    if real_only:
        E = A[:nu]
    else:
        E = np.concatenate([A[:nu], A[ns:ns + nu]])
    # Arnoldi weights: zero weight rows/cols (u, w at the walls, v at the top) drop out
    w = np.concatenate([np.sqrt(G.dyh), np.sqrt(G.dy), np.sqrt(G.dyh)])
    keep = w != 0.0
    if not real_only:
        keep = np.concatenate([keep, keep])
    M = E[np.ix_(keep, keep)]
    T = 0.0
    for _ in range(tend):
        T = T + dt
    return M, T


def block_eigs(pq):
    p, q = pq
    M, T = block_matrix(p, q)
    lam = np.linalg.eigvals(M)
    lam = lam[np.argsort(-np.abs(lam), kind="stable")]
    return p, q, lam[:40], T


def main():
    G = big_grid()
    dt = O.linearized_dt(G)
    tend = O.linearized_tend(dt)
    blocks = [(p, q) for p in range(PMAX + 1) for q in range(-QMAX, QMAX + 1) if not (p == 0 and q < 0)]
    lams, ps, qs = [], [], []
    T = None
    with ProcessPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        for p, q, lam, t in ex.map(block_eigs, blocks):
            T = t
            lams.append(lam)
            ps.append(np.full(lam.size, p)); qs.append(np.full(lam.size, q))
            print(f"block ({p},{q}): |lam|max {np.abs(lam[0]):.10f}", flush=True)
    lam = np.concatenate(lams)
    bp = np.concatenate(ps); bq = np.concatenate(qs)
    o = np.argsort(-np.abs(lam), kind="stable")
    lam, bp, bq = lam[o], bp[o], bq[o]
    cr, ci = O.eig_map(lam, T)
    np.savez(os.path.join(HERE, "arnoldi_64x97x64.npz"), lam=lam, cr=cr, ci=ci, p=bp, q=bq, t_total=T,
             dt=dt, tend=tend, pmax=PMAX, qmax=QMAX)
    for k in range(40):
        print(k, bp[k], bq[k], lam[k], cr[k], ci[k])


if __name__ == "__main__":
    main()
