"""Writes tests/golden/os_block_eigs.npz: leading Ritz values of the
linearized RK3 propagator (the Arnoldi matvec compute_time_deriv,
PFC/arnoldi.f90:485-567) about the laminar profile at Re = 7500,
restricted to the (kx = 1, kz = 0) Fourier block, computed with the oracle
and scipy's ARPACK (eigs, which='LM', tol 1e-10, seeded start vector).

The base flow has kx = kz = 0 only, so the propagator does not couple
Fourier blocks and the restriction is exact; it isolates the 2-D
Orr-Sommerfeld mode whose phase speed the reference hard-codes
(c_r = 0.24989154, PFC/setup.f90:1387).  Run: python tests/golden/make_os_fixture.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import chan_oracle as O  # noqa: E402

GRID = dict(nx=4, ny=97, nz=4, nx2=6, nz2=6, lx_over_pi=2.0, aspect=0.5, re=7500.0)
BLOCK = (1, 0)
K = 4


def grid():
    c = GRID
    return O.make_grid(c["nx"], c["ny"], c["nz"], c["nx2"], c["nz2"], c["lx_over_pi"], c["aspect"], c["re"],
                       stretch=2.75, fix_mass=0)


def block_operator(g, matvec):
    """Real operator on the (re, im) coefficients of the Fourier block BLOCK,
    built from a physical-space matvec of shape (3, ny+1, nz, nx)."""
    ix, iz = BLOCK
    nyp = g.ny + 1
    shp = (3, nyp, g.nz, g.nx)

    def s2p(c):
        C = np.zeros((3 * nyp, g.nxh, g.nz), complex)
        C[:, ix, iz] = c
        return O.small_c2r(g, C).reshape(shp)

    def p2s(f):
        return O.small_r2c(g, f.reshape(3 * nyp, g.nz, g.nx))[:, ix, iz]

    m = 2 * 3 * nyp

    def mv(x):
        w = p2s(matvec(s2p(x[:m // 2] + 1j * x[m // 2:])))
        return np.concatenate([w.real, w.imag])
    return m, mv


def leading_eigs(g, matvec, k=K):
    from scipy.sparse.linalg import LinearOperator, eigs
    m, mv = block_operator(g, matvec)
    lam = eigs(LinearOperator((m, m), matvec=mv, dtype=np.float64), k=k, which="LM", tol=1e-10,
               v0=np.random.default_rng(0).standard_normal(m), return_eigenvectors=False)
    return lam[np.argsort(-np.abs(lam), kind="stable")]


def main():
    g = grid()
    base = O.laminar_baseflow(g)
    dt = O.linearized_dt(g)
    tend = O.linearized_tend(dt)
    T = [0.0]

    def matvec(v):
        w, t = O.compute_time_deriv(g, v, base, dt, tend)
        T[0] = t
        return w
    lam = leading_eigs(g, matvec)
    cr, ci = O.eig_map(lam, T[0])
    np.savez(os.path.join(HERE, "os_block_eigs.npz"), lam=lam, t_total=T[0], dt=dt, tend=tend, cr=cr, ci=ci)
    print(lam, cr, ci)


if __name__ == "__main__":
    main()
