"""Arnoldi eigenvalues over the device matvec (chan_linearized_matvec) match
the oracle's to 1e-8: scipy's ARPACK on the (kx = 1, kz = 0) block of the
linearized propagator at Re = 7500, against tests/golden/os_block_eigs.npz
(the same computation over the oracle's matvec)."""
import importlib.util
import os

import numpy as np
import pytest

from helpers import make_channel
from oracle import chan_oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_orr_sommerfeld_eigenvalues_match_oracle():
    spec = importlib.util.spec_from_file_location("mk_os", os.path.join(HERE, "golden", "make_os_fixture.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    g = mk.grid()
    ch = make_channel(g)
    ch.set_baseflow(O.laminar_baseflow(g))
    dt = O.linearized_dt(g)
    tend = O.linearized_tend(dt)
    ch.set_scalars(dt=dt)
    T = [0.0]

    def matvec(v):
        w, t = ch.linearized_matvec(v, tend)
        T[0] = t
        return w
    lam = mk.leading_eigs(g, matvec)
    ref = np.load(os.path.join(HERE, "golden", "os_block_eigs.npz"))
    assert np.abs(lam - ref["lam"]).max() <= 1e-8 * np.abs(ref["lam"]).max()
    cr, ci = O.eig_map(lam, T[0])
    cr0, ci0 = ref["cr"], ref["ci"]
    assert np.abs(cr - cr0).max() <= 1e-8 and np.abs(ci - ci0).max() <= 1e-8
    ch.close()
