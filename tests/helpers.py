"""Shared helpers for the tests: package loader, grids and seeded inputs."""
import importlib.util
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "3d-navier-stokes-thesis-paper_amd")


def chanflow():
    """Import the package directory (its name is not a Python identifier)."""
    if "chanflow" in sys.modules:
        return sys.modules["chanflow"]
    spec = importlib.util.spec_from_file_location(
        "chanflow", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["chanflow"] = mod
    spec.loader.exec_module(mod)
    return mod


def synth():
    chanflow()
    import chanflow.synth as s  # noqa: E402
    return s


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


def grid_cfg0(O, **kw):
    """Config 0: Re_c=2500, box 2pi x 2 x pi, 32x65x32, 3/2 dealiasing, stretch 1.5."""
    args = dict(stretch=1.5, fix_mass=1, dt_flag=0)
    args.update(kw)
    return O.make_grid(32, 65, 32, 48, 48, 2.0, 2.0, 2500.0, **args)


def init_cfg0(O, g, seed=0, amp=1.0e-3):
    s = synth()
    return s.laminar_perturbed(g.nx, g.ny, g.nz, g.yh, g.y, seed=seed, amp=amp, kmax=4.0)


def make_channel(g, **kw):
    cf = chanflow()
    ch = cf.Channel(g.nx, g.ny, g.nz, g.nx2, g.nz2, g.lx, g.lz, g.re, beta=g.beta, fix_mass=g.fix_mass,
                    dt_flag=g.dt_flag, cfl_max=g.cfl_max, ctrlu=g.ctrlu, ctrlw=g.ctrlw, **kw)
    ch.upload_grid(g)
    return ch
