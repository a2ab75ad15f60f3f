"""chanflow: MI355X-native pseudo-spectral channel-flow time step.

Python host mirror of the C-ABI in ``include/chanflow.h`` (ctypes).  The
numerical work happens in ``libchanflow.so`` (hand-written HIP kernels for
gfx950); this module only moves arrays and calls entry points.  Importing it
fails loudly when the library has not been built: there is no CPU fallback.

Host arrays use the reference's Fortran layouts, exposed as C-ordered numpy
arrays with the axes reversed:

* ``data1`` complex(nz, nx/2+1, ny+1, 3)  <->  ndarray [3, ny+1, nx/2+1, nz]
* ``data2`` complex(ny+1, 3, nz, nx/2+1)  <->  ndarray [nx/2+1, nz, 3, ny+1]
* ``data4`` complex(ny+1, nz, nx/2+1)     <->  ndarray [nx/2+1, nz, ny+1]
* physical planes real(nx2, nz2, np)      <->  ndarray [np, nz2, nx2]
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# CHANFLOW_LIB names another build of the library (e.g. one compiled with a
# different instruction scheduler) to load instead of the in-tree one.
LIBPATH = os.environ.get("CHANFLOW_LIB") or os.path.join(_HERE, "libchanflow.so")

_dp = C.POINTER(C.c_double)


class ChanParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nx2", C.c_int), ("nz2", C.c_int),
                ("lx", C.c_double), ("lz", C.c_double), ("re", C.c_double), ("beta", C.c_double),
                ("fix_mass", C.c_int), ("dt_flag", C.c_int), ("cfl_max", C.c_double),
                ("ctrlu", C.c_double), ("ctrlw", C.c_double), ("bl_grid", C.c_int)]


class ChanOperators(C.Structure):
    _fields_ = [(n, _dp) for n in ("d2h", "d2n", "dm", "dh", "dn", "hn", "iu", "dyh", "dy", "kx", "kz",
                                   "buffer", "target_inflow")]


ENTRY_POINTS = (
    "chan_create", "chan_destroy", "chan_last_error", "chan_version", "chan_upload_operators",
    "chan_build_operators", "chan_set_state", "chan_get_state", "chan_get_planes", "chan_set_scalars",
    "chan_get_scalars", "chan_set_wall_bc", "chan_nonlinear", "chan_nonlinear_linearized",
    "chan_set_baseflow", "chan_rhs_solve", "chan_step", "chan_boundary", "chan_initialize_flow",
    "chan_remove_divergence", "chan_compute_initial_pressure", "chan_xz_fft_complex_to_real",
    "chan_xz_fft_real_to_complex", "chan_compute_divergence_of_fluxes", "chan_small_fft_complex_to_real",
    "chan_small_fft_real_to_complex", "chan_solve_helmholtz", "chan_solve_p", "chan_strip_divergence",
    "chan_linearized_matvec", "chan_divergence_max", "chan_mass_flux", "chan_set_timing",
    "chan_get_kernel_times", "chan_sync", "chan_io_last_error", "chan_savevel_write",
    "chan_savevel_read_header", "chan_savevel_read", "chan_save_velocity", "chan_read_velocity",
    "chan_set_coordinates", "chan_l2_norm", "chan_mode_energy", "chan_plane_moments",
    "chan_set_tuning", "chan_get_tuning", "chan_autotune", "chan_ensemble_create", "chan_ensemble_step",
    "chan_ensemble_invalidate", "chan_ensemble_autotune", "chan_ensemble_destroy", "chan_ensemble_last_error",
    "chan_mlp_create", "chan_mlp_destroy", "chan_mlp_last_error", "chan_mlp_num_params", "chan_mlp_set_params",
    "chan_mlp_get_params", "chan_mlp_infer", "chan_mlp_forward_fp32", "chan_mlp_gradient", "chan_mlp_train",
    "chan_vtk_write", "chan_write_vtk",
    "chan_les_configure", "chan_les_update", "chan_les_set_tau", "chan_les_get_tau", "chan_les_get_coefficient",
    "chan_les_get_fields", "chan_les_features", "chan_les_gated_points", "chan_jet_configure",
    "chan_strip_divergence_modes", "chan_subtract_pressure_gradient_modes",
    "chan_compute_divergence_of_fluxes_noalloc", "chan_les_filter_dns",
)

# bytes of the fixed per-GPU scratch array of chan_compute_divergence_of_fluxes_noalloc
CHAN_NOALLOC_SCRATCH_BYTES = 8 << 20

_lib = None
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


def load_library(path: str = LIBPATH):
    """Load libchanflow.so (raises OSError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"libchanflow.so not built at {path}; run __graft_entry__.build()")
    lib = C.CDLL(path)
    vp = C.c_void_p
    sig = {
        "chan_create": (C.c_int, [C.POINTER(ChanParams), C.POINTER(vp)]),
        "chan_destroy": (C.c_int, [vp]),
        "chan_last_error": (C.c_char_p, [vp]),
        "chan_version": (C.c_char_p, []),
        "chan_upload_operators": (C.c_int, [vp, C.POINTER(ChanOperators)]),
        "chan_build_operators": (C.c_int, [vp, C.c_double, _dp, _dp]),
        "chan_set_state": (C.c_int, [vp, _dp, _dp, _dp]),
        "chan_get_state": (C.c_int, [vp, _dp, _dp, _dp]),
        "chan_get_planes": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, _dp]),
        "chan_set_scalars": (C.c_int, [vp, _dp, _dp, _dp, _dp]),
        "chan_get_scalars": (C.c_int, [vp, _dp, _dp, _dp, _dp]),
        "chan_set_wall_bc": (C.c_int, [vp, _dp, _dp]),
        "chan_nonlinear": (C.c_int, [vp, C.c_int, C.c_int, _dp, _dp]),
        "chan_nonlinear_linearized": (C.c_int, [vp, C.c_int, C.c_int, _dp, _dp]),
        "chan_set_baseflow": (C.c_int, [vp, _dp]),
        "chan_rhs_solve": (C.c_int, [vp, C.c_int, C.c_int, _dp]),
        "chan_step": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "chan_boundary": (C.c_int, [vp]),
        "chan_initialize_flow": (C.c_int, [vp, C.c_int]),
        "chan_remove_divergence": (C.c_int, [vp]),
        "chan_compute_initial_pressure": (C.c_int, [vp, C.c_int]),
        "chan_xz_fft_complex_to_real": (C.c_int, [vp, _dp, _dp, C.c_int]),
        "chan_xz_fft_real_to_complex": (C.c_int, [vp, _dp, _dp, C.c_int]),
        "chan_compute_divergence_of_fluxes": (C.c_int, [vp, _dp, _dp, _dp, _dp, C.c_int]),
        "chan_compute_divergence_of_fluxes_noalloc": (C.c_int, [vp, vp, vp, vp, vp, C.c_int, vp]),
        "chan_les_filter_dns": (C.c_int, [vp, vp, _dp]),
        "chan_small_fft_complex_to_real": (C.c_int, [vp, _dp, _dp, C.c_int]),
        "chan_small_fft_real_to_complex": (C.c_int, [vp, _dp, _dp, C.c_int]),
        "chan_solve_helmholtz": (C.c_int, [vp, C.c_int, C.c_int, _dp, _dp, C.c_double, _dp, _dp, _dp,
                                           C.c_int, C.c_int]),
        "chan_solve_p": (C.c_int, [vp, C.c_int, _dp, _dp]),
        "chan_strip_divergence": (C.c_int, [vp, _dp, _dp, C.c_double]),
        "chan_strip_divergence_modes": (C.c_int, [vp, C.c_int, _dp, _dp, C.c_double, _ip, _ip]),
        "chan_subtract_pressure_gradient_modes": (C.c_int, [vp, C.c_int, _dp, _dp, C.c_double, _ip, _ip]),
        "chan_linearized_matvec": (C.c_int, [vp, _dp, _dp, C.c_int, _dp]),
        "chan_divergence_max": (C.c_int, [vp, _dp]),
        "chan_mass_flux": (C.c_int, [vp, _dp]),
        "chan_set_timing": (C.c_int, [vp, C.c_int]),
        "chan_get_kernel_times": (C.c_int, [vp, _dp, C.POINTER(C.c_long), C.c_int]),
        "chan_sync": (C.c_int, [vp]),
        "chan_io_last_error": (C.c_char_p, []),
        "chan_savevel_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                         C.c_double, C.c_double, C.c_double, _dp]),
        "chan_savevel_read_header": (C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                               C.POINTER(C.c_int), _dp, _dp, _dp, _dp, _dp]),
        "chan_savevel_read": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, _dp]),
        "chan_save_velocity": (C.c_int, [vp, C.c_char_p]),
        "chan_read_velocity": (C.c_int, [vp, C.c_char_p]),
        "chan_set_coordinates": (C.c_int, [vp, _dp]),
        "chan_l2_norm": (C.c_int, [vp, _dp, _dp]),
        "chan_mode_energy": (C.c_int, [vp, _dp]),
        "chan_plane_moments": (C.c_int, [vp, _dp]),
        "chan_set_tuning": (C.c_int, [vp, C.c_int, C.c_int]),
        "chan_get_tuning": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int)]),
        "chan_autotune": (C.c_int, [vp, C.c_int, _dp]),
        "chan_ensemble_create": (C.c_int, [C.c_int, C.POINTER(vp), C.c_int, C.c_int, C.POINTER(vp)]),
        "chan_ensemble_step": (C.c_int, [vp, C.c_int]),
        "chan_ensemble_invalidate": (C.c_int, [vp]),
        "chan_ensemble_autotune": (C.c_int, [vp, C.c_int, C.POINTER(C.c_double)]),
        "chan_ensemble_destroy": (C.c_int, [vp]),
        "chan_ensemble_last_error": (C.c_char_p, [vp]),
        "chan_mlp_create": (C.c_int, [C.POINTER(vp)]),
        "chan_mlp_destroy": (C.c_int, [vp]),
        "chan_mlp_last_error": (C.c_char_p, [vp]),
        "chan_mlp_num_params": (C.c_int, []),
        "chan_mlp_set_params": (C.c_int, [vp, _fp]),
        "chan_mlp_get_params": (C.c_int, [vp, _fp]),
        "chan_mlp_infer": (C.c_int, [vp, C.c_int, _fp, _fp]),
        "chan_mlp_forward_fp32": (C.c_int, [vp, C.c_int, _fp, _fp]),
        "chan_mlp_gradient": (C.c_int, [vp, C.c_int, _fp, _fp, _fp, _dp]),
        "chan_mlp_train": (C.c_int, [vp, C.c_int, _fp, _fp, C.c_int, C.c_int, C.c_double, _dp]),
        "chan_vtk_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp]),
        "chan_write_vtk": (C.c_int, [vp, C.c_char_p, _dp]),
        "chan_les_configure": (C.c_int, [vp, C.c_int, vp, C.c_double, C.c_int]),
        "chan_les_update": (C.c_int, [vp]),
        "chan_les_set_tau": (C.c_int, [vp, _dp]),
        "chan_les_get_tau": (C.c_int, [vp, _dp]),
        "chan_les_get_coefficient": (C.c_int, [vp, _dp]),
        "chan_les_get_fields": (C.c_int, [vp, _dp]),
        "chan_les_features": (C.c_int, [vp, _fp, _dp]),
        "chan_les_gated_points": (C.c_int, [vp, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
        "chan_jet_configure": (C.c_int, [vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                         C.c_double]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


class ChanError(RuntimeError):
    pass


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


KERNEL_KINDS = ("z-inverse", "x-pass", "z-forward", "solve", "other")


class Channel:
    """One channel instance resident on the GPU (wraps a chan_ctx)."""

    def __init__(self, nx, ny, nz, nx2, nz2, lx, lz, re, beta=0.0, fix_mass=1, dt_flag=0, cfl_max=0.95,
                 ctrlu=0.0, ctrlw=0.0):
        self.lib = load_library()
        self.nx, self.ny, self.nz, self.nx2, self.nz2 = nx, ny, nz, nx2, nz2
        self.nxh = nx // 2 + 1
        self.nyp = ny + 1
        p = ChanParams(nx, ny, nz, nx2, nz2, lx, lz, re, beta, fix_mass, dt_flag, cfl_max, ctrlu, ctrlw, 0)
        h = C.c_void_p()
        if self.lib.chan_create(C.byref(p), C.byref(h)) != 0:
            raise ChanError(self.lib.chan_last_error(None).decode())
        self.h = h
        self._keep = []

    # -- lifetime
    def close(self):
        if getattr(self, "h", None):
            self.lib.chan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise ChanError(self.lib.chan_last_error(self.h).decode())

    # -- operators
    def upload_operators(self, d2h, d2n, dm, dh, dn, hn, iu, dyh, dy, kx, kz, buffer=None, target_inflow=None):
        """Arrays in numpy form [i, j] of the Fortran (i, j) arrays: d2h (3, ny+1),
        d2n (3, ny), dm (3, 4), iu (2, ny), target_inflow (ny+1, 3); buffer [nz2, nx2]."""
        arrs = dict(d2h=_f64(np.asarray(d2h).T), d2n=_f64(np.asarray(d2n).T), dm=_f64(np.asarray(dm).T),
                    dh=_f64(dh), dn=_f64(dn), hn=_f64(hn), iu=_f64(np.asarray(iu).T), dyh=_f64(dyh), dy=_f64(dy),
                    kx=_f64(kx), kz=_f64(kz),
                    buffer=None if buffer is None else _f64(buffer),
                    target_inflow=None if target_inflow is None else _f64(np.asarray(target_inflow).T))
        o = ChanOperators(**{k: _ptr(v) for k, v in arrs.items()})
        self._chk(self.lib.chan_upload_operators(self.h, C.byref(o)))

    def upload_grid(self, g):
        """Upload the operators of an oracle-style grid object (same attribute names)."""
        self.upload_operators(g.d2h, g.d2n, g.dm, g.dh, g.dn, g.hn, g.iu, g.dyh, g.dy, g.kx, g.kz,
                              getattr(g, "buffer", None), getattr(g, "target_inflow", None))
        if getattr(g, "y", None) is not None:
            self.set_coordinates(g.y)

    def build_operators(self, stretch=2.75):
        y = np.empty(self.ny)
        yh = np.empty(self.ny + 1)
        self._chk(self.lib.chan_build_operators(self.h, stretch, _ptr(y), _ptr(yh)))
        return y, yh

    # -- state
    def set_state(self, data1=None, data2=None, data4=None):
        a1 = None if data1 is None else _c128(data1)
        a2 = None if data2 is None else _c128(data2)
        a4 = None if data4 is None else _c128(data4)
        self._chk(self.lib.chan_set_state(self.h, _ptr(a1), _ptr(a2), _ptr(a4)))

    def get_state(self):
        d1 = np.empty((3, self.nyp, self.nxh, self.nz), np.complex128)
        d2 = np.empty((self.nxh, self.nz, 3, self.nyp), np.complex128)
        d4 = np.empty((self.nxh, self.nz, self.nyp), np.complex128)
        self._chk(self.lib.chan_get_state(self.h, _ptr(d1), _ptr(d2), _ptr(d4)))
        return d1, d2, d4

    # -- launch-shape tuning
    TUNING_KEYS = ("xpass_wave", "xpass_threads", "zfwd_v2", "solve_lanes", "solve_bl", "zinv_kx", "zfwd_kx", "xpass_prefetch", "solve_dma", "xpass_chunk", "nl_chunks", "solve_seg", "xpass_rows")

    def set_tuning(self, key, value):
        k = self.TUNING_KEYS.index(key) if isinstance(key, str) else int(key)
        self._chk(self.lib.chan_set_tuning(self.h, k, int(value)))

    # -- wall jet
    def jet_configure(self, wall, amplitude, radius, xjet, zjet, forcing=0.0, frequency=0.0):
        """Gaussian jet at the bottom (wall=0) or top (wall=1) wall; amplitude 0 removes it."""
        self._chk(self.lib.chan_jet_configure(self.h, int(wall), float(amplitude), float(radius), float(xjet),
                                              float(zjet), float(forcing), float(frequency)))

    # -- subgrid closure (LES)
    LES_MODES = {"off": 0, "dynamic": 1, "learned": 2, "prescribed": 3}

    def les_configure(self, mode, net=None, beta=1.0, every_substep=True):
        """mode: 'off', 'dynamic' (dynamic Smagorinsky), 'learned' (Mlp `net`
        with the Smagorinsky fallback gate) or 'prescribed' (les_set_tau);
        every_substep=False evaluates the stress once per step."""
        m = self.LES_MODES[mode] if isinstance(mode, str) else int(mode)
        self._les_net = net
        self._chk(self.lib.chan_les_configure(self.h, m, net.h if net is not None else None, float(beta),
                                              1 if every_substep else 0))

    def les_filter_from(self, dns, tau=True):
        """Set this (coarse) channel's velocity to the filtered velocity of
        the fine channel `dns`; with tau, also return the exact subgrid stress
        [6][ny+1][nz2][nx2] of that filter (see chan_les_filter_dns)."""
        t = np.empty(self._phys_shape(6)) if tau else None
        self._chk(self.lib.chan_les_filter_dns(self.h, dns.h, _ptr(t)))
        return t

    def les_update(self):
        self._chk(self.lib.chan_les_update(self.h))

    def _phys_shape(self, nf):
        return (nf, self.nyp, self.nz2, self.nx2)

    def les_set_tau(self, tau):
        t = _f64(tau)
        if t.shape != self._phys_shape(6):
            raise ValueError(f"tau must have shape {self._phys_shape(6)}")
        self._chk(self.lib.chan_les_set_tau(self.h, _ptr(t)))

    def les_tau(self):
        t = np.empty(self._phys_shape(6))
        self._chk(self.lib.chan_les_get_tau(self.h, _ptr(t)))
        return t

    def les_coefficient(self):
        c = np.empty(self.nyp)
        self._chk(self.lib.chan_les_get_coefficient(self.h, _ptr(c)))
        return c

    def les_fields(self):
        f = np.empty(self._phys_shape(12))
        self._chk(self.lib.chan_les_get_fields(self.h, _ptr(f)))
        return f

    def les_features(self):
        n = (self.ny - 1) * self.nz2 * self.nx2
        x = np.empty((n, 10), np.float32)
        sc = np.empty(n)
        self._chk(self.lib.chan_les_features(self.h, x.ctypes.data_as(_fp), _ptr(sc)))
        return x, sc

    def les_gated_points(self):
        a, b = C.c_long(), C.c_long()
        self._chk(self.lib.chan_les_gated_points(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def get_tuning(self):
        out = {}
        for k, name in enumerate(self.TUNING_KEYS):
            v = C.c_int()
            self._chk(self.lib.chan_get_tuning(self.h, k, C.byref(v)))
            out[name] = v.value
        return out

    def autotune(self, steps=2):
        """Times every candidate launch shape over `steps` timesteps, keeps the
        fastest (state restored); returns ms per step of the chosen shape."""
        ms = C.c_double()
        self._chk(self.lib.chan_autotune(self.h, steps, C.byref(ms)))
        return ms.value

    # -- diagnostics
    def set_coordinates(self, y):
        self._chk(self.lib.chan_set_coordinates(self.h, _ptr(_f64(y))))

    def l2_norm(self):
        l2, drag = C.c_double(), C.c_double()
        self._chk(self.lib.chan_l2_norm(self.h, C.byref(l2), C.byref(drag)))
        return l2.value, drag.value

    def mode_energy(self):
        """Per-mode kinetic energy as [nx/2+1, nz]."""
        ke = np.empty((self.nz, self.nxh), np.float64)
        self._chk(self.lib.chan_mode_energy(self.h, _ptr(ke)))
        return ke.T.copy()

    def plane_moments(self):
        """[15, ny+1]: means, 2nd, 3rd, 4th central moments of u, v, w, fluxes <u_i' v'>."""
        out = np.empty((15, self.nyp), np.float64)
        self._chk(self.lib.chan_plane_moments(self.h, _ptr(out)))
        return out

    def write_vtk(self, path, yh):
        self._chk(self.lib.chan_write_vtk(self.h, os.fsencode(path), _ptr(_f64(yh))))

    # -- checkpoints (savevel.out / newvel.out)
    def save_velocity(self, path):
        self._chk(self.lib.chan_save_velocity(self.h, os.fsencode(path)))

    def read_velocity(self, path):
        self._chk(self.lib.chan_read_velocity(self.h, os.fsencode(path)))

    def get_data1(self):
        d1 = np.empty((3, self.nyp, self.nxh, self.nz), np.complex128)
        self._chk(self.lib.chan_get_state(self.h, _ptr(d1), None, None))
        return d1

    def set_scalars(self, dt=None, t_total=None, p_grad=None, d_um_dy=None):
        def box(v):
            return None if v is None else C.byref(C.c_double(v))
        dud = None if d_um_dy is None else _f64(d_um_dy)
        args = [box(dt), box(t_total), box(p_grad)]
        self._chk(self.lib.chan_set_scalars(self.h, *[a if a is not None else None for a in args], _ptr(dud)))

    def get_scalars(self):
        dt, tt, pg = C.c_double(), C.c_double(), C.c_double()
        dud = np.empty(self.ny)
        self._chk(self.lib.chan_get_scalars(self.h, C.byref(dt), C.byref(tt), C.byref(pg), _ptr(dud)))
        return dict(dt=dt.value, t_total=tt.value, p_grad=pg.value, d_um_dy=dud)

    def set_wall_bc(self, cbc_top=None, cbc_bot=None):
        t = None if cbc_top is None else _c128(cbc_top)
        b = None if cbc_bot is None else _c128(cbc_bot)
        self._chk(self.lib.chan_set_wall_bc(self.h, _ptr(t), _ptr(b)))

    def set_baseflow(self, base):
        """base: physical base flow [3, ny+1, nz2, nx2] (Fortran baseflow(nx2, nz2, ny+1, 3))."""
        self._chk(self.lib.chan_set_baseflow(self.h, _ptr(_f64(base))))

    # -- step phases
    def nonlinear(self, step, rk_step, dt=None, linearized=False):
        d = C.c_double(0.0 if dt is None else dt)
        cfl = C.c_double(0.0)
        f = self.lib.chan_nonlinear_linearized if linearized else self.lib.chan_nonlinear
        self._chk(f(self.h, step, rk_step, C.byref(d) if dt is not None else None, C.byref(cfl)))
        return (d.value if dt is not None else None), cfl.value

    def rhs_solve(self, step, rk_step):
        pg = C.c_double(0.0)
        self._chk(self.lib.chan_rhs_solve(self.h, step, rk_step, C.byref(pg)))
        return pg.value

    def step(self, nsteps=1, step0=1, linearized=False):
        self._chk(self.lib.chan_step(self.h, step0, nsteps, 1 if linearized else 0))

    def boundary(self):
        self._chk(self.lib.chan_boundary(self.h))

    def initialize_flow(self, linearized=False):
        self._chk(self.lib.chan_initialize_flow(self.h, 1 if linearized else 0))

    def remove_divergence(self):
        self._chk(self.lib.chan_remove_divergence(self.h))

    def compute_initial_pressure(self, linearized=False):
        self._chk(self.lib.chan_compute_initial_pressure(self.h, 1 if linearized else 0))

    # -- transforms (host arrays)
    def xz_c2r(self, cdata):
        cdata = _c128(cdata)
        npl = cdata.size // (self.nxh * self.nz)
        out = np.empty((npl, self.nz2, self.nx2))
        self._chk(self.lib.chan_xz_fft_complex_to_real(self.h, _ptr(cdata), _ptr(out), npl))
        return out

    def xz_r2c(self, data):
        data = _f64(data)
        npl = data.size // (self.nx2 * self.nz2)
        out = np.empty((npl, self.nxh, self.nz), np.complex128)
        self._chk(self.lib.chan_xz_fft_real_to_complex(self.h, _ptr(data), _ptr(out), npl))
        return out

    def div_fluxes_noalloc(self, fx_dev, fy_dev, fz_dev, out_dev, nplanes, stream=None):
        """compute_divergence_of_fluxes on device arrays (integer device
        addresses, e.g. a torch tensor's data_ptr()) with no allocation: fx,
        fy, fz real [nplanes][nz2][nx2], out complex [nplanes][nx/2+1][nz];
        queued on `stream` (a hipStream_t address; None: the context's
        stream) without synchronising."""
        self._chk(self.lib.chan_compute_divergence_of_fluxes_noalloc(self.h, fx_dev, fy_dev, fz_dev, out_dev,
                                                                     int(nplanes), stream))

    def div_fluxes(self, fx, fy, fz):
        fx, fy, fz = _f64(fx), _f64(fy), _f64(fz)
        npl = fx.size // (self.nx2 * self.nz2)
        out = np.empty((npl, self.nxh, self.nz), np.complex128)
        self._chk(self.lib.chan_compute_divergence_of_fluxes(self.h, _ptr(fx), _ptr(fy), _ptr(fz), _ptr(out), npl))
        return out

    def small_c2r(self, cdata):
        cdata = _c128(cdata)
        npl = cdata.size // (self.nxh * self.nz)
        out = np.empty((npl, self.nz, self.nx))
        self._chk(self.lib.chan_small_fft_complex_to_real(self.h, _ptr(cdata), _ptr(out), npl))
        return out

    def small_r2c(self, data):
        data = _f64(data)
        npl = data.size // (self.nx * self.nz)
        out = np.empty((npl, self.nxh, self.nz), np.complex128)
        self._chk(self.lib.chan_small_fft_real_to_complex(self.h, _ptr(data), _ptr(out), npl))
        return out

    # -- solvers (host arrays)
    def solve_helmholtz(self, f, alpha, diff, d2, topbc, botbc):
        """f [nsys, n] complex (Fortran f(n, nsys)); alpha [nsys]; d2 [3, n] numpy form."""
        f = _c128(f).copy()
        nsys, n = f.shape
        self._chk(self.lib.chan_solve_helmholtz(self.h, n, nsys, _ptr(f), _ptr(_f64(alpha)), diff,
                                                _ptr(_f64(np.asarray(d2).T)), _ptr(_c128(topbc)),
                                                _ptr(_c128(botbc)), 0, 0))
        return f

    def solve_p(self, p, alpha):
        p = _c128(p).copy()
        nsys = p.shape[0]
        self._chk(self.lib.chan_solve_p(self.h, nsys, _ptr(p), _ptr(_f64(alpha))))
        return p

    def strip_divergence(self, cu, coef):
        cu = _c128(cu).copy()
        cp = np.empty((self.nxh, self.nz, self.nyp), np.complex128)
        self._chk(self.lib.chan_strip_divergence(self.h, _ptr(cu), _ptr(cp), coef))
        return cu, cp

    def strip_divergence_modes(self, cu, coef, p, q):
        """strip_divergence per (kx,kz) column (PFC/solve.f90:14-50): cu
        complex [nsys][3][ny+1] (Fortran cu(ny+1, 3, nsys)), p, q the
        reference's 1-based x- and z-mode indices.  Returns (cu, cp)."""
        cu = _c128(cu).copy()
        nsys = cu.shape[0]
        cp = np.empty((nsys, self.nyp), np.complex128)
        pi = np.ascontiguousarray(p, dtype=np.int32)
        qi = np.ascontiguousarray(q, dtype=np.int32)
        self._chk(self.lib.chan_strip_divergence_modes(self.h, nsys, _ptr(cu), _ptr(cp), coef,
                                                       pi.ctypes.data_as(_ip), qi.ctypes.data_as(_ip)))
        return cu, cp

    def subtract_pressure_gradient_modes(self, crhs, cpress, coef, p, q):
        """subtract_pressure_gradient per column (PFC/solve.f90:54-77)."""
        crhs = _c128(crhs).copy()
        cpress = _c128(cpress)
        pi = np.ascontiguousarray(p, dtype=np.int32)
        qi = np.ascontiguousarray(q, dtype=np.int32)
        self._chk(self.lib.chan_subtract_pressure_gradient_modes(self.h, crhs.shape[0], _ptr(crhs), _ptr(cpress),
                                                                 coef, pi.ctypes.data_as(_ip),
                                                                 qi.ctypes.data_as(_ip)))
        return crhs

    def linearized_matvec(self, v, tend):
        v = _f64(v)
        w = np.empty_like(v)
        t = C.c_double()
        self._chk(self.lib.chan_linearized_matvec(self.h, _ptr(v), _ptr(w), tend, C.byref(t)))
        return w, t.value

    # -- diagnostics
    def divergence_max(self):
        d = C.c_double()
        self._chk(self.lib.chan_divergence_max(self.h, C.byref(d)))
        return d.value

    def mass_flux(self):
        d = C.c_double()
        self._chk(self.lib.chan_mass_flux(self.h, C.byref(d)))
        return d.value

    def set_timing(self, on=True):
        self._chk(self.lib.chan_set_timing(self.h, 1 if on else 0))

    def kernel_times(self):
        ms = (C.c_double * 5)()
        cnt = (C.c_long * 5)()
        self._chk(self.lib.chan_get_kernel_times(self.h, ms, cnt, 5))
        return {k: (ms[i], cnt[i]) for i, k in enumerate(KERNEL_KINDS)}

    def sync(self):
        self._chk(self.lib.chan_sync(self.h))


class Ensemble:
    """Independent Channel instances advanced together by one captured graph
    per timestep.  nstreams = 0 (the default): batched launches, every phase
    of a substage one launch per transform length over all members (members
    with the default launch shapes, no LES closure, jet or feedback gains);
    nstreams >= 1: each member's own kernels, members spread over that many
    concurrent streams."""

    def __init__(self, members, nstreams=0, linearized=False):
        self.lib = load_library()
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.h for m in self.members])
        h = C.c_void_p()
        if self.lib.chan_ensemble_create(len(self.members), arr, nstreams, int(linearized), C.byref(h)) != 0:
            raise ChanError(self.lib.chan_ensemble_last_error(None).decode())
        self.h = h

    def step(self, nsteps=1):
        if self.lib.chan_ensemble_step(self.h, nsteps) != 0:
            raise ChanError(self.lib.chan_ensemble_last_error(self.h).decode())

    def invalidate(self):
        self.lib.chan_ensemble_invalidate(self.h)

    def autotune(self, steps=3):
        """Batched ensembles: time each x-pass plane-range plan, then the tile
        shapes (z-rows per x-pass wavefront, modes per solve wavefront) over
        `steps` replayed steps each (every member's state restored after
        each), keep the fastest; returns its ms per step."""
        ms = C.c_double(0.0)
        if self.lib.chan_ensemble_autotune(self.h, steps, C.byref(ms)) != 0:
            raise ChanError(self.lib.chan_ensemble_last_error(self.h).decode())
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.chan_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _fptr(a):
    return a.ctypes.data_as(_fp)


class Mlp:
    """The learned closure network 10 -> 128 -> 128 -> 128 -> 6 (GELU) on the GPU."""

    def __init__(self, params=None):
        self.lib = load_library()
        h = C.c_void_p()
        if self.lib.chan_mlp_create(C.byref(h)) != 0:
            raise ChanError("chan_mlp_create failed")
        self.h = h
        self.nparams = self.lib.chan_mlp_num_params()
        if params is not None:
            self.set_params(params)

    def _chk(self, rc):
        if rc != 0:
            raise ChanError(self.lib.chan_mlp_last_error(self.h).decode())

    @staticmethod
    def flatten(W, b):
        return np.concatenate([np.concatenate([np.asarray(w).ravel(), np.asarray(bb).ravel()])
                               for w, bb in zip(W, b)]).astype(np.float32)

    def set_params(self, p):
        p = _f32(p)
        if p.size != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters")
        self._chk(self.lib.chan_mlp_set_params(self.h, _fptr(p)))

    def get_params(self):
        p = np.empty(self.nparams, np.float32)
        self._chk(self.lib.chan_mlp_get_params(self.h, _fptr(p)))
        return p

    def infer(self, x):
        x = _f32(x)
        y = np.empty((x.shape[0], 6), np.float32)
        self._chk(self.lib.chan_mlp_infer(self.h, x.shape[0], _fptr(x), _fptr(y)))
        return y

    def forward(self, x):
        x = _f32(x)
        y = np.empty((x.shape[0], 6), np.float32)
        self._chk(self.lib.chan_mlp_forward_fp32(self.h, x.shape[0], _fptr(x), _fptr(y)))
        return y

    def gradient(self, x, t):
        x, t = _f32(x), _f32(t)
        g = np.empty(self.nparams, np.float32)
        loss = C.c_double()
        self._chk(self.lib.chan_mlp_gradient(self.h, x.shape[0], _fptr(x), _fptr(t), _fptr(g), C.byref(loss)))
        return g, loss.value

    def train(self, x, t, epochs=1, batch=256, lr=1e-3):
        x, t = _f32(x), _f32(t)
        loss = C.c_double()
        self._chk(self.lib.chan_mlp_train(self.h, x.shape[0], _fptr(x), _fptr(t), epochs, batch, lr,
                                          C.byref(loss)))
        return loss.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.chan_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def savevel_write(path, data1, lx, lz, re, dt, t_total):
    """Write data1 complex [3][ny+1][nx/2+1][nz] (the Fortran data1(nz, nx/2+1, ny+1, 3))
    in the savevel.out format; nx is taken as 2*(nx/2+1 - 1)."""
    lib = load_library()
    d1 = _c128(data1)
    nc, nyp, nxh, nz = d1.shape
    if nc != 3:
        raise ValueError("data1 must have shape (3, ny+1, nx/2+1, nz)")
    rc = lib.chan_savevel_write(os.fsencode(path), 2 * (nxh - 1), nyp - 1, nz, lx, lz, re, dt, t_total, _ptr(d1))
    if rc != 0:
        raise ChanError(lib.chan_io_last_error().decode())


def savevel_read(path):
    """Read a savevel.out file: returns (header dict, data1 [3][ny+1][nx/2+1][nz])."""
    lib = load_library()
    n = [C.c_int() for _ in range(3)]
    r = [C.c_double() for _ in range(5)]
    p = os.fsencode(path)
    if lib.chan_savevel_read_header(p, *(C.byref(v) for v in n), *(C.byref(v) for v in r)) != 0:
        raise ChanError(lib.chan_io_last_error().decode())
    nx, ny, nz = (v.value for v in n)
    hdr = dict(nx=nx, ny=ny, nz=nz, lx=r[0].value, lz=r[1].value, re=r[2].value, dt=r[3].value,
               t_total=r[4].value)
    d1 = np.empty((3, ny + 1, nx // 2 + 1, nz), np.complex128)
    if lib.chan_savevel_read(p, nx, ny, nz, _ptr(d1)) != 0:
        raise ChanError(lib.chan_io_last_error().decode())
    return hdr, d1


def vtk_write(path, phys, lx, lz, yh):
    """Format phys [3][ny+1][nz][nx] (u, v, w on the yh planes) as output_hdf_dfsd's VTK file."""
    lib = load_library()
    ph = _f64(phys)
    _, nyp, nz, nx = ph.shape
    if lib.chan_vtk_write(os.fsencode(path), nx, nyp - 1, nz, lx, lz, _ptr(_f64(yh)), _ptr(ph)) != 0:
        raise ChanError(lib.chan_io_last_error().decode())


def version() -> str:
    return load_library().chan_version().decode()
