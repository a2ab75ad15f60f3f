"""The allocation-free entry point chan_compute_divergence_of_fluxes_noalloc
(device arrays in and out, x-transform intermediates in the library's fixed
per-GPU scratch array) against the allocating entry point
chan_compute_divergence_of_fluxes on the same inputs, on two calls in a row,
the second with fewer planes than the first: the results agree bit for bit
(the same kernels on the same data), and both agree with the oracle's
divergence of fluxes (PFC/transform.f90:573-690) to relative L2 1e-12.  A
call that needs more scratch than the array holds is refused with the
library's message."""
import numpy as np
import pytest

from helpers import chanflow, grid_cfg0, make_channel, rel_l2
from oracle import chan_oracle as O

pytestmark = pytest.mark.gpu


class _HipMem:
    """Device buffers through the HIP runtime itself (ctypes), so that the
    caller's memory comes from outside the library under test."""

    def __init__(self):
        import ctypes
        self.C = ctypes
        # the HIP runtime libchanflow.so is linked against (already loaded); torch
        # would bring a second HIP runtime of its own into the process
        self.lib = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.bufs = []

    def alloc(self, nbytes):
        p = self.C.c_void_p()
        assert self.lib.hipMalloc(self.C.byref(p), nbytes) == 0
        self.bufs.append(p.value)
        return p.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        d = self.alloc(arr.nbytes)
        assert self.lib.hipMemcpy(d, arr.ctypes.data, arr.nbytes, 1) == 0      # host to device
        return d

    def download(self, d, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.lib.hipMemcpy(out.ctypes.data, d, out.nbytes, 2) == 0     # device to host
        return out

    def free(self):
        for b in self.bufs:
            self.lib.hipFree(b)
        self.bufs = []


def test_noalloc_divergence_of_fluxes_matches_allocating_entry_point():
    cf = chanflow()
    g = grid_cfg0(O)
    ch = make_channel(g)
    rng = np.random.default_rng(17)
    mem = _HipMem()
    for npl in (40, 13):
        print(f"noalloc divergence of fluxes: {npl} planes of {g.nz2}x{g.nx2}", flush=True)
        fx, fy, fz = (rng.standard_normal((npl, g.nz2, g.nx2)) for _ in range(3))
        ref_alloc = ch.div_fluxes(fx, fy, fz)
        dx, dy, dz = (mem.upload(f) for f in (fx, fy, fz))
        dout = mem.alloc(npl * g.nxh * g.nz * 16)
        ch.sync()
        ch.div_fluxes_noalloc(dx, dy, dz, dout, npl)
        ch.sync()
        got = mem.download(dout, (npl, g.nxh, g.nz), np.complex128)
        assert np.array_equal(got, ref_alloc)
        ref = O.div_fluxes(g, fx, fy, fz)
        assert rel_l2(got, ref) <= 1e-12
        mem.free()
    # more planes than the scratch array holds
    aps = g.nz2 * ((g.nxh + 7) // 8 * 8)
    too_many = cf.CHAN_NOALLOC_SCRATCH_BYTES // (16 * 2 * aps) + 1
    big = mem.alloc(too_many * g.nz2 * g.nx2 * 8)
    outb = mem.alloc(too_many * g.nxh * g.nz * 16)
    with pytest.raises(cf.ChanError, match="scratch"):
        ch.div_fluxes_noalloc(big, big, big, outb, too_many)
    mem.free()
    ch.close()
