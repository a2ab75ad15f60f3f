"""C-ABI checks that need no GPU: the library loads, exports every entry
point declared in include/chanflow.h, and validates parameters before it
touches the device."""
import ctypes
import os
import re

import pytest

from helpers import ROOT, chanflow

HEADER = os.path.join(ROOT, "include", "chanflow.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(chan_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "chan_create" in syms and "chan_step" in syms and "chan_linearized_matvec" in syms
    assert len(syms) >= 30


def test_library_exports_every_declared_symbol():
    cf = chanflow()
    lib = cf.load_library()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    # the Python binding wraps exactly the declared set
    assert sorted(cf.ENTRY_POINTS) == declared_symbols()


def test_version_string():
    cf = chanflow()
    v = cf.version()
    assert v.startswith("chanflow") and "gfx950" in v and "fp64" in v


@pytest.mark.parametrize("bad,msg", [
    (dict(nx=31), "even"),
    (dict(ny=4), "ny must be"),
    (dict(nx2=30), "nx2 >= nx"),
    (dict(nz2=50), "2^a 3^b"),
    (dict(bl_grid=1), "bl_grid"),
    (dict(re=-1.0), "re/lx/lz"),
])
def test_create_validates_parameters(bad, msg):
    cf = chanflow()
    lib = cf.load_library()
    p = dict(nx=32, ny=65, nz=32, nx2=48, nz2=48, lx=6.28, lz=3.14, re=2500.0, beta=0.0, fix_mass=1,
             dt_flag=0, cfl_max=0.95, ctrlu=0.0, ctrlw=0.0, bl_grid=0)
    p.update(bad)
    prm = cf.ChanParams(**p)
    h = ctypes.c_void_p()
    assert lib.chan_create(ctypes.byref(prm), ctypes.byref(h)) != 0
    assert msg in lib.chan_last_error(None).decode()
    assert not h.value


def test_null_context_is_rejected():
    cf = chanflow()
    lib = cf.load_library()
    assert lib.chan_sync(None) != 0
    assert lib.chan_destroy(None) == 0


def test_null_ensemble_is_rejected():
    """Ensemble entry points refuse a null ensemble before any device call."""
    import ctypes as C
    lib = chanflow().load_library()
    ms = C.c_double(-1.0)
    assert lib.chan_ensemble_step(None, 1) != 0
    assert lib.chan_ensemble_autotune(None, 2, C.byref(ms)) != 0
    assert ms.value == -1.0
    assert lib.chan_ensemble_invalidate(None) != 0
    assert lib.chan_ensemble_destroy(None) == 0


C_CLIENT = r"""
#include <stdio.h>
#include <string.h>
#include "chanflow.h"
int main(void) {
  chan_params p = {32, 65, 31, 48, 48, 6.28, 3.14, 2500.0, 0.0, 1, 0, 0.95, 0.0, 0.0, 0};  /* nz odd */
  chan_ctx* ctx = NULL;
  if (strncmp(chan_version(), "chanflow", 8) != 0) return 2;
  if (chan_create(&p, &ctx) == 0 || ctx != NULL) return 3;
  printf("%s\n", chan_last_error(NULL));
  return chan_destroy(NULL);
}
"""


def test_plain_c_client(tmp_path):
    """A C99 program includes the header, links libchanflow.so and calls
    the ABI: chan_create rejects an odd nz with the library's message, with no
    device touched."""
    import shutil
    import subprocess
    cf = chanflow()
    lib_path = cf.load_library()._name
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    src = tmp_path / "client.c"
    src.write_text(C_CLIENT)
    exe = tmp_path / "client"
    lib_dir = os.path.dirname(lib_path)
    subprocess.run([gcc, "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                    "-L", lib_dir, "-lchanflow", "-Wl,-rpath," + lib_dir, "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "even" in r.stdout


def test_noalloc_scratch_size_matches_header():
    import re as _re
    m = _re.search(r"#define\s+CHAN_NOALLOC_SCRATCH_BYTES\s+\((\d+)u\s*<<\s*(\d+)\)", open(HEADER).read())
    assert m, "CHAN_NOALLOC_SCRATCH_BYTES not found in the header"
    assert chanflow().CHAN_NOALLOC_SCRATCH_BYTES == int(m.group(1)) << int(m.group(2))


def test_noalloc_entry_point_rejects_bad_arguments_without_gpu():
    cf = chanflow()
    lib = cf.load_library()
    # a null context is refused before any device call
    assert lib.chan_compute_divergence_of_fluxes_noalloc(None, None, None, None, None, 1, None) != 0
