"""The C ABI: libffs_hip.so loads without a GPU, exports every symbol that
include/ffs/ffs.h declares, and the ctypes mirror of ffs_params matches the
C layout (checked against a gcc-compiled probe of the header)."""
import math
import ctypes
import os
import re
import subprocess

from conftest import REPO

HEADER = os.path.join(REPO, "include", "ffs", "ffs.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(ffs_\w+)\s*\(", src, flags=re.M)))


def test_header_declares_the_api():
    names = declared_functions()
    for n in ("ffs_create", "ffs_set_field", "ffs_get_field", "ffs_initialize", "ffs_advance", "ffs_stats_pop",
              "ffs_spectrum", "ffs_get_scalar", "ffs_last_error", "ffs_destroy"):
        assert n in names


def test_library_exports_every_declared_symbol(lib):
    for n in declared_functions():
        assert hasattr(lib, n), n


def test_ctypes_signatures_cover_header(lib):
    from ffs import _lib

    assert set(_lib.SIGNATURES) == set(declared_functions())


def test_params_layout_matches_c(tmp_path):
    from ffs._lib import FfsParams

    fields = [f[0] for f in FfsParams._fields_]
    probe = tmp_path / "probe.c"
    body = "\n".join(f'  printf("{f} %zu\\n", offsetof(ffs_params, {f}));' for f in fields)
    probe.write_text(f"""#include <stdio.h>
#include <stddef.h>
#include "ffs/ffs.h"
int main(void) {{
  printf("__sizeof %zu\\n", sizeof(ffs_params));
{body}
  return 0;
}}
""")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c99", f"-I{os.path.join(REPO, 'include')}", str(probe), "-o", str(exe)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    assert int(out["__sizeof"]) == ctypes.sizeof(FfsParams)
    for f in fields:
        assert int(out[f]) == getattr(FfsParams, f).offset, f


def test_defaults_and_version(lib):
    from ffs import _lib

    header = open(HEADER).read()
    assert lib.ffs_abi_version() == int(re.search(r"#define FFS_ABI_VERSION (\d+)", header).group(1))
    p = _lib.default_params()
    assert p.size == ctypes.sizeof(_lib.FfsParams)
    assert abs(p.cutoff - 2 / 3) < 1e-16 and p.cfl == 0.18 and p.integrator == _lib.FFS_ETD3
    assert p.variable_dt == 1 and math.isnan(p.smagorinsky_c) and math.isnan(p.les_cp)  # the model's defaults


def test_null_arguments_are_rejected(lib):
    assert lib.ffs_create(None, None) == -1
    assert lib.ffs_initialize(None) == -1
    assert lib.ffs_advance(None, 1, None, None) == -1
    lib.ffs_destroy(None)  # no-op


def test_bad_params_fail_cleanly_without_device(lib):
    """Validation happens before any device call, so it runs here too."""
    from ffs import _lib

    p = _lib.default_params()
    p.nx = p.ny = p.nz = 16
    p.size = 12  # wrong ABI size
    h = ctypes.c_void_p()
    rc = lib.ffs_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == -1 and not h.value
    assert b"size" in lib.ffs_last_error(None)


def test_unsupported_sizes_fail_at_creation(lib):
    """An uncompiled transform length, or an x pass that cannot hold a row's fields, fails in
    ffs_create / ffs_footprint with FFS_ERR_ARG before any device work (so this runs without a GPU)."""
    import pytest

    import ffs

    base = dict(nx=32, ny=32, nz=32, nu=0.01, dt=0.01, variable_dt=False)
    assert ffs.footprint(ffs.Params(**base)) > 0
    # 2^a 3^b 5^c lengths are compiled from a list (16 ... 1024); 56, 2048 and 100 are not on it
    for key, bad in (("nx", 56), ("ny", 2048), ("nz", 100)):
        assert ffs.footprint(ffs.Params(**{**base, key: 80})) > 0
        p = ffs.Params(**{**base, key: bad})
        with pytest.raises(RuntimeError) as e:
            ffs.footprint(p)
        assert "not a compiled transform length" in str(e.value), str(e.value)
        with pytest.raises(RuntimeError) as e:
            ffs.Simulation(p)
        assert "not a compiled transform length" in str(e.value), str(e.value)
        assert e.value.code == -1  # FFS_ERR_ARG, before any HIP call
    # Vreman (reads w) with the density's and the passive scalar's LES fluxes: 10 field pairs per row,
    # more than the general x pass of length 512 holds (8); at length 256 it fits
    heavy = dict(les_model="Vreman", density=True, pr=1.0, density_gradient=1.0, z_acceleration=1.0,
                 reference_density=1.0, density_les="eddy", passive_scalar=True, scalar_pr=1.0, scalar_les="eddy")
    with pytest.raises(RuntimeError) as e:
        ffs.footprint(ffs.Params(**{**base, "nx": 512, **heavy}))
    assert "field pairs per row" in str(e.value), str(e.value)
    assert ffs.footprint(ffs.Params(**{**base, "nx": 256, **heavy})) > 0
