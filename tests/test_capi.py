"""CPU checks of the C-ABI library: it loads, exports every entry point the
header declares, and rejects invalid configurations before touching a GPU."""
import ctypes

import pytest

from pip_amd import _lib
from pip_amd.config import PIPConfig


def test_library_exports_every_header_symbol():
    L = _lib.lib()
    names = _lib.header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert b"gfx950" in L.pip_version()


def test_snapshot_library_exports_every_header_symbol():
    import os
    import re

    from pip_amd import build, snapshot
    if not os.path.exists(os.path.join(build.HDF5_PREFIX, "include", "hdf5.h")):
        pytest.skip("no HDF5 C library in this image")
    if not os.path.exists(build.IOLIB):
        build.build_io()
    src = open(os.path.join(os.path.dirname(_lib.HEADER), "pipio.h")).read()
    names = sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*\*?\s*(pipio_\w+)\s*\(", src, re.M)))
    assert names == ["pipio_last_error", "pipio_read", "pipio_write"]
    L = snapshot.lib()
    for n in names:
        assert hasattr(L, n), n


@pytest.mark.parametrize("bad", [
    dict(ndim=3), dict(nvar_m=9), dict(t_order=4), dict(flag_ir=1), dict(ix=6),
    dict(flag_bnd=(0, 1, 1, 1)), dict(dx=0.0), dict(flag_col=8), dict(flag_col=0),
    dict(flag_col=3, ac_field=1), dict(flag_pip_imp=2), dict(flag_pip_imp=1, pip_imp_factor=0.0),
    dict(ix=40000),
])
def test_create_rejects_invalid_config(bad):
    cfg = PIPConfig(ndim=1, nx=64)
    c = _lib.cfg_to_c(cfg)
    for k, v in bad.items():
        if k == "flag_bnd":
            for i in range(4):
                c.flag_bnd[i] = v[i]
        else:
            setattr(c, k, v)
    h = ctypes.c_void_p()
    rc = _lib.lib().pip_create(ctypes.byref(c), ctypes.byref(h))
    assert rc in (-1, -4)
    assert h.value is None
    assert len(_lib.lib().pip_last_error()) > 0


def test_null_arguments_are_errors():
    L = _lib.lib()
    assert L.pip_create(None, None) == -1
    assert L.pip_destroy(None) == 0
    assert L.pip_sync(None) == -1
    assert L.pip_ens_set_mode(None, 1) == -1
    # argument checks come before any device call
    assert L.pip_check_arith(None, None, None, None, 4) == -1
    assert L.pip_check_arith(None, None, None, None, -1) == -1
    assert L.pip_check_arith(None, None, None, None, 0) == 0
