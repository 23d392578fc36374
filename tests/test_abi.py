"""The C-ABI library loads and exports every entry point include/lag2d.h
declares (no device calls)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "lag2d.h")).read()
    return sorted(set(re.findall(r"\b(lag2d_[a-z_]+)\s*\(", src)))


def test_header_lists_entry_points():
    names = declared()
    assert "lag2d_step" in names and "lag2d_rhs" in names and "lag2d_create" in names


def test_library_exports_all_declared_symbols():
    import ctypes
    from lag2d_amd import lib
    if not os.path.exists(lib.LIBPATH):
        pytest.fail("liblag2d.so not built (python -m lag2d_amd.build)")
    so = ctypes.CDLL(lib.LIBPATH)
    for n in declared():
        assert hasattr(so, n), n
    assert set(declared()) <= set(lib.EXPORTS)


def test_unsupported_config_is_rejected_without_device():
    """lag2d_create validates switches before touching the device."""
    import ctypes
    import numpy as np
    from lag2d_amd import lib
    L = lib.load()
    cfg = lib.Config(ndegr=3, npoly=1, ncurv=0, nstag=2, ndens=2, nlimi=6, ncase=6, nrz=0, nfint=1,
                     nriem=1, ndt=1, gamlg=1.4, cfl=0.3, cimpd=1.0, cdrho=0.0, tend=0.2)
    ipq = np.array([1, 2, 3, 4], dtype=np.int32)
    bf = np.zeros(5, dtype=np.int32)
    e1 = np.array([1, 1, 1, 1], dtype=np.int32)
    e2 = np.array([0, 1, 2, 3, 4], dtype=np.int32)
    h = ctypes.c_void_p()
    P = ctypes.POINTER(ctypes.c_int)
    rc = L.lag2d_create(ctypes.byref(cfg), 1, 4, 0, ipq.ctypes.data_as(P), bf.ctypes.data_as(P),
                        e1.ctypes.data_as(P), e2.ctypes.data_as(P), ctypes.byref(h))
    assert rc == 2
    assert b"ndens" in L.lag2d_last_error()


@pytest.mark.parametrize("change,needle", [
    ({"nrz": 1}, b"nrz=0 or 2"),
    ({"nrz": 2, "ncurv": 1, "nfint": 4, "npoly": 2, "ndegr": 6}, b"nrz=2: straight P1"),
    ({"nrz": 2, "ncase": 12}, b"ncase"),
    ({"nrz": 2, "code": 25}, b"slip walls"),
])
def test_unsupported_rz_config_is_rejected_without_device(change, needle):
    """The RZ path takes straight P1 cells with slip walls; other RZ
    switches are refused before the device is touched."""
    import ctypes
    import numpy as np
    from lag2d_amd import lib
    L = lib.load()
    kw = dict(ndegr=3, npoly=1, ncurv=0, nstag=2, ndens=1, nlimi=6, ncase=7, nrz=0, nfint=1,
              nriem=1, ndt=1, gamlg=1.4, cfl=0.3, cimpd=1.0, cdrho=0.0, tend=0.2)
    code = change.pop("code", 22)
    kw.update(change)
    cfg = lib.Config(**kw)
    ipq = np.array([1, 2, 3, 4], dtype=np.int32)
    bf = np.array([1, 2, code, 221, 0], dtype=np.int32)
    e1 = np.array([1, 1, 1, 1], dtype=np.int32)
    e2 = np.array([0, 1, 2, 3, 4], dtype=np.int32)
    h = ctypes.c_void_p()
    P = ctypes.POINTER(ctypes.c_int)
    rc = L.lag2d_create(ctypes.byref(cfg), 1, 4, 1, ipq.ctypes.data_as(P), bf.ctypes.data_as(P),
                        e1.ctypes.data_as(P), e2.ctypes.data_as(P), ctypes.byref(h))
    assert rc == 2
    assert needle in L.lag2d_last_error()


@pytest.mark.parametrize("e2,needle", [
    ([1, 1, 2, 3, 4], b"esuv2(1)"),
    ([-1, 1, 2, 3, 4], b"esuv2(1)"),
    ([0, 2, 1, 3, 4], b"monotone"),
])
def test_bad_csr_is_rejected_without_device(e2, needle):
    """A CSR whose first offset is not 0 or whose offsets decrease would
    make the host-side neighbour builders index esuv1 out of bounds; it is
    refused before any of them (or the device) runs."""
    import ctypes
    import numpy as np
    from lag2d_amd import lib
    L = lib.load()
    cfg = lib.Config(ndegr=3, npoly=1, ncurv=0, nstag=2, ndens=1, nlimi=6, ncase=6, nrz=0, nfint=1,
                     nriem=1, ndt=1, gamlg=1.4, cfl=0.3, cimpd=1.0, cdrho=0.0, tend=0.2)
    ipq = np.array([1, 2, 3, 4], dtype=np.int32)
    bf = np.zeros(5, dtype=np.int32)
    e1 = np.array([1, 1, 1, 1], dtype=np.int32)
    e2 = np.array(e2, dtype=np.int32)
    h = ctypes.c_void_p()
    P = ctypes.POINTER(ctypes.c_int)
    rc = L.lag2d_create(ctypes.byref(cfg), 1, 4, 0, ipq.ctypes.data_as(P), bf.ctypes.data_as(P),
                        e1.ctypes.data_as(P), e2.ctypes.data_as(P), ctypes.byref(h))
    assert rc == 1
    assert needle in L.lag2d_last_error()


@pytest.mark.parametrize("change,needle", [
    ({"npoly": 1, "ndegr": 3}, b"ncurv=1/nfint=5/P2"),
    ({"ncase": 13}, b"nfint=5: ncase"),
    ({"code": 21}, b"nfint=5: slip walls"),
])
def test_unsupported_nfint5_config_is_rejected_without_device(change, needle):
    """The nfint=5 branch (face Gauss points) runs P2 with slip and fixed
    walls; its other switches are refused before the device is touched."""
    import ctypes
    import numpy as np
    from lag2d_amd import lib
    L = lib.load()
    kw = dict(ndegr=6, npoly=2, ncurv=1, nstag=3, ndens=1, nlimi=6, ncase=7, nrz=0, nfint=5,
              nriem=1, ndt=1, gamlg=1.4, cfl=0.25, cimpd=1.0, cdrho=1.0, tend=1.0)
    code = change.pop("code", 22)
    kw.update(change)
    cfg = lib.Config(**kw)
    ipq = np.arange(1, 10, dtype=np.int32)
    bf = np.array([1, 2, code, 221, 5], dtype=np.int32)
    e1 = np.ones(9, dtype=np.int32)
    e2 = np.arange(0, 10, dtype=np.int32)
    h = ctypes.c_void_p()
    P = ctypes.POINTER(ctypes.c_int)
    rc = L.lag2d_create(ctypes.byref(cfg), 1, 9, 1, ipq.ctypes.data_as(P), bf.ctypes.data_as(P),
                        e1.ctypes.data_as(P), e2.ctypes.data_as(P), ctypes.byref(h))
    assert rc == 2
    assert needle in L.lag2d_last_error()


def test_ncase8_is_rejected_without_device():
    """ncase 8 lowers the CFL number mid-run (Getrhslag.f90:1785), which the
    library does not reproduce: it refuses the problem rather than drift."""
    import ctypes
    import numpy as np
    from lag2d_amd import lib
    L = lib.load()
    cfg = lib.Config(ndegr=3, npoly=1, ncurv=0, nstag=2, ndens=1, nlimi=6, ncase=8, nrz=0, nfint=1,
                     nriem=1, ndt=1, gamlg=1.4, cfl=0.3, cimpd=1.0, cdrho=0.0, tend=0.2)
    ipq = np.array([1, 2, 3, 4], dtype=np.int32)
    bf = np.zeros(5, dtype=np.int32)
    e1 = np.array([1, 1, 1, 1], dtype=np.int32)
    e2 = np.array([0, 1, 2, 3, 4], dtype=np.int32)
    h = ctypes.c_void_p()
    P = ctypes.POINTER(ctypes.c_int)
    rc = L.lag2d_create(ctypes.byref(cfg), 1, 4, 0, ipq.ctypes.data_as(P), bf.ctypes.data_as(P),
                        e1.ctypes.data_as(P), e2.ctypes.data_as(P), ctypes.byref(h))
    assert rc == 2
    assert b"ncase 8" in L.lag2d_last_error()


def test_fortran_interface_compiles(tmp_path):
    """The ISO_C_BINDING module (lag2d_amd/fortran/lag2d_c.f90) compiles on
    its own: it depends on nothing of the reference."""
    import shutil
    import subprocess
    fc = shutil.which("amdflang") or shutil.which("flang") or shutil.which("gfortran")
    if fc is None:
        import pytest
        pytest.skip("no Fortran compiler")
    src = os.path.join(ROOT, "lag2d_amd", "fortran", "lag2d_c.f90")
    r = subprocess.run([fc, "-c", src, "-o", str(tmp_path / "lag2d_c.o")], cwd=tmp_path,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
