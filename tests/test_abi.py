"""The C-ABI library loads and exports every entry point include/lbhip.h declares (no GPU needed); without a HIP
device the product path raises instead of falling back to the CPU."""
import ctypes
import os
import re

import pytest

from lagbound_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "lbhip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(lb_[a-z_A-Z0-9]+)\s*\(", src)))


def test_header_matches_binding():
    assert _declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(_lib.LIBPATH)
    for name in _declared():
        assert hasattr(lib, name), name


def test_version_and_kernel_names():
    L = _lib.load()
    assert b"gfx950" in L.lb_version()
    names = [L.lb_kernel_name(k) for k in range(_lib.LB_NKERNELS)]
    assert b"lb_k_rows" in names and L.lb_kernel_name(99) == b""


def test_bad_arguments_rejected_without_device():
    L = _lib.load()
    ctx = ctypes.c_void_p()
    assert L.lb_create(ctypes.byref(ctx), 0, 0, 12) == -1      # unsupported rank
    assert L.lb_create(ctypes.byref(ctx), 0, 7, 16) == -1      # unsupported dtype
    assert L.lb_create(None, 0, 0, 16) == -1


def test_no_cpu_fallback():
    """Where there is no HIP device, creating an engine raises (nothing silently runs on the CPU)."""
    from lagbound_amd.engine import Engine
    try:
        e = Engine(r=16, dtype="f32")
    except _lib.LBError:
        return
    e.close()
    pytest.skip("a HIP device is present")


def test_certificate_library_exports_every_symbol():
    """liblbcert.so (the host part of the sparse certificate) exports every entry point include/lbcert.h declares."""
    from lagbound_amd import sparse_cert
    src = open(os.path.join(ROOT, "include", "lbcert.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = sorted(set(re.findall(r"\b(lbc_[a-z_A-Z0-9]+)\s*\(", src)))
    assert names == ["lbc_core", "lbc_factor", "lbc_free", "lbc_info", "lbc_md_order", "lbc_order"]
    lib = ctypes.CDLL(sparse_cert.CERTLIB)
    for name in names:
        assert hasattr(lib, name), name


def test_verifier_library_exports_every_symbol():
    """liblbverify.so (the independent verifier's factorization) exports every entry point include/lbverify.h
    declares."""
    from lagbound_amd import verify
    src = open(os.path.join(ROOT, "include", "lbverify.h")).read()
    names = sorted(set(re.findall(r"\b(lbv_[a-z_A-Z0-9]+)\s*\(", src)))
    assert names == ["lbv_ldl"]
    lib = ctypes.CDLL(verify.VERIFYLIB)
    for name in names:
        assert hasattr(lib, name), name


def test_host_bookkeeping_library_exports_every_symbol():
    """liblbhost.so (the search's host bookkeeping) exports every entry point include/lbhost.h declares, and the
    binding lists the same names."""
    from lagbound_amd import _host
    src = open(os.path.join(ROOT, "include", "lbhost.h")).read()
    src = re.sub(r"//[^\n]*", "", src)
    names = sorted(set(re.findall(r"\b(lbh_[a-z_A-Z0-9]+)\s*\(", src)))
    assert names == sorted(_host.EXPORTS)
    lib = ctypes.CDLL(_host.LIBPATH)
    for name in names:
        assert hasattr(lib, name), name
