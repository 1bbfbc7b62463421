"""CPU checks of the C ABI library: it loads, exports every symbol the header
declares, and validates options without touching a GPU."""
import ctypes as C
import os
import re

import pytest

from abhsf_amd import capi

from conftest import ROOT


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "abhsf_c.h")).read()
    return sorted(set(re.findall(r"^(?:abhsf_status|const char \*|int64_t)\s*(abhsf_\w+)\s*\(", src, re.M)))


def test_header_declares_expected_entry_points():
    assert declared_symbols() == sorted(capi.EXPORTED)


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    for name in declared_symbols():
        assert hasattr(L, name), name
    assert b"gfx950" in L.abhsf_version()


def test_library_has_gfx950_code_object():
    data = open(capi.LIB_PATH, "rb").read()
    assert b"gfx950" in data


@pytest.mark.parametrize("bs,policy,prec,variant,ok", [
    (64, 0, 0, 0, True), (2, 1, 1, 2, True), (256, 0, 0, 1, True),
    (1, 0, 0, 0, False), (48, 0, 0, 0, False), (512, 0, 0, 0, False),
    (64, 0, 0, 3, True), (64, 0, 0, 4, True), (64, 2, 0, 0, False), (64, 0, 2, 0, False),
    (64, 0, 0, 5, False),
])
def test_create_validates_options(bs, policy, prec, variant, ok):
    opts = capi.Options(bs, policy, prec, variant, 0, 0, None)
    h = C.c_void_p(0)
    rc = capi.lib().abhsf_create(C.byref(opts), C.byref(h))
    if ok:
        assert rc == capi.OK and h.value
        assert capi.lib().abhsf_destroy(h) == capi.OK
    else:
        assert rc == capi.ERR_INVALID and not h.value
        assert capi.lib().abhsf_last_error()


def test_null_arguments_and_state_errors():
    L = capi.lib()
    assert L.abhsf_create(None, None) == capi.ERR_INVALID
    opts = capi.Options(8, 0, 0, 0, 0, 0, None)
    h = C.c_void_p(0)
    assert L.abhsf_create(C.byref(opts), C.byref(h)) == capi.OK
    st = capi.Stats()
    assert L.abhsf_get_stats(h, C.byref(st)) == capi.ERR_STATE
    assert L.abhsf_spmv(h, 1.0, None, 1.0, None) == capi.ERR_INVALID
    assert b"null" in L.abhsf_last_error()
    # bad shapes are rejected before any device work
    assert L.abhsf_build_csr(h, 0, 5, 0, None, None, None, 0) == capi.ERR_INVALID
    assert L.abhsf_destroy(h) == capi.OK


def test_package_fails_loudly_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(capi, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(capi, "_lib", None)
    with pytest.raises(OSError):
        capi.lib()


def test_symmetry_argument_checks():
    """Checked before any GPU work: unknown symmetry, non-square and FP32
    symmetric storage are refused."""
    import numpy as np
    from abhsf_amd import AbhsfError, AbhsfMatrix
    M = AbhsfMatrix(block_size=16)
    with pytest.raises(AbhsfError):
        capi.call("abhsf_set_symmetry", M._h, 7)
    rp = np.array([0, 1, 1], dtype=np.int64)
    col = np.array([0], dtype=np.int32)
    with pytest.raises(AbhsfError, match="square"):
        M.from_csr(2, 3, rp, col, np.ones(1), symmetry="symmetric")
    M32 = AbhsfMatrix(block_size=16, precision="f32")
    with pytest.raises(AbhsfError, match="FP64"):
        M32.from_csr(2, 2, rp, col, np.ones(1), symmetry="symmetric")
