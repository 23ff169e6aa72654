"""CPU checks of the C ABI boundary (no compute calls) and host-side logic."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import helpers as Hp
from ed_code_amd import lattices as L

ROOT = Hp.ROOT


def _declared_symbols():
    src = open(os.path.join(ROOT, "include", "ed_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ed_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from ed_code_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from ed_code_amd import build
        build.build()
    lib = C.CDLL(_lib.LIB_PATH)
    syms = _declared_symbols()
    assert len(syms) >= 18
    for s in syms:
        assert hasattr(lib, s), s
    # the Python binding covers exactly the declared set
    assert set(_lib.SIGNATURES) == set(syms)


def test_last_error_without_device_is_safe():
    from ed_code_amd import _lib
    lib = _lib.lib()
    # error string accessor never touches the GPU
    assert lib.ed_last_error() is not None


def test_argument_checks_before_any_device_call():
    """Entry points reject bad arguments with ED_EINVAL before they touch the
    GPU (so these run without one)."""
    from ed_code_amd import _lib
    lib = _lib.lib()
    EINVAL = 1
    n = C.c_int64(-1)
    assert lib.ed_unwrap(None, None, 0, None, None, C.byref(n)) == EINVAL
    assert lib.ed_pair_correlations(None, None, None, 1, None, None, None) == EINVAL
    # CSR product: negative size, row pointers not starting at 0, decreasing,
    # and a column out of range
    v = np.zeros(4)
    w = np.zeros(4)
    assert lib.ed_csr_hv_host(-1, 2, None, None, None, None, None) == EINVAL
    rp = np.array([1, 1, 2], np.int64)
    col = np.array([0, 1], np.int64)
    val = np.zeros(4)
    assert lib.ed_csr_hv_host(2, 2, rp.ctypes.data, col.ctypes.data, val.ctypes.data, v.ctypes.data,
                              w.ctypes.data) == EINVAL
    rp = np.array([0, 2, 1], np.int64)
    assert lib.ed_csr_hv_host(2, 2, rp.ctypes.data, col.ctypes.data, val.ctypes.data, v.ctypes.data,
                              w.ctypes.data) == EINVAL
    rp = np.array([0, 1, 2], np.int64)
    bad = np.array([0, 5], np.int64)
    assert lib.ed_csr_hv_host(2, 2, rp.ctypes.data, bad.ctypes.data, val.ctypes.data, v.ctypes.data,
                              w.ctypes.data) == EINVAL
    assert b"column 5" in lib.ed_last_error()
    # an empty matrix is a no-op
    assert lib.ed_csr_hv_host(0, 0, None, None, None, None, None) == 0


def test_lattice_definitions():
    sq = L.square(4, 8, J2=0.5)
    assert sq.n_sites == 32 and len(sq.bonds) == 128 and sq.orders() == (4, 8)
    tri = L.triangular(6, 5, Ch=0.1)
    assert len(tri.bonds) == 90 and len(tri.triangles) == 60 and len(tri.chiral) == 180
    t40 = L.tilted_square((6, 2), (-2, 6), J2=0.5)
    assert t40.n_sites == 40 and len(t40.bonds) == 160 and t40.orders() == (20, 20)
    assert len(Hp.group_elements(t40.T1, t40.T2)) == 40
    for m in (sq, tri, t40, L.triangular(3, 9), L.chain(24)):
        T1, T2 = np.asarray(m.T1), np.asarray(m.T2)
        assert np.array_equal(T1[T2], T2[T1])                      # translations commute
        assert sorted(T1.tolist()) == list(range(m.n_sites))
        # every bond is mapped onto a bond by both translations (translation invariance)
        bset = {frozenset((x, y)) for (x, y, _, _) in m.bonds}
        for (x, y, _, _) in m.bonds:
            assert frozenset((int(T1[x]), int(T1[y]))) in bset
            assert frozenset((int(T2[x]), int(T2[y]))) in bset


def test_bowtie_expansion_counts():
    tri = L.triangular(4, 3, ChCh=0.2)
    # 9 entries per ordered pair of neighbouring triangles
    assert len(tri.bowties) % 9 == 0 and len(tri.bowties) > 0


def test_sz_rule():
    assert L.n_ones_for_sz(32, 0.0) == 16
    assert L.n_ones_for_sz(27, 0.0) == 13
    assert L.n_ones_for_sz(27, 1.0) == 14
    assert L.n_ones_for_sz(24, -1.0) == 11


def test_bench_refuses_multi_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr
