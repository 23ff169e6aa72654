"""ctypes bindings for the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads oracle/libedoracle.so (the restatement) and, when built,
oracle/_ref/libedref.so (the reference's own basis + H*v slice).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libedoracle.so")
REF_SO = os.path.join(HERE, "_ref", "libedref.so")

i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C")


class Terms(C.Structure):
    _fields_ = [
        ("lam", C.c_double), ("nbonds", C.c_int32), ("bond_xy", C.c_void_p), ("bond_J", C.c_void_p),
        ("bond_phase", C.c_void_p), ("Ch", C.c_double), ("nchiral", C.c_int32), ("chiral_ijk", C.c_void_p),
        ("chiral_phase", C.c_void_p), ("ChCh", C.c_double), ("nbowtie", C.c_int32), ("bowtie", C.c_void_p),
    ]


def make_terms(model):
    """Returns (Terms, keepalive) for a lattices.Model."""
    a = model.term_arrays()
    keep = [np.ascontiguousarray(a[k]) for k in ("bond_xy", "bond_J", "bond_phase", "chiral_ijk", "chiral_phase", "bowtie")]
    t = Terms(a["lam"], a["nbonds"], keep[0].ctypes.data, keep[1].ctypes.data, keep[2].ctypes.data,
              a["Ch"], a["nchiral"], keep[3].ctypes.data, keep[4].ctypes.data,
              a["ChCh"], a["nbowtie"], keep[5].ctypes.data)
    return t, keep


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run `make -C oracle`)")
        L = C.CDLL(ORACLE_SO)
        L.or_cexp.argtypes = [C.c_double, C.c_double, f64p]
        L.or_n_choose_k.restype = C.c_int64
        L.or_n_choose_k.argtypes = [C.c_int, C.c_int]
        L.or_gosper.argtypes = [C.c_int, C.c_int, i64p]
        L.or_lin_tables.argtypes = [C.c_int, i64p, i64p]
        L.or_translate.restype = C.c_int64
        L.or_translate.argtypes = [C.c_int64, i32p, C.c_int]
        L.or_LxLy.argtypes = [i32p, i32p, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.or_representative.restype = C.c_int64
        L.or_representative.argtypes = [C.c_int64, i32p, i32p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                        C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                        C.POINTER(C.c_int32)]
        L.or_basis_build.restype = C.c_void_p
        L.or_basis_build.argtypes = [C.c_int, C.c_int, i32p, i32p, C.c_double, C.c_double]
        L.or_basis_build3.restype = C.c_void_p
        L.or_basis_build3.argtypes = [C.c_int, C.c_int, i32p, i32p, i32p, C.c_double, C.c_double, C.c_double]
        L.or_basis_l3.restype = C.c_int32
        L.or_basis_l3.argtypes = [C.c_void_p]
        L.or_representative3.restype = C.c_int64
        L.or_representative3.argtypes = [C.c_int64, i32p, i32p, i32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                         C.c_double, C.c_double] + [C.POINTER(C.c_int32)] * 5
        L.or_basis_free.argtypes = [C.c_void_p]
        L.or_basis_info.argtypes = [C.c_void_p, i64p]
        L.or_basis_lin_error.restype = C.c_int32
        L.or_basis_lin_error.argtypes = [C.c_void_p]
        L.or_basis_copy.argtypes = [C.c_void_p] + [C.c_void_p] * 7
        L.or_hv.argtypes = [C.c_void_p, C.POINTER(Terms), f64p, f64p]
        L.or_hv_rows.argtypes = [C.c_void_p, C.POINTER(Terms), f64p, f64p, C.c_int64, C.c_int64]
        L.or_count_targets.argtypes = [C.c_void_p, C.POINTER(Terms), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.or_make_h.argtypes = [C.c_void_p, C.POINTER(Terms), f64p]
        L.or_start_vector.argtypes = [C.c_uint64, C.c_int64, f64p]
        L.or_density_matrix.argtypes = [C.c_int64, i64p, f64p, C.c_int64, i64p, f64p, C.c_int, i32p, C.c_int, f64p]
        L.or_tables_new.restype = C.c_void_p
        L.or_tables_new.argtypes = [C.c_int, C.c_int, i32p, i32p, C.c_double, C.c_double, C.c_int64, C.c_int64,
                                    C.c_int64, C.c_int]
        L.or_tables_ptrs.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        L.or_hv_rows_search.argtypes = [C.c_int, i32p, i32p, C.c_double, C.c_double, C.c_int64, C.c_void_p,
                                        C.POINTER(Terms), i64p, f64p, C.c_int64, C.c_int64, i64p, f64p]
        L.or_rows_targets_search.restype = C.c_int64
        L.or_rows_targets_search.argtypes = [C.c_int, i32p, i32p, C.c_double, C.c_double, C.c_int64, C.c_void_p,
                                             C.POINTER(Terms), C.c_int64, i64p, C.c_int64, C.c_void_p]
        L.or_lanczos.restype = C.c_int64
        L.or_lanczos.argtypes = [C.c_void_p, C.POINTER(Terms), C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64,
                                 f64p, C.c_void_p]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"reference slice not built: {REF_SO}")
        L = C.CDLL(REF_SO)
        L.ref_n_choose_k.restype = C.c_int64
        L.ref_n_choose_k.argtypes = [C.c_int, C.c_int]
        L.ref_constrained_dets.argtypes = [C.c_int, C.c_int, i64p]
        L.ref_LxLy.argtypes = [i32p, i32p, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.ref_representative.restype = C.c_int64
        L.ref_representative.argtypes = [C.c_int64, i32p, i32p, C.c_int, C.c_double, C.c_double,
                                         C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                         C.POINTER(C.c_int32)]
        L.ref_basis_build.restype = C.c_void_p
        L.ref_basis_build.argtypes = [C.c_int, C.c_int, i32p, i32p, C.c_double, C.c_double]
        L.ref_basis_build3.restype = C.c_void_p
        L.ref_basis_build3.argtypes = [C.c_int, C.c_int, i32p, i32p, i32p, C.c_double, C.c_double, C.c_double]
        L.ref_basis_free.argtypes = [C.c_void_p]
        L.ref_basis_info.argtypes = [C.c_void_p, i64p]
        L.ref_basis_copy.argtypes = [C.c_void_p] + [C.c_void_p] * 7
        L.ref_hv.argtypes = [C.c_void_p, C.c_void_p, f64p, f64p, C.c_int64]
        L.ref_make_h.argtypes = [C.c_void_p, C.c_void_p, f64p]
        L.ref_unwrap.restype = C.c_int64
        L.ref_unwrap.argtypes = [C.c_void_p, f64p, C.c_int64, C.c_void_p, C.c_void_p]
        L.ref_tables_new.restype = C.c_void_p
        L.ref_tables_new.argtypes = [C.c_int, i32p, i32p, C.c_double, C.c_double, C.c_int64, C.c_int64, C.c_int64,
                                     C.c_int]
        L.ref_tables_ptrs.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        _ref = L
    return _ref


def _ptr(a):
    return None if a is None else a.ctypes.data


class OracleBasis:
    """Oracle restatement of pConstruct_States (Momentum_Basis.cpp:645-778)."""

    def __init__(self, model, n_ones: int, kx: int = 0, ky: int = 0, kz: int = 0):
        self.model = model
        self.n_ones = n_ones
        self.px, self.py = model.momentum(kx, ky)
        L = lib()
        if getattr(model, "T3", None) is not None:
            # three generators: Momentum_Basis.cpp:781-897, pz as Lanczos_sector.cpp:240
            self.pz = float(2 * kz) * math.pi / model.order3()
            self.h = L.or_basis_build3(model.n_sites, n_ones, np.ascontiguousarray(model.T1, np.int32),
                                       np.ascontiguousarray(model.T2, np.int32),
                                       np.ascontiguousarray(model.T3, np.int32), self.px, self.py, self.pz)
        else:
            self.h = L.or_basis_build(model.n_sites, n_ones, np.ascontiguousarray(model.T1, np.int32),
                                      np.ascontiguousarray(model.T2, np.int32), self.px, self.py)
        info = np.zeros(8, np.int64)
        L.or_basis_info(self.h, info)
        self.N, _, self.L1, self.L2, self.half, self.bits_right, self.C, self.D = [int(x) for x in info]
        self.terms, self._keep = make_terms(model)

    def __del__(self):
        try:
            lib().or_basis_free(self.h)
        except Exception:
            pass

    def tables(self, full=True):
        D, C = self.D, self.C
        reps = np.zeros(D, np.int64)
        norm = np.zeros(D)
        norm_inv = np.zeros(D)
        rep_loc = np.zeros(C, np.int64) if full else None
        sph = np.zeros(2 * C) if full else None
        nh = 1 << self.half
        Ia = np.zeros(nh, np.int64)
        Ib = np.zeros(nh, np.int64)
        lib().or_basis_copy(self.h, _ptr(reps), _ptr(norm), _ptr(norm_inv), _ptr(rep_loc), _ptr(sph), _ptr(Ia), _ptr(Ib))
        out = dict(reps=reps, norm=norm, norm_inv=norm_inv, Ia=Ia, Ib=Ib)
        if full:
            out["rep_loc"] = rep_loc
            out["state_ph"] = sph.view(np.complex128)
        return out

    def set_model(self, model):
        self.model = model
        self.terms, self._keep = make_terms(model)

    def hv(self, v: np.ndarray) -> np.ndarray:
        v = np.ascontiguousarray(v, np.complex128)
        w = np.zeros(self.D, np.complex128)
        lib().or_hv(self.h, C.byref(self.terms), v.view(np.float64), w.view(np.float64))
        return w

    def hv_rows(self, v, r0, r1):
        v = np.ascontiguousarray(v, np.complex128)
        w = np.zeros(self.D, np.complex128)
        lib().or_hv_rows(self.h, C.byref(self.terms), v.view(np.float64), w.view(np.float64), r0, r1)
        return w

    def make_h(self) -> np.ndarray:
        H = np.zeros(self.D * self.D, np.complex128)
        lib().or_make_h(self.h, C.byref(self.terms), H.view(np.float64))
        return H.reshape(self.D, self.D, order="F")

    def count_targets(self):
        g = C.c_int64(); t = C.c_int64()
        lib().or_count_targets(self.h, C.byref(self.terms), C.byref(g), C.byref(t))
        return g.value, t.value

    def unwrap(self, v):
        """Unwrap_pState (Momentum_Basis.cpp:978-1037; three generators
        :901-975): for each representative, the reference's walk (T1 steps
        outer, T2 steps cumulative inside, T3 steps cumulative innermost), the
        phase a running product of exp(-i px), exp(-i py), exp(-i pz) per step,
        coefficient v_i * ph / Norm_i, the first visit of a state kept
        (add2ids, idsClass.cpp:33-53); entries with |re| or |im| > 1e-8,
        sorted by state.  Pure Python: small sectors only."""
        t = self.tables(full=False)
        m = self.model
        T1, T2 = np.asarray(m.T1), np.asarray(m.T2)
        three = getattr(m, "T3", None) is not None
        T3 = np.asarray(m.T3) if three else None
        L1, L2 = m.orders()
        L3 = m.order3() if three else 1
        def e(p):   # std::exp(complex(0, -p)): the C library's cexp (not (cos, sin): they differ in the last bit)
            out = np.zeros(2)
            lib().or_cexp(0.0, -p, out)
            return complex(out[0], out[1])

        ex, ey, ez = e(self.px), e(self.py), e(self.pz if three else 0.0)

        def tr(s, T):
            out = 0
            i = 0
            while s:
                if s & 1:
                    out |= 1 << int(T[i])
                s >>= 1
                i += 1
            return out

        def cmul(a, b):   # one rounding per product and per sum, as the C library's complex multiply
            return complex(a.real * b.real - a.imag * b.imag, a.real * b.imag + a.imag * b.real)

        out = []
        for i, rep in enumerate(t["reps"]):
            nrm = float(t["norm"][i])
            vi = complex(v[i])
            seen = {}
            sid = int(rep)
            seen[sid] = complex(vi.real / nrm, vi.imag / nrm)
            ph = complex(1.0, 0.0)

            def add(s, c):
                if s not in seen:
                    seen[s] = complex(c.real / nrm, c.imag / nrm)

            for x in range(L1):
                if x > 0:
                    sid = tr(sid, T1)
                    ph = cmul(ph, ex)
                    add(sid, cmul(vi, ph))
                for y in range(L2):
                    if y > 0:
                        sid = tr(sid, T2)
                        ph = cmul(ph, ey)
                        add(sid, cmul(vi, ph))
                    for z in range(1, L3):
                        sid = tr(sid, T3)
                        ph = cmul(ph, ez)
                        add(sid, cmul(vi, ph))
            out.extend((s, c) for s, c in seen.items() if abs(c.real) > 1e-8 or abs(c.imag) > 1e-8)
        out.sort(key=lambda e_: e_[0])
        return np.array([s for s, _ in out], np.int64), np.array([c for _, c in out], np.complex128)

    def lanczos(self, nev=1, iterations=100, ncycles=10, tol=1e-11, seed=1234, want_vectors=False):
        ev = np.zeros(nev)
        vecs = np.zeros((nev, self.D), np.complex128) if want_vectors else None
        n = lib().or_lanczos(self.h, C.byref(self.terms), nev, iterations, ncycles, tol, seed, ev,
                             None if vecs is None else vecs.ctypes.data)
        return (ev, vecs, n) if want_vectors else ev


def start_vector(seed: int, D: int) -> np.ndarray:
    v = np.zeros(2 * D)
    lib().or_start_vector(seed, D, v)
    return v.view(np.complex128)


class RefBasis:
    """The reference's own pConstruct_States / sHv / makeH (compiled slice)."""

    def __init__(self, model, n_ones: int, kx: int = 0, ky: int = 0, kz: int = 0):
        self.model = model
        self.px, self.py = model.momentum(kx, ky)
        R = ref()
        if getattr(model, "T3", None) is not None:
            self.pz = float(2 * kz) * math.pi / model.order3()
            self.h = R.ref_basis_build3(model.n_sites, n_ones, np.ascontiguousarray(model.T1, np.int32),
                                        np.ascontiguousarray(model.T2, np.int32),
                                        np.ascontiguousarray(model.T3, np.int32), self.px, self.py, self.pz)
        else:
            self.h = R.ref_basis_build(model.n_sites, n_ones, np.ascontiguousarray(model.T1, np.int32),
                                       np.ascontiguousarray(model.T2, np.int32), self.px, self.py)
        info = np.zeros(4, np.int64)
        R.ref_basis_info(self.h, info)
        self.D, self.C, self.nIa, self.bits_right = [int(x) for x in info]
        self.terms, self._keep = make_terms(model)

    def __del__(self):
        try:
            ref().ref_basis_free(self.h)
        except Exception:
            pass

    def tables(self):
        D, C = self.D, self.C
        reps = np.zeros(D, np.int64); norm = np.zeros(D); norm_inv = np.zeros(D)
        rep_loc = np.zeros(C, np.int64); sph = np.zeros(2 * C)
        Ia = np.zeros(self.nIa, np.int64); Ib = np.zeros(self.nIa, np.int64)
        ref().ref_basis_copy(self.h, _ptr(reps), _ptr(norm), _ptr(norm_inv), _ptr(rep_loc), _ptr(sph), _ptr(Ia), _ptr(Ib))
        return dict(reps=reps, norm=norm, norm_inv=norm_inv, rep_loc=rep_loc, state_ph=sph.view(np.complex128),
                    Ia=Ia, Ib=Ib)

    def set_model(self, model):
        self.model = model
        self.terms, self._keep = make_terms(model)

    def hv(self, v, rows=-1):
        v = np.ascontiguousarray(v, np.complex128)
        w = np.zeros(self.D, np.complex128)
        ref().ref_hv(self.h, C.addressof(self.terms), v.view(np.float64), w.view(np.float64), rows)
        return w

    def make_h(self):
        H = np.zeros(self.D * self.D, np.complex128)
        ref().ref_make_h(self.h, C.addressof(self.terms), H.view(np.float64))
        return H.reshape(self.D, self.D, order="F")

    def unwrap(self, v):
        v = np.ascontiguousarray(v, np.complex128)
        n = ref().ref_unwrap(self.h, v.view(np.float64), 0, None, None)
        st = np.zeros(n, np.int64)
        co = np.zeros(2 * n)
        ref().ref_unwrap(self.h, v.view(np.float64), n, _ptr(st), _ptr(co))
        return st, co.view(np.complex128)


class RefSHv:
    """The reference sHv run over tables supplied from outside (timing only)."""

    def __init__(self, model, px, py, C_, D, nlin, bits_right):
        R = ref()
        self.model = model
        self.D = D
        self.h = R.ref_tables_new(model.n_sites, np.ascontiguousarray(model.T1, np.int32),
                                  np.ascontiguousarray(model.T2, np.int32), px, py, C_, D, nlin, bits_right)
        ptrs = (C.c_void_p * 7)()
        R.ref_tables_ptrs(self.h, ptrs)
        self.ptrs = [p for p in ptrs]
        self.terms, self._keep = make_terms(model)

    def __del__(self):
        try:
            ref().ref_basis_free(self.h)
        except Exception:
            pass

    def set_model(self, model):
        self.model = model
        self.terms, self._keep = make_terms(model)

    def hv_rows(self, v, rows):
        v = np.ascontiguousarray(v, np.complex128)
        w = np.zeros(self.D, np.complex128)
        ref().ref_hv(self.h, C.addressof(self.terms), v.view(np.float64), w.view(np.float64), rows)
        return w


class OracleTables:
    """The oracle's row routines (or_hv, or_hv_rows, or_count_targets) over
    pConstruct_States tables supplied from outside (or_tables_new): fill the
    arrays through .ptrs (reps, norm, norm_inv, rep_loc, state_ph, Ia, Ib),
    e.g. with the product's export, whose tables a test pins separately."""

    def __init__(self, model, n_ones, px, py, C_, D, nlin, bits_right):
        self.model = model
        self.D = D
        self.h = lib().or_tables_new(model.n_sites, n_ones, np.ascontiguousarray(model.T1, np.int32),
                                     np.ascontiguousarray(model.T2, np.int32), px, py, C_, D, nlin, bits_right)
        ptrs = (C.c_void_p * 7)()
        lib().or_tables_ptrs(self.h, ptrs)
        self.ptrs = [p for p in ptrs]
        self.terms, self._keep = make_terms(model)

    def __del__(self):
        try:
            lib().or_basis_free(self.h)
        except Exception:
            pass

    def set_model(self, model):
        self.model = model
        self.terms, self._keep = make_terms(model)

    def hv(self, v):
        v = np.ascontiguousarray(v, np.complex128)
        w = np.zeros(self.D, np.complex128)
        lib().or_hv(self.h, C.byref(self.terms), v.view(np.float64), w.view(np.float64))
        return w

    def count_targets(self):
        g = C.c_int64(); t = C.c_int64()
        lib().or_count_targets(self.h, C.byref(self.terms), C.byref(g), C.byref(t))
        return g.value, t.value


def rows_targets_search(model, px, py, reps, rows):
    """Sorted indices of v that the rows gather from (their targets and the
    rows themselves), without full-space tables (or_rows_targets_search)."""
    rows = np.ascontiguousarray(rows, np.int64)
    t, keep = make_terms(model)
    args = (model.n_sites, np.ascontiguousarray(model.T1, np.int32), np.ascontiguousarray(model.T2, np.int32), px, py,
            len(reps), reps.ctypes.data, C.byref(t), len(rows), rows)
    n = lib().or_rows_targets_search(*args, 0, None)
    out = np.zeros(n, np.int64)
    lib().or_rows_targets_search(*args, n, out.ctypes.data)
    return out


def hv_rows_search(model, px, py, reps, v, rows, vidx=None):
    """(H v)[rows] of a two-generator sector without full-space tables
    (or_hv_rows_search): representative_p + findstate per target.  reps:
    int64 array (may be a large memory view); v: the whole vector, or with
    vidx its entries at those sorted indices."""
    assert reps.dtype == np.int64 and reps.flags.c_contiguous
    rows = np.ascontiguousarray(rows, np.int64)
    v = np.ascontiguousarray(v, np.complex128)
    vidx = np.arange(len(v), dtype=np.int64) if vidx is None else np.ascontiguousarray(vidx, np.int64)
    t, keep = make_terms(model)
    w = np.zeros(len(rows), np.complex128)
    lib().or_hv_rows_search(model.n_sites, np.ascontiguousarray(model.T1, np.int32),
                            np.ascontiguousarray(model.T2, np.int32), px, py, len(reps), reps.ctypes.data, C.byref(t),
                            vidx, v.view(np.float64), len(vidx), len(rows), rows, w.view(np.float64))
    return w


def representative(model, s, px, py):
    """or_representative (representative_p, Momentum_Basis.cpp:472-543):
    (rep, lx, ly, p_id, lid)."""
    L1, L2 = model.orders()
    lx, ly, pid, lid = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
    r = lib().or_representative(int(s), np.ascontiguousarray(model.T1, np.int32),
                                np.ascontiguousarray(model.T2, np.int32), model.n_sites, L1, L2, px, py,
                                C.byref(lx), C.byref(ly), C.byref(pid), C.byref(lid))
    return int(r), lx.value, ly.value, pid.value, lid.value


def cexp(re, im):
    """std::exp of complex(re, im) as the reference computes state_ph (glibc cexp)."""
    out = np.zeros(2)
    lib().or_cexp(re, im, out)
    return complex(out[0], out[1])


def density_matrix(st1, c1, st2, c2, cut, n_cut_ones=None):
    """density_matrixij / density_matrixij_Sz (analysis/src/Entropy.cpp:170-265)
    of two unwrapped vectors (states sorted, complex coefficients)."""
    st1 = np.ascontiguousarray(st1, np.int64)
    st2 = np.ascontiguousarray(st2, np.int64)
    c1 = np.ascontiguousarray(c1, np.complex128)
    c2 = np.ascontiguousarray(c2, np.complex128)
    cut = np.ascontiguousarray(cut, np.int32)
    nco = -1 if n_cut_ones is None else int(n_cut_ones)
    m = (1 << len(cut)) if nco < 0 else math.comb(len(cut), nco)
    dm = np.zeros(2 * m * m)
    lib().or_density_matrix(len(st1), st1, c1.view(np.float64), len(st2), st2, c2.view(np.float64), len(cut), cut,
                            nco, dm)
    return dm.view(np.complex128).reshape(m, m, order="F")
