"""MI355X (gfx950) matrix-free H*v in the momentum- and Sz-symmetrised basis.

Python mirror of the reference's sector interface (hiteshjc/ED_Code):

    Lattice(T1, T2)                  translation group (read_input_file T1=/T2=)
    Basis(lattice, n_ones, kx, ky)   pConstruct_States   (Momentum_Basis.cpp:645-778)
    Operator(basis, model)           terms of sHv        (H_v.cpp:192-357)
    Operator.hv(v, w)                sHv on device tensors or host arrays
    Operator.lanczos(...)            lanczos_evecs       (lanczos.cpp:133-349)
    sector_lanczos(...)              sector_lanczos      (Lanczos_sector.cpp:15-198)

All compute runs in ed_code_amd/libed_amd.so through its C ABI
(include/ed_amd.h); there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional

import numpy as np

from . import lattices
from ._lib import (ED_C64, ED_C128, ED_LZ_FULL, ED_LZ_GRAPH, ED_LZ_LEAN, ED_LZ_REFERENCE, ED_LZ_SELECTIVE, ED_LZ_THICK, EDError, LanczosParams, LanczosStats, Terms,
                   check, lib)
from .lattices import Model, n_ones_for_sz

__all__ = ["Lattice", "Basis", "Operator", "BlockOperator", "MultiBlockOperator", "sector_lanczos", "entanglement_entropies", "lattices", "Model",
           "EDError", "ED_C64", "ED_C128",
           "ED_LZ_FULL", "ED_LZ_REFERENCE", "ED_LZ_LEAN", "ED_LZ_SELECTIVE", "ED_LZ_THICK", "n_ones_for_sz"]

_initialised = False


def init(device: int = 0):
    """Selects the device.  torch's HIP runtime is started first when torch
    sees a GPU: the library and torch each load a HIP runtime, and torch's
    must initialise before the library places anything on the device."""
    global _initialised
    if not _initialised:
        try:
            import torch
            if torch.cuda.is_available():
                torch.cuda.init()
        except ImportError:
            pass
        check(lib().ed_init(device))
        _initialised = True


class Lattice:
    def __init__(self, T1, T2, T3=None):
        init()
        self.T1 = np.ascontiguousarray(T1, np.int32)
        self.T2 = np.ascontiguousarray(T2, np.int32)
        self.T3 = None if T3 is None else np.ascontiguousarray(T3, np.int32)
        if self.T1.ndim != 1 or self.T1.shape != self.T2.shape or (self.T3 is not None and self.T3.shape != self.T1.shape):
            # read_input_file.cpp:423-427 rejects T1/T2 of different lengths
            raise EDError(1, "T1 and T2 must have the same length")
        self.N = len(self.T1)
        h = C.c_void_p()
        check(lib().ed_lattice_create3(self.N, self.T1.ctypes.data, self.T2.ctypes.data,
                                       None if self.T3 is None else self.T3.ctypes.data, C.byref(h)))
        self.h = h
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().ed_lattice_orders(h, C.byref(a), C.byref(b)))
        check(lib().ed_lattice_order3(h, C.byref(c)))
        self.L1, self.L2, self.L3 = a.value, b.value, c.value

    @classmethod
    def of(cls, model: Model) -> "Lattice":
        return cls(model.T1, model.T2, getattr(model, "T3", None))

    def __del__(self):
        if getattr(self, "h", None) and lib is not None:   # module globals are gone at interpreter exit
            lib().ed_lattice_destroy(self.h)
            self.h = None


class Basis:
    """Momentum basis of (n_ones, kx, ky), resident on the GPU."""

    def __init__(self, lattice: Lattice, n_ones: int, kx: int = 0, ky: int = 0, kz: int = 0):
        self.lattice = lattice
        self.n_ones, self.kx, self.ky, self.kz = n_ones, kx, ky, kz
        h = C.c_void_p()
        check(lib().ed_basis_build3(lattice.h, n_ones, kx, ky, kz, C.byref(h)))
        self.h = h
        d = C.c_int64()
        check(lib().ed_basis_dim(h, C.byref(d)))
        self.D = d.value

    def momenta(self):
        """(px, py) = 2 pi (kx/L1, ky/L2) as Lanczos_sector.cpp:57 forms them;
        (px, py, pz) on three-generator lattices (:240)."""
        L = self.lattice
        px = float(2 * self.kx) * math.pi / L.L1
        py = float(2 * self.ky) * math.pi / L.L2
        if L.T3 is None:
            return px, py
        return px, py, float(2 * self.kz) * math.pi / L.L3

    def export(self):
        """(pblock_states, norm, norm_inv) as host arrays."""
        D = self.D
        reps = np.zeros(D, np.int64)
        norm = np.zeros(D)
        norm_inv = np.zeros(D)
        check(lib().ed_basis_export(self.h, reps.ctypes.data, norm.ctypes.data, norm_inv.ctypes.data))
        return reps, norm, norm_inv

    def export_full(self):
        """Full-space rep_loc, state_ph, Ia, Ib, bits_right (reference layout)."""
        Cn, nl = C.c_int64(), C.c_int64()
        check(lib().ed_basis_full_size(self.h, C.byref(Cn), C.byref(nl)))
        rep_loc = np.zeros(Cn.value, np.int64)
        sph = np.zeros(2 * Cn.value)
        Ia = np.zeros(nl.value, np.int64)
        Ib = np.zeros(nl.value, np.int64)
        br = C.c_int32()
        check(lib().ed_basis_export_full(self.h, rep_loc.ctypes.data, sph.ctypes.data, Ia.ctypes.data,
                                         Ib.ctypes.data, C.byref(br)))
        return dict(rep_loc=rep_loc, state_ph=sph.view(np.complex128), Ia=Ia, Ib=Ib, bits_right=br.value)

    def pair_correlations(self, pairs, ket, bra=None):
        """<bra| Sz_i Sz_j |ket> and <bra| (S+_i S-_j + S-_i S+_j)/2 |ket> for each
        site pair (i, j), from complex128 device vectors of this basis (bra
        defaults to ket), computed in the momentum basis (no unwrapping; the
        reference's analysis/src/pspincorr.cpp:80-145 unwraps).  Returns two
        complex arrays (zz, pm)."""
        import torch
        bra = ket if bra is None else bra
        for x in (ket, bra):
            if x.dtype != torch.complex128 or not x.is_cuda or x.numel() != self.D:
                raise ValueError("vectors must be complex128 CUDA tensors of length D")
        ij = np.ascontiguousarray(np.asarray(pairs, np.int32).reshape(-1, 2))
        zz = np.zeros(2 * len(ij))
        pm = np.zeros(2 * len(ij))
        torch.cuda.synchronize()
        check(lib().ed_pair_correlations(self.h, C.c_void_p(bra.data_ptr()), C.c_void_p(ket.data_ptr()), len(ij),
                                         ij.ctypes.data, zz.ctypes.data, pm.ctypes.data))
        return zz.view(np.complex128), pm.view(np.complex128)

    def triangle_orbit(self, tri):
        """Distinct translates g(i), g(j), g(k) of the oriented triangle tri
        under the lattice's translations (cyclic rotations of a triple are the
        same triangle)."""
        L = self.lattice
        gens = [L.T1, L.T2] + ([L.T3] if L.T3 is not None else [])
        seen, todo, out = set(), [tuple(int(s) for s in tri)], []
        while todo:
            t = todo.pop()
            key = min(t[r:] + t[:r] for r in range(3))
            if key in seen:
                continue
            seen.add(key)
            out.append(t)
            todo.extend(tuple(int(T[s]) for s in t) for T in gens)
        return out

    def scalar_chirality(self, triangles, ket, bra=None):
        """<bra| S_i . (S_j x S_k) |ket> for each oriented triangle (i, j, k),
        from complex128 device vectors of this basis (bra defaults to ket), in
        the momentum basis without unwrapping (the reference unwraps the
        vectors first: analysis/src/chirality_exp.cpp:22-72 applies one cyclic
        term per call).  Both vectors have the sector's momentum, so every
        translate of a triangle has the same matrix element: each value is
        that of the operator averaged over the triangle's translation orbit,
        one H*v per triangle of the operator made of the three cyclic chiral
        entries (read_input_file.cpp:262-273, phases 1) of every translate.
        Those entries of one triangle (i, j, k) with Ch = 1 form
        -S_i . (S_j x S_k) (bit 1 = spin up, the gather convention of
        H_v.cpp), so the operator takes Ch = -1 / orbit size; the chiral part
        of a Hamiltonian with coupling Ch is -Ch * sum of these values."""
        import torch
        bra = ket if bra is None else bra
        for x in (ket, bra):
            if x.dtype != torch.complex128 or not x.is_cuda or x.numel() != self.D:
                raise ValueError("vectors must be complex128 CUDA tensors of length D")
        L = self.lattice
        out = np.zeros(len(triangles), np.complex128)
        done = {}   # triangles of one orbit share the value
        for q, tri in enumerate(triangles):
            orb = self.triangle_orbit(tri)
            key = frozenset(min(t[r:] + t[:r] for r in range(3)) for t in orb)
            if key in done:
                out[q] = out[done[key]]
                continue
            done[key] = q
            model = Model(name="chirality", n_sites=L.N, T1=L.T1, T2=L.T2, T3=L.T3, lam=0.0, bonds=[],
                          Ch=-1.0 / len(orb), chiral=lattices.chiral_entries(orb, []))
            w = Operator(self, model).hv(ket)
            out[q] = complex(torch.vdot(bra, w).item())
        return out

    def unwrap(self, v):
        """Full-Sz-basis components of the momentum-basis vector v (complex128
        device tensor): (states int64, coefficients complex128), ascending by
        state, entries above 1e-8 (Unwrap_pState, Momentum_Basis.cpp:978-1037)."""
        import torch
        if v.dtype != torch.complex128 or not v.is_cuda or v.numel() != self.D:
            raise ValueError("v must be a complex128 CUDA tensor of length D")
        torch.cuda.synchronize()
        n = C.c_int64()
        check(lib().ed_unwrap(self.h, C.c_void_p(v.data_ptr()), 0, None, None, C.byref(n)))
        st = np.zeros(n.value, np.int64)
        co = np.zeros(2 * n.value)
        if n.value:
            check(lib().ed_unwrap(self.h, C.c_void_p(v.data_ptr()), n.value, st.ctypes.data, co.ctypes.data,
                                  C.byref(n)))
        return st, co.view(np.complex128)

    def density_matrix(self, cut, ket, bra=None, n_cut_ones=None):
        """Reduced density matrix of the cut (a list of sites) from complex128
        device vectors of this basis: dm[j, k] = sum_e ket(j, e) conj(bra(k, e))
        over the configurations e of the other sites (bra defaults to ket),
        computed on the GPU (density_matrixij, analysis/src/Entropy.cpp:170-205).
        j, k are configurations of the cut, bit l = site cut[l]; with
        n_cut_ones, the block of cut configurations with that many up spins,
        in ascending order (density_matrixij_Sz, :207-265)."""
        import torch
        bra = ket if bra is None else bra
        for x in (ket, bra):
            if x.dtype != torch.complex128 or not x.is_cuda or not x.is_contiguous() or x.numel() != self.D:
                raise ValueError("vectors must be contiguous complex128 CUDA tensors of length D")
        cut = np.ascontiguousarray(np.asarray(cut, np.int32).ravel())
        nco = -1 if n_cut_ones is None else int(n_cut_ones)
        m = (1 << len(cut)) if nco < 0 else math.comb(len(cut), nco)
        dm = np.zeros(2 * m * m)
        torch.cuda.synchronize()
        check(lib().ed_density_matrix(self.h, C.c_void_p(ket.data_ptr()), C.c_void_p(bra.data_ptr()), len(cut),
                                      cut.ctypes.data, nco, dm.ctypes.data))
        return dm.view(np.complex128).reshape(m, m, order="F")

    def __del__(self):
        if getattr(self, "h", None) and lib is not None:   # module globals are gone at interpreter exit
            lib().ed_basis_destroy(self.h)
            self.h = None


def entanglement_entropies(dm):
    """(Renyi-2, von Neumann) entropies in bits of a density matrix
    (Renyi_Entropy and Von_Neumann_Entropy, analysis/src/Entropy.cpp:307-355),
    the eigenvalues from rocSOLVER on the GPU."""
    a = np.asfortranarray(np.asarray(dm, np.complex128))
    if a.ndim != 2 or a.shape[0] != a.shape[1]:
        raise ValueError("dm must be a square matrix")
    r2, vn = C.c_double(), C.c_double()
    check(lib().ed_entanglement_entropies(a.ctypes.data, a.shape[0], C.byref(r2), C.byref(vn)))
    return r2.value, vn.value


def _terms(model: Model):
    a = model.term_arrays()
    keep = [np.ascontiguousarray(a[k]) for k in ("bond_xy", "bond_J", "bond_phase", "chiral_ijk", "chiral_phase",
                                                 "bowtie")]
    t = Terms(a["lam"], a["nbonds"], keep[0].ctypes.data, keep[1].ctypes.data, keep[2].ctypes.data, a["Ch"],
              a["nchiral"], keep[3].ctypes.data, keep[4].ctypes.data, a["ChCh"], a["nbowtie"], keep[5].ctypes.data)
    return t, keep


class Operator:
    """H restricted to one momentum sector; terms uploaded once."""

    def __init__(self, basis: Basis, model: Model):
        self.basis = basis
        self.model = model
        t, keep = _terms(model)
        h = C.c_void_p()
        check(lib().ed_op_create(basis.h, C.byref(t), C.byref(h)))
        self.h = h
        self.D = basis.D

    def hv(self, v, w=None, stream=None):
        """w = H v.  v, w: torch tensors on the GPU (complex64 or complex128)
        or host numpy complex128 arrays (copied through the device)."""
        if isinstance(v, np.ndarray):
            v = np.ascontiguousarray(v, np.complex128)
            if v.ndim != 1 or v.size != self.D:
                raise ValueError(f"v must hold D = {self.D} values, got shape {v.shape}")
            if w is None:
                w = np.zeros(self.D, np.complex128)
            elif (not isinstance(w, np.ndarray) or w.dtype != np.complex128 or w.ndim != 1 or w.size != self.D
                  or not w.flags.c_contiguous or not w.flags.writeable):
                raise ValueError("w must be a writeable C-contiguous complex128 array of length D")
            if self.D and np.shares_memory(v, w):
                raise ValueError("v and w must not alias")
            check(lib().ed_hv_host(self.h, v.ctypes.data, w.ctypes.data))
            return w
        dt = self._check_device(v, w)
        if w is None:
            w = v.new_empty(v.shape)
        import torch
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(lib().ed_hv(self.h, dt, C.c_void_p(v.data_ptr()), C.c_void_p(w.data_ptr()), C.c_void_p(st)))
        return w

    def _check_device(self, v, w):
        """dtype code of a device H*v: v (and w when given) contiguous CUDA
        tensors of D complex64 or complex128 values, w distinct from v."""
        import torch
        if v.dtype == torch.complex128:
            dt = ED_C128
        elif v.dtype == torch.complex64:
            dt = ED_C64
        else:
            raise TypeError(f"unsupported dtype {v.dtype}")
        if not v.is_cuda or not v.is_contiguous() or v.numel() != self.D:
            raise ValueError("v must be a contiguous device tensor of length D")
        if w is not None:
            if w.dtype != v.dtype or not w.is_cuda or not w.is_contiguous() or w.numel() != self.D:
                raise ValueError("w must match v")
            if self.D and w.data_ptr() == v.data_ptr():
                raise ValueError("v and w must not alias")
        return dt

    def tune(self, path: int = 0, block: int = 0):
        """Kernel choice: path 0 auto (tuned on the first large H*v), 1 generic
        walk, 2 compiled-shape rotation kernel (K=4), 3 run-time-shape
        rotation kernel, 4 compiled-shape K=2 at 8 waves, 5 compiled-shape K=8,
        6 two-row 64-bit kernel, 7 compiled-shape K=4 software-pipelined,
        8 two-row kernel gathering u by row through the directory, 9 / 10
        compiled-shape with every row split over two / four lanes, 11 as 9
        with K=4 per lane, 12 stored element list (built by the first H*v:
        each element's image element and family kept in HBM, then streamed;
        uniform real bond families at k = 0 only); block 0 = default."""
        check(lib().ed_op_tune(self.h, path, block))

    def elist_stats(self):
        """(elements, 4-byte slots) of the stored element list; (0, 0) before
        the first H*v on path 12 builds it."""
        e, s = C.c_int64(), C.c_int64()
        check(lib().ed_op_elist_stats(self.h, C.byref(e), C.byref(s)))
        return e.value, s.value

    def tuning(self, dtype: str = "c128"):
        """(path, variant, block) the next H*v of this precision runs."""
        p, v, b = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().ed_op_tuning(self.h, ED_C64 if dtype == "c64" else ED_C128, C.byref(p), C.byref(v), C.byref(b)))
        return p.value, v.value, b.value

    def hv_profile(self, v, w, stream=None) -> float:
        """ed_hv with HIP events around the row kernel; returns its ms."""
        import torch
        if w is None:
            raise ValueError("hv_profile needs an output tensor w")
        dt = self._check_device(v, w)
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        ms = C.c_float()
        check(lib().ed_hv_profile(self.h, dt, C.c_void_p(v.data_ptr()), C.c_void_p(w.data_ptr()), C.c_void_p(st),
                                  C.byref(ms)))
        return ms.value

    def count_targets(self):
        g, t = C.c_int64(), C.c_int64()
        check(lib().ed_op_count_targets(self.h, C.byref(g), C.byref(t)))
        return g.value, t.value

    def make_h(self) -> np.ndarray:
        H = np.zeros(self.D * self.D, np.complex128)
        check(lib().ed_make_h(self.h, H.ctypes.data))
        return H.reshape(self.D, self.D, order="F")

    def dense_eigh(self, want_vectors=False):
        """All eigenvalues (and eigenvectors) of the sector matrix, on the GPU."""
        ev = np.zeros(self.D)
        vecs = np.zeros(self.D * self.D, np.complex128) if want_vectors else None
        check(lib().ed_dense_eigh(self.h, ev.ctypes.data, None if vecs is None else vecs.ctypes.data))
        return (ev, vecs.reshape(self.D, self.D, order="F")) if want_vectors else ev

    def lanczos(self, nev=1, iterations=100, ncycles=10, tol=1e-11, seed=1234, precision=ED_C128,
                mode=ED_LZ_REFERENCE, want_vectors=False, selective=False, graph=False):
        flags = (ED_LZ_SELECTIVE if selective else 0) | (ED_LZ_GRAPH if graph else 0)
        p = LanczosParams(nev, iterations, ncycles, precision, mode, flags, tol, seed)
        ev = np.zeros(nev)
        vecs = np.zeros((nev, self.D), np.complex128) if want_vectors else None
        st = LanczosStats()
        check(lib().ed_lanczos(self.h, C.byref(p), ev.ctypes.data, None if vecs is None else vecs.ctypes.data,
                               C.byref(st)))
        n = min(nev, self.D)
        stats = dict(hv_calls=st.hv_calls, hv_seconds=st.hv_seconds, total_seconds=st.total_seconds,
                     residual=list(st.residual[:min(n, 16)]), rayleigh=list(st.rayleigh[:min(n, 16)]),
                     eval_error=list(st.eval_error[:min(n, 64)]))
        return (ev[:n], vecs, stats) if want_vectors else (ev[:n], stats)

    def __del__(self):
        if getattr(self, "h", None) and lib is not None:   # module globals are gone at interpreter exit
            lib().ed_op_destroy(self.h)
            self.h = None


class BlockOperator:
    """M coupling instances (models with the same bond / triangle lists,
    different couplings) on one sector: W[:, m] = H_m V[:, m] in one pass.
    Vectors are (D, M) row-major torch tensors on the GPU (element (i, m) at
    i*M + m)."""

    def __init__(self, basis: Basis, models):
        self.basis = basis
        self.models = list(models)
        self.M = len(self.models)
        ts, self._keep = [], []
        for m in self.models:
            t, keep = _terms(m)
            ts.append(t)
            self._keep.append(keep)
        arr = (Terms * self.M)(*ts)
        h = C.c_void_p()
        check(lib().ed_block_create(basis.h, self.M, arr, C.byref(h)))
        self.h = h
        self.D = basis.D

    @property
    def fast(self) -> bool:
        return bool(lib().ed_block_fast(self.h))

    def hv(self, V, W=None, stream=None):
        import torch
        if V.dtype == torch.complex128:
            dt = ED_C128
        elif V.dtype == torch.complex64:
            dt = ED_C64
        else:
            raise TypeError(f"unsupported dtype {V.dtype}")
        if not V.is_cuda or not V.is_contiguous() or tuple(V.shape) != (self.D, self.M):
            raise ValueError("V must be a contiguous (D, M) device tensor")
        if W is None:
            W = torch.empty_like(V)
        if W.dtype != V.dtype or not W.is_contiguous() or W.shape != V.shape:
            raise ValueError("W must match V")
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(lib().ed_block_hv(self.h, dt, C.c_void_p(V.data_ptr()), C.c_void_p(W.data_ptr()), C.c_void_p(st)))
        return W

    def lanczos(self, nev=1, iterations=200, tol=1e-10, seed=1234, precision=ED_C64, mode=ED_LZ_FULL, ncycles=1):
        """Lowest nev eigenvalues of every instance: array (M, nev), stats.
        mode ED_LZ_FULL: one run per eigenvalue with deflation (iterations
        steps each); ED_LZ_THICK: thick restart with iterations Krylov
        vectors and at most ncycles restarts."""
        p = LanczosParams(nev, iterations, ncycles, precision, mode, 0, tol, seed)
        n = min(nev, self.D)
        ev = np.zeros((self.M, n))
        st = LanczosStats()
        check(lib().ed_block_lanczos(self.h, C.byref(p), ev.ctypes.data, C.byref(st)))
        err = np.zeros(self.M * n)
        k = min(self.M * n, 64)
        err[:k] = st.eval_error[:k]
        stats = dict(hv_calls=st.hv_calls, hv_seconds=st.hv_seconds, total_seconds=st.total_seconds,
                     residual=list(st.residual[:min(self.M, 16)]), rayleigh=list(st.rayleigh[:min(self.M, 16)]),
                     eval_error=err.reshape(self.M, n))
        return ev, stats

    def __del__(self):
        if getattr(self, "h", None) and lib is not None:   # module globals are gone at interpreter exit
            lib().ed_block_destroy(self.h)
            self.h = None


class MultiBlockOperator:
    """M coupling instances (models with one term structure) on each of
    several momentum sectors (bases of one lattice and one n_ones), applied
    in one launch over all sectors' row tiles.  Vectors are (rows, M)
    row-major torch tensors on the GPU: the sectors' (D_s, M) blocks
    stacked in sector order (rows = sum of D_s); sector s occupies rows
    [row0[s], row0[s] + D_s)."""

    def __init__(self, bases, models):
        self.bases = list(bases)
        self.models = list(models)
        self.M = len(self.models)
        ts, self._keep = [], []
        for m in self.models:
            t, keep = _terms(m)
            ts.append(t)
            self._keep.append(keep)
        arr = (Terms * self.M)(*ts)
        hb = (C.c_void_p * len(self.bases))(*[b.h.value for b in self.bases])
        h = C.c_void_p()
        check(lib().ed_mblock_create(hb, len(self.bases), self.M, arr, C.byref(h)))
        self.h = h
        self.row0 = []
        for s in range(len(self.bases) + 1):
            r0, d = C.c_int64(), C.c_int64()
            check(lib().ed_mblock_rows(h, s, C.byref(r0), C.byref(d)))
            self.row0.append(r0.value if s < len(self.bases) else d.value)
        self.rows = self.row0[-1]

    @property
    def fast(self) -> bool:
        return bool(lib().ed_mblock_fast(self.h))

    def split(self, V):
        """The per-sector (D_s, M) views of a stacked vector."""
        return [V[self.row0[s]:self.row0[s + 1]] for s in range(len(self.bases))]

    def hv(self, V, W=None, stream=None):
        import torch
        if V.dtype == torch.complex128:
            dt = ED_C128
        elif V.dtype == torch.complex64:
            dt = ED_C64
        else:
            raise TypeError(f"unsupported dtype {V.dtype}")
        if not V.is_cuda or not V.is_contiguous() or tuple(V.shape) != (self.rows, self.M):
            raise ValueError("V must be a contiguous (rows, M) device tensor")
        if W is None:
            W = torch.empty_like(V)
        if W.dtype != V.dtype or not W.is_contiguous() or W.shape != V.shape or W.data_ptr() == V.data_ptr():
            raise ValueError("W must match V and not alias it")
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(lib().ed_mblock_hv(self.h, dt, C.c_void_p(V.data_ptr()), C.c_void_p(W.data_ptr()), C.c_void_p(st)))
        return W

    def lanczos(self, nev=1, iterations=200, tol=1e-10, seed=1234, precision=ED_C64):
        """Lowest nev eigenvalues of every (sector, instance), all advanced
        together: arrays (nsec, M, nev) of eigenvalues and of their FP64
        error bounds, and stats."""
        S = len(self.bases)
        n = min(nev, min(b.D for b in self.bases))
        p = LanczosParams(nev, iterations, 1, precision, ED_LZ_FULL, 0, tol, seed)
        ev = np.zeros((S, self.M, n))
        err = np.zeros((S, self.M, n))
        st = LanczosStats()
        check(lib().ed_mblock_lanczos(self.h, C.byref(p), ev.ctypes.data, err.ctypes.data, C.byref(st)))
        stats = dict(hv_calls=st.hv_calls, hv_seconds=st.hv_seconds, total_seconds=st.total_seconds,
                     eval_error=err)
        return ev, stats

    def __del__(self):
        if getattr(self, "h", None) and lib is not None:   # module globals are gone at interpreter exit
            lib().ed_mblock_destroy(self.h)
            self.h = None


def sector_lanczos(model: Model, kx: int, ky: int, cSz: float, n_sec_ev: int = 2, tol: float = 1e-11,
                   iterations: int = 100, ncycles: int = 10, seed: int = 1234, precision: int = ED_C128,
                   dense_limit: int = 10000):
    """Lanczos_sector.cpp:15-198: n_ones from cSz (integer n_spins/2), k
    check, basis, dense diagonalisation for D <= 10000 (all D eigenvalues,
    exact_evecs, lanczos.cpp:81-131) else Lanczos with the reference
    schedule.  Returns (energies, D)."""
    n_ones = n_ones_for_sz(model.n_sites, cSz)
    lat = Lattice.of(model)
    if kx >= lat.L1 or ky >= lat.L2 or kx < 0 or ky < 0:
        raise EDError(1, "Invalid kx or ky")
    basis = Basis(lat, n_ones, kx, ky)
    if basis.D == 0:
        raise EDError(4, "No states in psector")
    op = Operator(basis, model)
    if basis.D <= dense_limit:
        return op.dense_eigh(), basis.D
    ev, _ = op.lanczos(nev=n_sec_ev, iterations=iterations, ncycles=ncycles, tol=tol, seed=seed,
                       precision=precision, mode=ED_LZ_REFERENCE)
    return ev, basis.D
