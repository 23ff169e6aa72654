"""The benchmark-shape kernels at full size against the reference and the
oracle.

The sectors of the headline and batch workloads (32-site 4x8 square torus,
30-site 6x5 triangular torus with chirality at k = 0 and k != 0, 27-site 3x9
triangular torus, 40-site tilted cluster) run the compiled-shape kernels
TorusC<32,4,8>, TorusC<30,6,5>, TorusC<27,3,9> and Rows2C<40,20>, which the
small parity cases never reach.  Bar: G (distinct targets) exact, basis
tables bit-exact, complex128 H*v within 1e-12 and complex64 within 1e-6
relative (vector 2-norm).

Where the reference's full-space tables fit in host memory (C <= 6.0e8
states: 14.4 GB), the product exports them and the reference's own sHv
(H_v.cpp:192-357, compiled from its sources into oracle/_ref) runs over every
row; the exported tables are pinned by the oracle's representative_p on
sampled states.  The 40-site sector (C = 1.4e11) is checked on sampled rows
with the oracle's table-free row formula."""
import ctypes as C
import math

import numpy as np
import pytest

import helpers as Hp
import pyoracle as P
from ed_code_amd import lattices as L

pytestmark = pytest.mark.gpu
needs_ref = pytest.mark.skipif(not P.ref_available(), reason="reference slice (oracle/_ref) not built")


def _ed():
    import ed_code_amd as ed
    ed.init(0)
    return ed


def _unrank(r, N, n):
    """State of combinatorial-number-system rank r among weight-n words."""
    s = 0
    for t in range(n, 0, -1):
        lo, hi = t - 1, N - 1            # largest p with C(p, t) <= r
        while lo < hi:
            mid = (lo + hi + 1) // 2
            if math.comb(mid, t) <= r:
                lo = mid
            else:
                hi = mid - 1
        s |= 1 << lo
        r -= math.comb(lo, t)
    return s


def _view(ptr, ctype, n):
    return np.ctypeslib.as_array((ctype * n).from_address(ptr))


class Exported:
    """pConstruct_States' tables exported by the GPU into the reference
    harness (RefSHv) and copied into the oracle (OracleTables)."""

    def __init__(self, basis, model, n, kx, ky):
        from ed_code_amd._lib import check, lib
        Cn, nl, br = C.c_int64(), C.c_int64(), C.c_int32()
        check(lib().ed_basis_full_size(basis.h, C.byref(Cn), C.byref(nl)))
        check(lib().ed_basis_export_full(basis.h, None, None, None, None, C.byref(br)))
        self.C, self.D, self.nl = Cn.value, basis.D, nl.value
        self.px, self.py = model.momentum(kx, ky)
        self.ref = P.RefSHv(model, self.px, self.py, self.C, self.D, self.nl, br.value)
        p = self.ref.ptrs
        check(lib().ed_basis_export(basis.h, p[0], p[1], p[2]))
        check(lib().ed_basis_export_full(basis.h, p[3], p[4], p[5], p[6], C.byref(br)))
        self.orc = P.OracleTables(model, n, self.px, self.py, self.C, self.D, self.nl, br.value)
        sizes = [self.D * 8] * 3 + [self.C * 8, self.C * 16, self.nl * 8, self.nl * 8]
        for dst, src, nb in zip(self.orc.ptrs, p, sizes):
            C.memmove(dst, src, nb)
        self.reps = _view(p[0], C.c_int64, self.D)
        self.norm = _view(p[1], C.c_double, self.D)
        self.rep_loc = _view(p[3], C.c_int64, self.C)
        self.state_ph = _view(p[4], C.c_double, 2 * self.C)


def _pin_sampled(model, n, E, nsample=3000, seed=0):
    """The exported tables against the oracle's representative_p
    (Momentum_Basis.cpp:472-543) on sampled states: rep_loc and state_ph of
    random ranks bit-exact, sampled representatives valid with their norm."""
    N = model.n_sites
    L1, L2 = model.orders()
    rng = np.random.default_rng(seed)
    assert np.all(np.diff(E.reps) > 0)
    for r in rng.integers(0, E.C, nsample):
        rep, lx, ly, pid, lid = P.representative(model, _unrank(int(r), N, n), E.px, E.py)
        j = int(np.searchsorted(E.reps, rep))
        want = j if (pid and j < E.D and E.reps[j] == rep) else -1
        assert E.rep_loc[r] == want, (r, rep)
        ph = P.cexp(0.0, -E.px * (lx % L1) - E.py * (ly % L2))
        assert (E.state_ph[2 * r], E.state_ph[2 * r + 1]) == (ph.real, ph.imag), r
    for i in rng.integers(0, E.D, nsample):
        rep, _, _, pid, lid = P.representative(model, int(E.reps[i]), E.px, E.py)
        assert rep == E.reps[i] and pid == 1
        assert E.norm[i] == math.sqrt(L1 * L2 // lid)


def _relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _gpu_hv(op, v, dtype):
    import torch
    vd = torch.from_numpy(v).to("cuda").to(dtype)
    return op.hv(vd).to(torch.complex128).cpu().numpy()


FULL = [("sq4x8_k00", L.square(4, 8, J1=1.0, J2=0.5), 16, (0, 0), 18784998),
        ("tri6x5_ch_k00", L.triangular(6, 5, Ch=0.15), 15, (0, 0), None),
        ("tri6x5_ch_k23", L.triangular(6, 5, Ch=0.2), 15, (2, 3), None)]


@needs_ref
@pytest.mark.parametrize("name,model,n,k,dim", FULL, ids=[c[0] for c in FULL])
def test_full_sector_vs_reference_shv(gpu_ready, name, model, n, k, dim):
    """Every row of the benchmark sector: the GPU's complex128 H*v (the
    autotuned compiled-shape kernel) against the reference's own sHv on the
    exported tables, complex64 likewise; G against the oracle's count on the
    same tables; the tables pinned by sampled representative_p."""
    import torch
    ed = _ed()
    print(f"[case] full {name}", flush=True)
    basis = ed.Basis(ed.Lattice.of(model), n, *k)
    if dim is not None:
        assert basis.D == dim
    op = ed.Operator(basis, model)
    E = Exported(basis, model, n, *k)
    _pin_sampled(model, n, E)
    print(f"[case] {name} tables pinned, D={basis.D}", flush=True)
    assert op.count_targets() == E.orc.count_targets()
    v = P.start_vector(17, basis.D)
    wr = E.ref.hv_rows(v, basis.D)
    print(f"[case] {name} reference sHv done", flush=True)
    w128 = _gpu_hv(op, v, torch.complex128)
    assert op.tuning("c128")[0] == 2            # the compiled-shape kernel ran
    assert _relerr(w128, wr) <= 1e-12
    assert _relerr(_gpu_hv(op, v, torch.complex64), wr) <= 1e-6
    # every compiled variant against the reference; the one-lane variants
    # bit-identical to each other, the split ones (rows over two and four
    # lanes) bit-identical per lane count
    out = {}
    for path in (2, 4, 5, 7, 9, 10, 11):
        op.tune(path)
        out[path] = _gpu_hv(op, v, torch.complex128)
        assert _relerr(out[path], wr) <= 1e-12, path
        assert _relerr(_gpu_hv(op, v, torch.complex64), wr) <= 1e-6, path
    for path in (4, 5, 7):
        assert np.array_equal(out[path], out[2]), path
    assert np.array_equal(out[11], out[9])
    for b in (256, 512):
        for path in (9, 10, 11):
            op.tune(path, b)
            assert np.array_equal(_gpu_hv(op, v, torch.complex128), out[path]), (path, b)
    if name == "sq4x8_k00":
        # the stored element list: the same elements, order and FMAs as path 2
        op.tune(12)
        assert np.array_equal(_gpu_hv(op, v, torch.complex128), out[2])
        ne, slots = op.elist_stats()
        assert ne >= op.count_targets()[0] and slots >= ne


@needs_ref
def test_block_kernel_tri6x5_vs_reference_shv(gpu_ready):
    """The M=4 block kernel on the 30-site triangular torus at k = (1,2) with
    four chiral couplings against the reference's sHv of each instance."""
    import torch
    ed = _ed()
    print("[case] block tri6x5 M=4 k=(1,2)", flush=True)
    chs = [0.05, 0.12, 0.21, 0.3]
    models = [L.triangular(6, 5, Ch=c) for c in chs]
    basis = ed.Basis(ed.Lattice.of(models[0]), 15, 1, 2)
    blk = ed.BlockOperator(basis, models)
    assert blk.fast
    E = Exported(basis, models[0], 15, 1, 2)
    V = np.stack([P.start_vector(30 + m, basis.D) for m in range(4)], axis=1)
    W = blk.hv(torch.from_numpy(np.ascontiguousarray(V)).to("cuda")).cpu().numpy()
    W64 = blk.hv(torch.from_numpy(np.ascontiguousarray(V)).to("cuda").to(torch.complex64))
    W64 = W64.to(torch.complex128).cpu().numpy()
    for m, mod in enumerate(models):
        E.ref.set_model(mod)
        wr = E.ref.hv_rows(np.ascontiguousarray(V[:, m]), basis.D)
        assert _relerr(W[:, m], wr) <= 1e-12, m
        assert _relerr(W64[:, m], wr) <= 1e-6, m


def test_block_kernel_tri6x5_m8_vs_single(gpu_ready):
    """M=8 on the 30-site triangular torus with chirality: the pair tables
    of eight instances and the rank tables take 78 KB of LDS, which one
    gfx950 workgroup may hold, so the shared kernel runs; each instance's
    column against the single-instance operator of its couplings."""
    import torch
    ed = _ed()
    print("[case] block tri6x5 M=8 k=(1,2)", flush=True)
    chs = [0.02, 0.05, 0.09, 0.12, 0.16, 0.21, 0.25, 0.3]
    models = [L.triangular(6, 5, Ch=c) for c in chs]
    basis = ed.Basis(ed.Lattice.of(models[0]), 15, 1, 2)
    blk = ed.BlockOperator(basis, models)
    assert blk.fast
    V = np.stack([P.start_vector(40 + m, basis.D) for m in range(8)], axis=1)
    Vd = torch.from_numpy(np.ascontiguousarray(V)).to("cuda")
    W = blk.hv(Vd).cpu().numpy()
    W64 = blk.hv(Vd.to(torch.complex64)).to(torch.complex128).cpu().numpy()
    for m, mod in enumerate(models):
        ws = _gpu_hv(ed.Operator(basis, mod), np.ascontiguousarray(V[:, m]), torch.complex128)
        assert _relerr(W[:, m], ws) <= 1e-12, m
        assert _relerr(W64[:, m], ws) <= 1e-6, m


@pytest.mark.parametrize("k", [(0, 0), (1, 4)], ids=["k00", "k14"])
def test_tri3x9_vs_oracle_full(gpu_ready, k):
    """27-site 3x9 triangular torus with chirality, "Sz=0" (n_ones = 13 as
    Lanczos_sector.cpp:45-46 computes it): the oracle's pConstruct_States
    in full, tables bit-exact, H*v and G."""
    import torch
    ed = _ed()
    model = L.triangular(3, 9, Ch=0.15)
    n = L.n_ones_for_sz(27, 0.0)
    assert n == 13
    print(f"[case] tri3x9 k={k}", flush=True)
    o = P.OracleBasis(model, n, *k)
    t = o.tables()
    basis = ed.Basis(ed.Lattice.of(model), n, *k)
    assert basis.D == o.D
    reps, norm, norm_inv = basis.export()
    assert np.array_equal(reps, t["reps"]) and np.array_equal(norm, t["norm"])
    assert np.array_equal(norm_inv, t["norm_inv"])
    full = basis.export_full()
    assert np.array_equal(full["rep_loc"], t["rep_loc"])
    assert np.array_equal(full["state_ph"].view(np.float64), t["state_ph"].view(np.float64))
    op = ed.Operator(basis, model)
    assert op.count_targets() == o.count_targets()
    v = P.start_vector(23, o.D)
    wo = o.hv(v)
    assert _relerr(_gpu_hv(op, v, torch.complex128), wo) <= 1e-12
    assert op.tuning("c128")[0] == 2
    assert _relerr(_gpu_hv(op, v, torch.complex64), wo) <= 1e-6


def test_tilted40_two_row_kernel_sampled_rows(gpu_ready):
    """The 40-site tilted cluster (periods (6,2), (-2,6)) relabelled to two
    rows of 20 sites, Sz=0, k=0, J2=0.5: the two-row 64-bit kernel
    (Rows2C<40,20>) with complex64 vectors on all D = 3,446,177,112 rows,
    checked on 256 sampled rows against the oracle's table-free row formula
    (representative_p + findstate per target, FP64) on the same complex64
    input; sampled representatives pinned by representative_p."""
    import torch
    ed = _ed()
    model, _ = L.regular_relabel(L.tilted_square((6, 2), (-2, 6), J1=1.0, J2=0.5))
    print("[case] tilted40 basis", flush=True)
    basis = ed.Basis(ed.Lattice.of(model), 20, 0, 0)
    assert basis.D == 3446177112
    reps = np.empty(basis.D, np.int64)
    from ed_code_amd._lib import check, lib
    check(lib().ed_basis_export(basis.h, reps.ctypes.data, None, None))
    print("[case] tilted40 reps exported", flush=True)
    rng = np.random.default_rng(4)
    for i in rng.integers(0, basis.D, 400):
        rep, _, _, pid, _ = P.representative(model, int(reps[i]), 0.0, 0.0)
        assert rep == reps[i] and pid == 1
        assert i == 0 or reps[i - 1] < reps[i]
    op = ed.Operator(basis, model)
    assert op.tuning("c64")[0] == 6
    g = torch.Generator(device="cuda").manual_seed(8)
    v = torch.randn(basis.D, dtype=torch.complex64, device="cuda", generator=g)
    w = op.hv(v)
    rows = np.unique(rng.integers(0, basis.D, 256))
    need = P.rows_targets_search(model, 0.0, 0.0, reps, rows)
    vn = v[torch.from_numpy(need).to("cuda")].to(torch.complex128).cpu().numpy()
    wg = w[torch.from_numpy(rows).to("cuda")].to(torch.complex128).cpu().numpy()
    del v, w
    torch.cuda.empty_cache()
    print(f"[case] tilted40 H*v done, {len(need)} entries of v needed", flush=True)
    wo = P.hv_rows_search(model, 0.0, 0.0, reps, vn, rows, vidx=need)
    assert np.isfinite(wo).all()
    assert _relerr(wg, wo) <= 1e-6
