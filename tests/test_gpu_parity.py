"""GPU parity: the HIP library (through its C ABI) against the oracle and
the reference's golden vectors.  Bar: basis tables bit-exact; H*v within
1e-12 relative (complex128) and 1e-6 relative (complex64); eigenvalues
within 1e-10 of dense diagonalisation."""
import math

import numpy as np
import pytest

import helpers as Hp
import pyoracle as P
from ed_code_amd import lattices as L

pytestmark = pytest.mark.gpu

GOLDEN = Hp.golden_files()


def _ed():
    import ed_code_amd as ed
    ed.init(0)
    return ed


def _hv(op, v, dtype):
    import torch
    vd = torch.from_numpy(np.ascontiguousarray(v)).to("cuda").to(dtype)
    return op.hv(vd).to(torch.complex128).cpu().numpy()


@pytest.mark.parametrize("path", GOLDEN, ids=[p.split("/")[-1][:-4] for p in GOLDEN])
def test_gpu_vs_golden(gpu_ready, path):
    import torch
    ed = _ed()
    m, d = Hp.load_golden(path)
    print(f"[case] {path}", flush=True)
    basis = ed.Basis(ed.Lattice.of(m), int(d["n_ones"]), int(d["kx"]), int(d["ky"]), int(d["kz"]))
    reps, norm, norm_inv = basis.export()
    assert np.array_equal(reps, d["reps"])
    assert np.array_equal(norm, d["norm"]) and np.array_equal(norm_inv, d["norm_inv"])
    full = basis.export_full()
    assert np.array_equal(full["rep_loc"], d["rep_loc"])
    bad = np.nonzero(full["state_ph"] != d["state_ph"])[0]
    if len(bad):
        st = Hp.states(m.n_sites, int(d["n_ones"]))
        for q in bad[:8]:
            print(f"[mismatch] rank {q} state {st[q]:#x} gpu {full['state_ph'][q].real.hex()} {full['state_ph'][q].imag.hex()}"
                  f" ref {d['state_ph'][q].real.hex()} {d['state_ph'][q].imag.hex()}", flush=True)
    assert len(bad) == 0, f"{len(bad)} state_ph entries differ"
    assert np.array_equal(full["Ia"], d["Ia"]) and np.array_equal(full["Ib"], d["Ib"])
    assert full["bits_right"] == int(d["bits_right"])
    op = ed.Operator(basis, m)
    w = _hv(op, d["v"], torch.complex128)
    assert np.linalg.norm(w - d["w"]) <= 1e-12 * np.linalg.norm(d["w"])
    w32 = _hv(op, d["v"], torch.complex64)
    assert np.linalg.norm(w32 - d["w"]) <= 1e-6 * np.linalg.norm(d["w"])
    wh = op.hv(np.ascontiguousarray(d["v"]))
    assert np.linalg.norm(wh - d["w"]) <= 1e-12 * np.linalg.norm(d["w"])
    H = op.make_h()
    assert np.abs(np.linalg.eigvalsh(H) - d["eigvals"]).max() < 1e-10
    # G, the distinct-target count of the roofline model, against the oracle
    o = P.OracleBasis(m, int(d["n_ones"]), int(d["kx"]), int(d["ky"]), int(d["kz"]))
    assert op.count_targets() == o.count_targets()


CASES = [
    ("chain20_k0", L.chain(20), 10, 0, 0),
    ("chain21_k4", L.chain(21, phases=[0.2] * 21), 10, 4, 0),
    ("sq4x4_j2_k12", L.square(4, 4, J2=0.5), 8, 1, 2),
    ("sq4x5_j2_k00", L.square(4, 5, J2=0.5), 10, 0, 0),
    ("tri4x4_ch_k13", L.triangular(4, 4, Ch=0.2), 8, 1, 3),
    ("tri3x6_chch_k02", L.triangular(3, 6, Ch=0.2, ChCh=0.1), 9, 0, 2),
    ("tilted10_k55", L.tilted_square((3, 1), (-1, 3), J2=0.5), 4, 5, 5),
    ("tilted20_k00", L.tilted_square((4, 2), (-2, 4), J2=0.5), 10, 0, 0),
]


@pytest.mark.parametrize("name,model,n,kx,ky", CASES, ids=[c[0] for c in CASES])
def test_gpu_vs_oracle(gpu_ready, name, model, n, kx, ky):
    import torch
    ed = _ed()
    print(f"[case] {name}", flush=True)
    o = P.OracleBasis(model, n, kx, ky)
    t = o.tables()
    basis = ed.Basis(ed.Lattice.of(model), n, kx, ky)
    assert basis.D == o.D
    reps, norm, _ = basis.export()
    assert np.array_equal(reps, t["reps"]) and np.array_equal(norm, t["norm"])
    full = basis.export_full()
    assert np.array_equal(full["rep_loc"], t["rep_loc"])
    assert np.array_equal(full["state_ph"].view(np.float64), t["state_ph"].view(np.float64))
    op = ed.Operator(basis, model)
    rng = np.random.default_rng(11)
    v = rng.standard_normal(o.D) + 1j * rng.standard_normal(o.D)
    wo = o.hv(v)
    assert np.linalg.norm(_hv(op, v, torch.complex128) - wo) <= 1e-12 * np.linalg.norm(wo)
    assert np.linalg.norm(_hv(op, v, torch.complex64) - wo) <= 1e-6 * np.linalg.norm(wo)
    assert op.count_targets() == o.count_targets()


def test_gpu_empty_sector_and_errors(gpu_ready):
    ed = _ed()
    m = L.tilted_square((3, 1), (-1, 3))
    lat = ed.Lattice.of(m)
    b = ed.Basis(lat, 5, 1, 1)
    assert b.D == 0
    op = ed.Operator(b, m)
    with pytest.raises(ed.EDError) as e:
        op.lanczos(nev=1)
    assert e.value.code == 4
    with pytest.raises(ed.EDError):
        ed.Basis(lat, 5, 10, 0)           # k out of range
    with pytest.raises(ed.EDError):
        ed.Basis(lat, 11, 0, 0)           # n_ones > N
    with pytest.raises(ed.EDError):
        ed.Lattice([1, 0, 2], [0, 1])     # length mismatch is caught by the ABI


@pytest.mark.parametrize("name,model,n,dim", [
    ("chain24", L.chain(24), 12, 112720),
    ("tri3x9", L.triangular(3, 9), 13, 742900),
])
def test_gpu_sector_dims_and_hermiticity(gpu_ready, name, model, n, dim):
    """Full-size sectors: Burnside dimensions, and Hermiticity / linearity of H."""
    import torch
    ed = _ed()
    print(f"[case] {name}", flush=True)
    basis = ed.Basis(ed.Lattice.of(model), n, 0, 0)
    assert basis.D == dim
    op = ed.Operator(basis, model)
    g = torch.Generator(device="cuda").manual_seed(5)
    x = torch.randn(basis.D, dtype=torch.complex128, device="cuda", generator=g)
    y = torch.randn(basis.D, dtype=torch.complex128, device="cuda", generator=g)
    hx, hy = op.hv(x), op.hv(y)
    lhs = torch.vdot(x, hy).item()
    rhs = torch.vdot(hx, y).item()
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
    h2 = op.hv(2.0 * x + 3.0 * y)
    assert torch.linalg.vector_norm(h2 - (2.0 * hx + 3.0 * hy)).item() < 1e-12 * torch.linalg.vector_norm(h2).item()


def test_gpu_chain24_matches_oracle_full(gpu_ready):
    import torch
    ed = _ed()
    model = L.chain(24)
    o = P.OracleBasis(model, 12, 0, 0)
    basis = ed.Basis(ed.Lattice.of(model), 12, 0, 0)
    reps, norm, _ = basis.export()
    t = o.tables(full=False)
    assert np.array_equal(reps, t["reps"]) and np.array_equal(norm, t["norm"])
    op = ed.Operator(basis, model)
    v = P.start_vector(3, o.D)
    wo = o.hv(v)
    assert np.linalg.norm(_hv(op, v, torch.complex128) - wo) <= 1e-12 * np.linalg.norm(wo)


def test_gpu_sq4x8_full_size(gpu_ready):
    """The benchmark sector at full size: dimension, target count, hermiticity."""
    import torch
    ed = _ed()
    model = L.square(4, 8, J1=1.0, J2=0.5)
    basis = ed.Basis(ed.Lattice.of(model), 16, 0, 0)
    assert basis.D == 18784998
    op = ed.Operator(basis, model)
    g = torch.Generator(device="cuda").manual_seed(9)
    x = torch.randn(basis.D, dtype=torch.complex128, device="cuda", generator=g)
    y = torch.randn(basis.D, dtype=torch.complex128, device="cuda", generator=g)
    lhs = torch.vdot(x, op.hv(y)).item()
    rhs = torch.vdot(op.hv(x), y).item()
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
    # single-precision path agrees with the double path on the same input
    w64 = op.hv(x)
    w32 = op.hv(x.to(torch.complex64)).to(torch.complex128)
    assert torch.linalg.vector_norm(w32 - w64).item() < 1e-6 * torch.linalg.vector_norm(w64).item()
    # the representatives ascend (every row against the reference's sHv:
    # test_gpu_fullsize.py)
    reps, _, _ = basis.export()
    assert np.all(np.diff(reps) > 0)


def _nonuniform_square():
    m = L.square(4, 4, J2=0.5)
    x, y, J, ph = m.bonds[3]
    m.bonds[3] = (x, y, 0.7, ph)          # one J1 bond differs: the per-member real class
    return m


FAST = [
    ("chain16_k5", L.chain(16, phases=[0.3] * 16), 8, 5, 0),
    ("sq4x4nu_k00", _nonuniform_square(), 8, 0, 0),
    ("sq4x4_k13", L.square(4, 4, J2=0.5, lam=0.7), 8, 1, 3),
    ("sq4x5_k00", L.square(4, 5, J2=0.5), 10, 0, 0),
    ("tri4x4ch_k12", L.triangular(4, 4, Ch=0.3), 8, 1, 2),
    ("tri6x3ch_k31", L.triangular(6, 3, Ch=0.2, J2=0.1), 9, 3, 1),
]


@pytest.mark.parametrize("name,model,n,kx,ky", FAST, ids=[c[0] for c in FAST])
def test_gpu_rotation_fast_path_vs_generic(gpu_ready, name, model, n, kx, ky):
    """Both kernel paths against the oracle, complex128 and complex64."""
    import torch
    ed = _ed()
    print(f"[case] {name}", flush=True)
    o = P.OracleBasis(model, n, kx, ky)
    basis = ed.Basis(ed.Lattice.of(model), n, kx, ky)
    op = ed.Operator(basis, model)
    rng = np.random.default_rng(12)
    v = rng.standard_normal(o.D) + 1j * rng.standard_normal(o.D)
    wo = o.hv(v)
    out = {}
    for path in (2, 4, 5, 7, 9, 10, 11, 3, 1):
        op.tune(path)
        out[path] = _hv(op, v, torch.complex128)
        assert np.linalg.norm(out[path] - wo) <= 1e-12 * np.linalg.norm(wo), path
        assert np.linalg.norm(_hv(op, v, torch.complex64) - wo) <= 1e-6 * np.linalg.norm(wo), path
    # the one-lane compiled variants accumulate in the same order
    assert np.array_equal(out[2], out[4]) and np.array_equal(out[2], out[5]) and np.array_equal(out[2], out[7])
    # the split variants (two and four lanes per row) sum each lane's share
    # separately: the same sums to rounding
    for path in (9, 10, 11):
        assert np.linalg.norm(out[path] - out[2]) <= 1e-14 * np.linalg.norm(out[2]), path
    # two lanes per row with four or eight elements in flight: the same order
    assert np.array_equal(out[9], out[11])


STORED = [
    ("sq4x4_k00", L.square(4, 4, J2=0.5, lam=0.7), 8),
    ("sq4x5_k00", L.square(4, 5, J2=0.5), 10),
    ("sq4x5_k00_n7", L.square(4, 5, J2=0.3, lam=1.3), 7),
    ("chain16_k00", L.chain(16, J=0.8), 8),
]


@pytest.mark.parametrize("name,model,n", STORED, ids=[c[0] for c in STORED])
def test_gpu_stored_element_list(gpu_ready, name, model, n):
    """Path 12 (the stored element list) against the oracle and bit-identical
    to path 2 (same elements, same order, same FMAs); the list holds one
    element per antiparallel bond of every representative, counted here from
    the oracle's representatives and the model's bonds."""
    import torch
    ed = _ed()
    print(f"[case] stored list {name}", flush=True)
    o = P.OracleBasis(model, n, 0, 0)
    basis = ed.Basis(ed.Lattice.of(model), n, 0, 0)
    op = ed.Operator(basis, model)
    rng = np.random.default_rng(21)
    v = rng.standard_normal(o.D) + 1j * rng.standard_normal(o.D)
    wo = o.hv(v)
    op.tune(2)
    w2 = _hv(op, v, torch.complex128)
    w2f = _hv(op, v, torch.complex64)
    op.tune(12)
    assert op.elist_stats() == (0, 0)
    w = _hv(op, v, torch.complex128)
    assert op.tuning("c128")[0] == 12 and op.tuning("c64")[0] == 12
    assert np.linalg.norm(w - wo) <= 1e-12 * np.linalg.norm(wo)
    assert np.array_equal(w, w2)
    assert np.array_equal(_hv(op, v, torch.complex64), w2f)
    reps = o.tables(full=False)["reps"]
    anti = np.zeros(o.D, np.int64)
    for x, y, J, _ in model.bonds:
        if abs(J) > 1e-16:
            anti += ((reps >> x) & 1) != ((reps >> y) & 1)
    ne, slots = op.elist_stats()
    assert ne == int(anti.sum())
    groups = [anti[g:g + 64].max() * 64 for g in range(0, o.D, 64)]
    assert slots == int(np.sum(groups))
    # a second H*v reuses the list
    assert np.array_equal(_hv(op, v, torch.complex128), w2)


def test_gpu_stored_element_list_refused(gpu_ready):
    """The stored list covers uniform real bond families at k = 0; complex
    coefficients (k != 0, chiral terms) are refused by the tune call."""
    ed = _ed()
    m = L.triangular(4, 4, Ch=0.3)
    for k in ((0, 0), (1, 2)):
        op = ed.Operator(ed.Basis(ed.Lattice.of(m), 8, *k), m)
        with pytest.raises(RuntimeError, match="stored element list"):
            op.tune(12)
    m = L.square(4, 4, J2=0.5)
    op = ed.Operator(ed.Basis(ed.Lattice.of(m), 8, 1, 3), m)
    with pytest.raises(RuntimeError, match="stored element list"):
        op.tune(12)


def test_gpu_autotune_picks_identical_result(gpu_ready):
    """On a sector above the tuning threshold (30-site triangular torus,
    D ~ 5.2e6) the automatic path times the variants on the first H*v; its
    result is bit-identical to each fixed variant, and the choice is reported."""
    import torch
    ed = _ed()
    m = L.triangular(6, 5, Ch=0.1)
    print("[case] tri6x5 autotune", flush=True)
    basis = ed.Basis(ed.Lattice.of(m), 15, 0, 0)
    assert basis.D > 1 << 20
    op = ed.Operator(basis, m)
    assert op.tuning("c64")[1] == -1
    g = torch.Generator(device="cuda").manual_seed(3)
    v = torch.randn(basis.D, dtype=torch.complex64, device="cuda", generator=g)
    w_auto = op.hv(v)
    path, var, block = op.tuning("c64")
    assert path == 2 and var in (0, 1, 2, 3, 4, 5, 6) and block in (256, 512)
    assert op.tuning("c128")[1] == -1          # tuned per precision
    # variants that sum a row in the same order as the pick: bit-identical;
    # the others (one lane per row against rows split over lanes): to rounding
    same = {0: (2, 4, 5, 7), 1: (2, 4, 5, 7), 2: (2, 4, 5, 7), 3: (2, 4, 5, 7), 4: (9, 11), 5: (10,), 6: (9, 11)}[var]
    for p in (2, 4, 5, 7, 9, 10, 11):
        for b in (256, 512):
            op.tune(p, b)
            w = op.hv(v)
            if p in same:
                assert torch.equal(w, w_auto), (p, b)
            else:
                assert torch.linalg.norm(w - w_auto) <= 1e-6 * torch.linalg.norm(w_auto), (p, b)
    op.tune(0, 256)                            # fixed block: only the variant is tuned
    w = op.hv(v)
    assert torch.linalg.norm(w - w_auto) <= 1e-6 * torch.linalg.norm(w_auto)
    assert op.tuning("c64")[2] == 256


def test_gpu_autotune_picks_kept_for_the_process(gpu_ready):
    """A second operator of the same shape, class and sector size takes the
    pick the first one measured (no timing: its first H*v is bit-identical
    to the first operator's); after ed_tune_cache_clear a new operator times
    the candidates again and keeps one of them."""
    import torch
    ed = _ed()
    from ed_code_amd._lib import lib
    m = L.triangular(6, 5, Ch=0.07)
    print("[case] tri6x5 autotune cache", flush=True)
    basis = ed.Basis(ed.Lattice.of(m), 15, 1, 0)
    assert basis.D > 1 << 20
    lib().ed_tune_cache_clear()
    g = torch.Generator(device="cuda").manual_seed(5)
    v = torch.randn(basis.D, dtype=torch.complex64, device="cuda", generator=g)
    a = ed.Operator(basis, m)
    wa = a.hv(v)
    pick = a.tuning("c64")
    assert pick[0] == 2 and pick[1] >= 0
    b = ed.Operator(basis, m)
    assert b.tuning("c64")[1] == -1
    wb = b.hv(v)
    assert b.tuning("c64") == pick
    assert torch.equal(wa, wb)
    lib().ed_tune_cache_clear()
    c = ed.Operator(basis, m)
    wc = c.hv(v)
    p2 = c.tuning("c64")
    assert p2[0] == 2 and p2[1] in (0, 1, 2, 3, 4, 5, 6) and p2[2] in (256, 512)
    assert torch.linalg.norm(wc - wa) <= 1e-6 * torch.linalg.norm(wa)


TWO_ROW = [("tilted20_k00", (4, 2), (-2, 4), 10, (0, 0)), ("tilted20_k71", (4, 2), (-2, 4), 10, (7, 1)),
           ("tilted20_n9_k31", (4, 2), (-2, 4), 9, (3, 1))]


@pytest.mark.parametrize("name,P1,P2,n,k", TWO_ROW, ids=[c[0] for c in TWO_ROW])
def test_gpu_two_row_64bit_path(gpu_ready, name, P1, P2, n, k):
    """Relabelled tilted 20-site cluster through the two-row 64-bit kernel
    (forced with path 6; the 32-bit rotation kernel also applies at N=20)
    against the oracle, both precisions."""
    import torch
    ed = _ed()
    m, _ = L.regular_relabel(L.tilted_square(P1, P2, J2=0.5, lam=0.9))
    print(f"[case] two-row {name}", flush=True)
    o = P.OracleBasis(m, n, *k)
    basis = ed.Basis(ed.Lattice.of(m), n, *k)
    assert basis.D == o.D
    op = ed.Operator(basis, m)
    op.tune(6)
    assert op.tuning("c128")[0] == 6
    rng = np.random.default_rng(21)
    v = rng.standard_normal(o.D) + 1j * rng.standard_normal(o.D)
    wo = o.hv(v)
    w6 = _hv(op, v, torch.complex128)
    assert np.linalg.norm(w6 - wo) <= 1e-12 * np.linalg.norm(wo)
    assert np.linalg.norm(_hv(op, v, torch.complex64) - wo) <= 1e-6 * np.linalg.norm(wo)
    # path 8: the same kernel gathering u by row through the directory (the
    # layout it falls back to when the wide gather image does not fit)
    op.tune(8)
    assert op.tuning("c128")[0] == 8
    w8 = _hv(op, v, torch.complex128)
    assert np.array_equal(w8, w6)
    assert np.linalg.norm(_hv(op, v, torch.complex64) - wo) <= 1e-6 * np.linalg.norm(wo)


CUBIC = [("cubic2x3x4_k123", L.cubic(2, 3, 4, J2=0.4, lam=0.8), 12, (1, 2, 3)),
         ("cubic3x3x2_k201", L.cubic(3, 3, 2), 9, (2, 0, 1)),
         ("cubic2x2x5_k015", L.cubic(2, 2, 5, J2=0.2), 10, (0, 1, 4))]


@pytest.mark.parametrize("name,model,n,k", CUBIC, ids=[c[0] for c in CUBIC])
def test_gpu_three_generators_vs_oracle(gpu_ready, name, model, n, k):
    """Three translation generators (Momentum_Basis.cpp:781-898, H_v.cpp:360):
    representatives, norms, full tables and H*v against the oracle."""
    import torch
    ed = _ed()
    print(f"[case] {name}", flush=True)
    o = P.OracleBasis(model, n, *k)
    t = o.tables()
    basis = ed.Basis(ed.Lattice.of(model), n, *k)
    assert basis.D == o.D > 0
    reps, norm, _ = basis.export()
    assert np.array_equal(reps, t["reps"]) and np.array_equal(norm, t["norm"])
    full = basis.export_full()
    assert np.array_equal(full["rep_loc"], t["rep_loc"])
    assert np.array_equal(full["state_ph"].view(np.float64), t["state_ph"].view(np.float64))
    op = ed.Operator(basis, model)
    rng = np.random.default_rng(12)
    v = rng.standard_normal(o.D) + 1j * rng.standard_normal(o.D)
    wo = o.hv(v)
    assert np.linalg.norm(_hv(op, v, torch.complex128) - wo) <= 1e-12 * np.linalg.norm(wo)
    assert np.linalg.norm(_hv(op, v, torch.complex64) - wo) <= 1e-6 * np.linalg.norm(wo)


def test_gpu_three_generators_full_size_sum_rule(gpu_ready):
    """4x4x2 cube, 32 sites, Sz=0: every orbit of size s lands in exactly s of
    the 32 momentum sectors, so the sector dimensions sum to C(32,16); H is
    Hermitian at k != 0."""
    import math
    import torch
    ed = _ed()
    model = L.cubic(4, 4, 2, J2=0.3)
    lat = ed.Lattice.of(model)
    total = 0
    for kx in range(4):
        for ky in range(4):
            for kz in range(2):
                total += ed.Basis(lat, 16, kx, ky, kz).D
    assert total == math.comb(32, 16)
    basis = ed.Basis(lat, 16, 1, 3, 1)
    op = ed.Operator(basis, model)
    g = torch.Generator(device="cuda").manual_seed(6)
    x = torch.randn(basis.D, dtype=torch.complex128, device="cuda", generator=g)
    y = torch.randn(basis.D, dtype=torch.complex128, device="cuda", generator=g)
    lhs, rhs = torch.vdot(x, op.hv(y)).item(), torch.vdot(op.hv(x), y).item()
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


@pytest.mark.parametrize("path", GOLDEN, ids=[p.split("/")[-1][:-4] for p in GOLDEN])
def test_gpu_unwrap_vs_golden(gpu_ready, path):
    """ed_unwrap against the reference's own Unwrap_pState output recorded in
    the fixture (states and coefficients bit for bit)."""
    import torch
    ed = _ed()
    m, d = Hp.load_golden(path)
    print(f"[case] unwrap {path.split('/')[-1]}", flush=True)
    basis = ed.Basis(ed.Lattice.of(m), int(d["n_ones"]), int(d["kx"]), int(d["ky"]), int(d["kz"]))
    v = torch.tensor(d["v"], dtype=torch.complex128, device="cuda")
    st, co = basis.unwrap(v)
    assert np.array_equal(st, d["unwrap_states"])
    assert np.array_equal(co.view(np.float64), d["unwrap_coef"].view(np.float64))


@pytest.mark.parametrize("name,model,n,k", [("chain20_k3", L.chain(20, phases=[0.1] * 20), 10, (3, 0)),
                                            ("tri4x4_k13", L.triangular(4, 4, Ch=0.2), 8, (1, 3)),
                                            ("tilted20_k71", L.tilted_square((4, 2), (-2, 4)), 10, (7, 1)),
                                            ("cubic2x2x4_k013", L.cubic(2, 2, 4), 8, (0, 1, 3))])
def test_gpu_unwrap_vs_oracle(gpu_ready, name, model, n, k):
    import torch
    ed = _ed()
    print(f"[case] unwrap {name}", flush=True)
    o = P.OracleBasis(model, n, *k)
    basis = ed.Basis(ed.Lattice.of(model), n, *k)
    rng = np.random.default_rng(21)
    v = rng.normal(size=o.D) + 1j * rng.normal(size=o.D)
    st, co = basis.unwrap(torch.tensor(v, dtype=torch.complex128, device="cuda"))
    so, cob = o.unwrap(v)
    assert np.array_equal(st, so)
    assert np.array_equal(co.view(np.float64), cob.view(np.float64))
