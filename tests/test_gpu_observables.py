"""Two-site correlations of momentum-basis vectors (ed_pair_correlations)
against the same expectation values taken in the full Sz basis, as the
reference's analysis programs take them (analysis/src/pspincorr.cpp:80-145,
szsz_op / sxsxsysy_op at analysis/src/chirality_exp.cpp:232-289), on vectors
mapped to the full basis by an independent projector."""
import numpy as np
import pytest

import helpers as Hp
import pyoracle as P
from ed_code_amd import lattices as L

pytestmark = pytest.mark.gpu


def _ed():
    import ed_code_amd as ed
    ed.init(0)
    return ed


def _full_ops(st, i, j):
    """Sz_i Sz_j (diagonal) and (S+_i S-_j + S-_i S+_j)/2 on the state list."""
    idx = {s: a for a, s in enumerate(st)}
    zz = np.array([((s >> i) & 1) - 0.5 for s in st]) * np.array([((s >> j) & 1) - 0.5 for s in st])
    pm = np.zeros((len(st), len(st)))
    for a, s in enumerate(st):
        if ((s >> i) & 1) != ((s >> j) & 1):
            pm[idx[s ^ (1 << i) ^ (1 << j)], a] = 0.5
    return np.diag(zz), pm


CASES = [(L.square(4, 4, J2=0.4), 8, (1, 2), [(0, 1), (0, 5), (0, 10), (3, 12)]),
         (L.triangular(4, 3, Ch=0.2), 6, (0, 0), [(0, 1), (2, 7), (0, 6)]),
         (L.chain(12), 6, (3, 0), [(0, 1), (0, 6), (5, 2)])]


@pytest.mark.parametrize("model,n,k,pairs", CASES, ids=[f"{c[0].name}_k{c[2][0]}{c[2][1]}" for c in CASES])
def test_pair_correlations_vs_full_basis(gpu_ready, model, n, k, pairs):
    import torch
    import ed_code_amd as ed
    print(f"[case] correlations {model.name} k={k}", flush=True)
    o = P.OracleBasis(model, n, *k)
    basis = ed.Basis(ed.Lattice.of(model), n, *k)
    H, st = Hp.full_space_h(model, n)
    Pm = Hp.momentum_projector(model, n, *k, st)
    assert Pm.shape[1] == o.D == basis.D
    # the projector's convention is the library's: an eigenvector of the sector
    # maps to an eigenvector of the full H
    ev, vec = np.linalg.eigh(o.make_h())
    psi = Pm @ vec[:, 0]
    assert np.linalg.norm(H @ psi - ev[0] * psi) < 1e-9
    rng = np.random.default_rng(3)
    bra = rng.normal(size=o.D) + 1j * rng.normal(size=o.D)
    kets = [vec[:, 0], vec[:, 1] if o.D > 1 else vec[:, 0], rng.normal(size=o.D) + 1j * rng.normal(size=o.D)]
    for ket, b in ((kets[0], kets[0]), (kets[1], kets[0]), (kets[2], bra)):
        tk = torch.tensor(ket, dtype=torch.complex128, device="cuda")
        tb = torch.tensor(b, dtype=torch.complex128, device="cuda")
        zz, pm = basis.pair_correlations(pairs, tk, tb)
        for p, (i, j) in enumerate(pairs):
            Zf, Pf = _full_ops(st, i, j)
            fk, fb = Pm @ ket, Pm @ b
            ref_zz, ref_pm = np.vdot(fb, Zf @ fk), np.vdot(fb, Pf @ fk)
            assert abs(zz[p] - ref_zz) < 1e-12 * max(1.0, np.linalg.norm(fk) * np.linalg.norm(fb)), (i, j)
            assert abs(pm[p] - ref_pm) < 1e-12 * max(1.0, np.linalg.norm(fk) * np.linalg.norm(fb)), (i, j)


def test_pair_correlations_sum_rules(gpu_ready):
    """Ground state of the 4x4 Heisenberg torus (Sz=0, k=0, a singlet):
    sum_{j != 0} <S_0.S_j> = -3/4 (total spin zero), and the energy is
    E0 = 16 (<S_0.S_1> + <S_0.S_4>) (16 bonds along x and 16 along y)."""
    import torch
    import ed_code_amd as ed
    model = L.square(4, 4)
    basis = ed.Basis(ed.Lattice.of(model), 8, 0, 0)
    op = ed.Operator(basis, model)
    ev, vecs, _ = op.lanczos(nev=1, iterations=200, ncycles=4, tol=1e-12, want_vectors=True)
    psi = torch.tensor(vecs[0], dtype=torch.complex128, device="cuda")
    pairs = [(0, j) for j in range(1, 16)]
    zz, pm = basis.pair_correlations(pairs, psi)
    sdots = zz.real + pm.real              # S_i.S_j = Sz Sz + (S+S- + S-S+)/2
    assert abs(sdots.sum() + 0.75) < 1e-9
    assert abs(16 * (sdots[0] + sdots[3]) - ev[0]) < 1e-9
    assert np.abs(zz.imag).max() < 1e-12 and np.abs(pm.imag).max() < 1e-12


def test_unwrap_to_files(gpu_ready, tmp_path):
    """Eigenvector files of a solved sector: the binary twins hold exactly the
    GPU unwrap and the momentum vector, the text files parse back to them."""
    import torch
    import ed_code_amd as ed
    from ed_code_amd import evfiles as E
    model = L.triangular(4, 3, Ch=0.2)
    basis = ed.Basis(ed.Lattice.of(model), 6, 1, 1)
    op = ed.Operator(basis, model)
    ev, vecs, _ = op.lanczos(nev=1, iterations=150, ncycles=4, tol=1e-12, want_vectors=True)
    psi = torch.tensor(vecs[0], dtype=torch.complex128, device="cuda")
    prefix = str(tmp_path / "tri")
    E.unwrap_to_files(basis, psi, prefix, 0, 0.0, (1, 1), 1.0, ev[0])
    st, co = basis.unwrap(psi)
    b = E.read_bin(prefix + "_EV0.bin")
    assert np.array_equal(b["states"], st) and np.array_equal(b["coeffs"], co)
    ts, tc = E.read_ev(prefix + "_EV0.txt")
    assert np.array_equal(ts, st) and np.abs(tc - co).max() < 1e-15
    assert abs(np.linalg.norm(co) - 1) < 1e-9
    pb = E.read_bin(prefix + "_pEV0.bin")
    reps, norm, _ = basis.export()
    assert np.array_equal(pb["states"], reps) and np.array_equal(pb["coeffs"], vecs[0])
    n_p, p, r, nm, c = E.read_pev(prefix + "_pEV0.txt")
    assert n_p == basis.D and np.array_equal(r, reps) and np.abs(c - vecs[0]).max() < 1e-15


@pytest.mark.parametrize("k", [(0, 0), (1, 1)])
def test_scalar_chirality_vs_full_basis(gpu_ready, k):
    """<bra| S_i.(S_j x S_k) |ket> of momentum-basis vectors (orbit-averaged
    operator, no unwrapping) against the spin-matrix operator in the full
    basis, on the ground state, an excited state and random vectors of a
    chiral triangular torus; the ground-state energy then splits into the bond
    and chiral parts: E0 = sum_b J (lam <SzSz> + <S+S-+h.c.>/2) - Ch sum_t chi_t."""
    import torch
    import ed_code_amd as ed
    model, n = L.triangular(4, 3, Ch=0.3), 6
    print(f"[case] chirality {model.name} k={k}", flush=True)
    o = P.OracleBasis(model, n, *k)
    basis = ed.Basis(ed.Lattice.of(model), n, *k)
    H, st = Hp.full_space_h(model, n)
    Pm = Hp.momentum_projector(model, n, *k, st)
    ev, vec = np.linalg.eigh(o.make_h())
    rng = np.random.default_rng(5)
    tris = [model.triangles[0], model.triangles[1], model.triangles[7]]
    X = [Hp.scalar_chirality_matrix(model.n_sites, st, t) for t in tris]
    r1 = rng.normal(size=o.D) + 1j * rng.normal(size=o.D)
    r2 = rng.normal(size=o.D) + 1j * rng.normal(size=o.D)
    for ket, bra in ((vec[:, 0], vec[:, 0]), (vec[:, 1], vec[:, 0]), (r1, r2)):
        tk = torch.tensor(ket, dtype=torch.complex128, device="cuda")
        tb = torch.tensor(bra, dtype=torch.complex128, device="cuda")
        chi = basis.scalar_chirality(tris, tk, tb)
        fk, fb = Pm @ ket, Pm @ bra
        scale = max(1.0, np.linalg.norm(fk) * np.linalg.norm(fb))
        for q in range(len(tris)):
            assert abs(chi[q] - np.vdot(fb, X[q] @ fk)) < 1e-12 * scale, (k, q)
    # energy split of the ground state
    psi = torch.tensor(vec[:, 0], dtype=torch.complex128, device="cuda")
    pairs = [(x, y) for (x, y, _J, _ph) in model.bonds]
    zz, pm = basis.pair_correlations(pairs, psi)
    e_bonds = sum(J * (model.lam * zz[b].real + pm[b].real) for b, (_x, _y, J, _ph) in enumerate(model.bonds))
    chi = basis.scalar_chirality(model.triangles, psi)
    assert abs(e_bonds - model.Ch * chi.real.sum() - ev[0]) < 1e-9
    assert np.abs(chi.imag).max() < 1e-12


DM_CASES = [("chain12_k3", L.chain(12, phases=[0.15] * 12), 6, (3, 0), [0, 1, 2, 3, 4, 5]),
            ("tri4x4_ch_k13", L.triangular(4, 4, Ch=0.2), 8, (1, 3), [0, 1, 4, 5, 9, 14]),
            ("sq4x4_j2_k12", L.square(4, 4, J2=0.5), 7, (1, 2), [3, 0, 7, 12, 10]),
            ("tilted10_k17", L.tilted_square((3, 1), (-1, 3)), 5, (1, 7), [0, 2, 4, 6, 8])]


@pytest.mark.parametrize("name,model,n,k,cut", DM_CASES, ids=[c[0] for c in DM_CASES])
def test_density_matrix_vs_oracle(gpu_ready, name, model, n, k, cut):
    """ed_density_matrix (unwrapped vectors laid out as one matrix per cut
    Sz, one zgemm per block) against the oracle's density_matrixij and
    density_matrixij_Sz of the oracle-unwrapped vectors, and the entropies
    against their formulas."""
    import torch
    ed = _ed()
    print(f"[case] density matrix {name}", flush=True)
    o = P.OracleBasis(model, n, *k)
    basis = ed.Basis(ed.Lattice.of(model), n, *k)
    rng = np.random.default_rng(7)
    v1 = rng.normal(size=o.D) + 1j * rng.normal(size=o.D)
    v2 = rng.normal(size=o.D) + 1j * rng.normal(size=o.D)
    v1 /= np.linalg.norm(v1)
    v2 /= np.linalg.norm(v2)
    s1, c1 = o.unwrap(v1)
    s2, c2 = o.unwrap(v2)
    ket, bra = (torch.tensor(x, dtype=torch.complex128, device="cuda") for x in (v1, v2))
    dm = basis.density_matrix(cut, ket, bra)
    ref = P.density_matrix(s1, c1, s2, c2, cut)
    assert np.abs(dm - ref).max() <= 1e-13
    for a in range(len(cut) + 1):
        blk = basis.density_matrix(cut, ket, bra, n_cut_ones=a)
        assert np.abs(blk - P.density_matrix(s1, c1, s2, c2, cut, n_cut_ones=a)).max() <= 1e-13
    rho = basis.density_matrix(cut, ket)
    assert abs(np.trace(rho) - 1.0) < 1e-12
    r2, vn = ed.entanglement_entropies(rho)
    w = np.linalg.eigvalsh(rho)
    w = w[np.abs(w) > 1e-8]
    assert abs(r2 + np.log2(np.sum(np.abs(rho) ** 2))) < 1e-12
    assert abs(vn + np.sum(w * np.log2(w))) < 1e-10


def test_ground_state_half_cut_entropy(gpu_ready):
    """The 16-site Heisenberg ring's ground state (k = 0): the reduced density
    matrix of half the ring from ed_density_matrix against the numpy partial
    trace of the unwrapped vector; entropies from both."""
    import torch
    ed = _ed()
    model = L.chain(16)
    o = P.OracleBasis(model, 8, 0, 0)
    basis = ed.Basis(ed.Lattice.of(model), 8, 0, 0)
    op = ed.Operator(basis, model)
    ev, vecs, _ = op.lanczos(nev=1, iterations=200, ncycles=1, tol=1e-13, mode=ed.ED_LZ_FULL, want_vectors=True)
    x = torch.tensor(vecs[0], dtype=torch.complex128, device="cuda")
    cut = list(range(8))
    rho = basis.density_matrix(cut, x)
    st, co = o.unwrap(vecs[0])
    full = np.zeros(1 << 16, complex)
    full[st] = co
    psi = full.reshape(1 << 8, 1 << 8).T          # index = cut bits + 2^8 * rest bits
    ref = psi @ psi.conj().T
    assert np.abs(rho - ref).max() < 1e-12
    r2, vn = ed.entanglement_entropies(rho)
    w = np.linalg.eigvalsh(ref)
    w = w[np.abs(w) > 1e-8]
    assert abs(vn + np.sum(w * np.log2(w))) < 1e-9
    assert 1.0 < vn < 3.0 and 0.5 < r2 < vn       # a critical chain: S2 < S1, of order log L
