"""GPU Lanczos against dense diagonalisation, closed forms and the oracle.
Bar: complex128 Krylov storage within 1e-10 relative of the exact
eigenvalues (the drop-in contract), complex64 storage with the FP64
correction within 1e-9 relative (the precision target); every eigenvalue
within the FP64 error bound the solver reports for it."""
import math

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


def _bar(prec):
    return 1e-10 if prec == "c128" else 1e-9


def _within_bounds(lz, exact, st):
    """Each returned eigenvalue lies within its reported FP64 error bound of
    the exact one (the bound is the residual of its unit Ritz vector)."""
    err = np.abs(np.asarray(lz) - np.asarray(exact))
    bound = np.array(st["eval_error"][:len(err)])
    assert np.all(bound > 0)
    assert np.all(err <= bound + 1e-12), (err, bound)


@pytest.mark.parametrize("prec", ["c128", "c64"])
@pytest.mark.parametrize("mode", ["reference", "full"])
def test_lanczos_vs_dense(gpu_ready, prec, mode):
    ed = _ed()
    print(f"[case] lanczos {prec} {mode}", flush=True)
    model = L.square(4, 4, J2=0.5)
    o = P.OracleBasis(model, 8, 0, 0)
    ev = np.linalg.eigvalsh(o.make_h())
    basis = ed.Basis(ed.Lattice.of(model), 8, 0, 0)
    op = ed.Operator(basis, model)
    pr = ed.ED_C128 if prec == "c128" else ed.ED_C64
    md = ed.ED_LZ_REFERENCE if mode == "reference" else ed.ED_LZ_FULL
    lz, st = op.lanczos(nev=3, iterations=120, ncycles=6, tol=1e-12, seed=7, precision=pr, mode=md)
    assert np.abs(lz - ev[:3]).max() < _bar(prec) * np.abs(ev[:3]).max()
    assert st["residual"][0] < (1e-6 if prec == "c128" else 1e-3)
    _within_bounds(lz, ev[:3], st)


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_lanczos_selective_reorthogonalisation(gpu_ready, prec):
    """Full mode with the conditional second Gram-Schmidt pass: same
    eigenvalues as dense diagonalisation, fewer passes over the basis."""
    ed = _ed()
    model = L.triangular(4, 4, Ch=0.2, J2=0.3)
    o = P.OracleBasis(model, 8, 1, 2)
    ev = np.linalg.eigvalsh(o.make_h())
    basis = ed.Basis(ed.Lattice.of(model), 8, 1, 2)
    op = ed.Operator(basis, model)
    pr = ed.ED_C128 if prec == "c128" else ed.ED_C64
    lz, st = op.lanczos(nev=3, iterations=200, ncycles=1, tol=1e-13, seed=5, precision=pr, mode=ed.ED_LZ_FULL,
                        selective=True)
    assert np.abs(lz - ev[:3]).max() < _bar(prec) * np.abs(ev[:3]).max()
    assert st["residual"][0] < (1e-6 if prec == "c128" else 1e-3)
    _within_bounds(lz, ev[:3], st)


def test_lanczos_chain24_ground_state(gpu_ready):
    """chain24 Sz=0 k=0, 4 lowest deflated eigenvalues: c128 reference schedule
    against the oracle's Lanczos."""
    ed = _ed()
    model = L.chain(24)
    basis = ed.Basis(ed.Lattice.of(model), 12, 0, 0)
    op = ed.Operator(basis, model)
    lz, st = op.lanczos(nev=2, iterations=100, ncycles=4, tol=1e-12, seed=3)
    o = P.OracleBasis(model, 12, 0, 0)
    lo = o.lanczos(nev=2, iterations=100, ncycles=4, tol=1e-12, seed=3)
    assert np.abs(lz - lo).max() < 1e-10
    # Bethe-ansatz value for the 24-site Heisenberg ring: E0/N = -0.4438..., sanity band
    assert -0.45 * 24 < lz[0] < -0.44 * 24


def test_sector_lanczos_dense_and_lanczos_paths(gpu_ready):
    ed = _ed()
    # D <= 10000: dense path returns every eigenvalue
    m = L.chain(12)
    e, D = ed.sector_lanczos(m, 0, 0, 0.0)
    o = P.OracleBasis(m, 6, 0, 0)
    assert D == o.D and np.abs(e - np.linalg.eigvalsh(o.make_h())).max() < 1e-10
    # four-site ring E0 = -2
    e4, _ = ed.sector_lanczos(L.chain(4), 0, 0, 0.0)
    assert abs(e4[0] + 2.0) < 1e-12
    # odd N, Sz = 0 means n_ones = N//2
    assert ed.n_ones_for_sz(27, 0.0) == 13 and ed.n_ones_for_sz(27, 1.0) == 14


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_lean_lanczos_vs_dense(gpu_ready, prec):
    """Three-vector Lanczos without reorthogonalisation, Ritz vector from a
    second pass: lowest eigenvalue (FP64 Rayleigh quotient) against dense."""
    ed = _ed()
    model = L.triangular(4, 4, Ch=0.2, J2=0.3)
    o = P.OracleBasis(model, 8, 1, 2)
    e_dense = np.linalg.eigvalsh(o.make_h())[0]
    basis = ed.Basis(ed.Lattice.of(model), 8, 1, 2)
    op = ed.Operator(basis, model)
    pr = ed.ED_C128 if prec == "c128" else ed.ED_C64
    lz, st = op.lanczos(nev=1, iterations=300, ncycles=1, tol=1e-13, seed=9, precision=pr, mode=ed.ED_LZ_LEAN)
    assert abs(lz[0] - e_dense) < _bar(prec) * abs(e_dense)
    _within_bounds(lz, [e_dense], st)
    assert st["residual"][0] < (1e-6 if prec == "c128" else 1e-4)
    with pytest.raises(ed.EDError):
        op.lanczos(nev=2, iterations=50, mode=ed.ED_LZ_LEAN)


def test_lean_lanczos_two_row_path(gpu_ready):
    """Lean mode through the two-row 64-bit kernel (relabelled tilted 20-site
    cluster) against the full-reorthogonalisation solver on the same sector."""
    ed = _ed()
    m, _ = L.regular_relabel(L.tilted_square((4, 2), (-2, 4), J2=0.5))
    basis = ed.Basis(ed.Lattice.of(m), 10, 0, 0)
    op = ed.Operator(basis, m)
    op.tune(6)
    lean, st = op.lanczos(nev=1, iterations=400, ncycles=1, tol=1e-13, seed=2, precision=ed.ED_C64,
                          mode=ed.ED_LZ_LEAN)
    full, _ = op.lanczos(nev=1, iterations=300, ncycles=1, tol=1e-13, seed=2, precision=ed.ED_C128,
                         mode=ed.ED_LZ_FULL)
    assert abs(lean[0] - full[0]) < 1e-9 * abs(full[0])
    assert st["residual"][0] < 1e-4


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_graph_captured_lean_lanczos(gpu_ready, prec):
    """The lean solver with three steps captured as one hipGraph and replayed
    (device-side alpha/beta, no host round trips between checks): lowest
    eigenvalue against dense and against the eager lean solver; the graph is
    replayed many times in each run and the solve is run twice."""
    ed = _ed()
    model = L.triangular(4, 4, Ch=0.2, J2=0.3)
    o = P.OracleBasis(model, 8, 1, 2)
    e_dense = np.linalg.eigvalsh(o.make_h())[0]
    basis = ed.Basis(ed.Lattice.of(model), 8, 1, 2)
    op = ed.Operator(basis, model)
    pr = ed.ED_C128 if prec == "c128" else ed.ED_C64
    eager, _ = op.lanczos(nev=1, iterations=300, ncycles=1, tol=1e-13, seed=9, precision=pr, mode=ed.ED_LZ_LEAN)
    for _ in range(2):
        g, st = op.lanczos(nev=1, iterations=300, ncycles=1, tol=1e-13, seed=9, precision=pr, mode=ed.ED_LZ_LEAN,
                           graph=True)
        assert abs(g[0] - e_dense) < _bar(prec) * abs(e_dense)
        assert abs(g[0] - eager[0]) < _bar(prec) * abs(e_dense)
        assert st["residual"][0] < (1e-6 if prec == "c128" else 1e-4)
        assert st["hv_calls"] > 30


@pytest.mark.parametrize("selective", [False, True])
@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_graph_captured_full_lanczos(gpu_ready, prec, selective):
    """Full reorthogonalisation with one step captured as a hipGraph and
    replayed for every k (device-side step index, coefficients, alpha and
    beta; the selective second pass gated on the device): three eigenvalues
    of a complex sector against dense and against the eager solver, and two
    runs identical to the last bit (fixed-order reductions, replays that
    reuse one graph)."""
    ed = _ed()
    print(f"[case] graph full lanczos {prec} selective={selective}", flush=True)
    model = L.triangular(4, 4, Ch=0.2, J2=0.3)
    o = P.OracleBasis(model, 8, 1, 2)
    ev = np.linalg.eigvalsh(o.make_h())[:3]
    basis = ed.Basis(ed.Lattice.of(model), 8, 1, 2)
    op = ed.Operator(basis, model)
    pr = ed.ED_C128 if prec == "c128" else ed.ED_C64
    kw = dict(nev=3, iterations=150, ncycles=1, tol=1e-12, seed=5, precision=pr, mode=ed.ED_LZ_FULL,
              selective=selective)
    eager, _ = op.lanczos(**kw)
    runs = [op.lanczos(graph=True, **kw) for _ in range(2)]
    for g, st in runs:
        assert np.abs(g - ev).max() < _bar(prec) * np.abs(ev).max(), (g, ev)
        assert np.abs(g - eager).max() < _bar(prec) * np.abs(ev).max(), (g, eager)
        assert st["hv_calls"] > 30
        _within_bounds(g, ev, st)
    assert np.array_equal(runs[0][0], runs[1][0])


@pytest.mark.parametrize("prec", ["c128", "c64"])
@pytest.mark.parametrize("ncycles", [1, 6])
def test_graph_captured_reference_schedule(gpu_ready, prec, ncycles):
    """The reference schedule (restart cycles, deflation every step, the
    Gram-Schmidt pass in the last cycle, the jump to the last cycle once
    |E0 - E0_prev| < tol) with one step captured per kind of cycle and
    replayed: three eigenvalues of a complex sector against dense and the
    eager solver; two runs identical to the last bit."""
    ed = _ed()
    print(f"[case] graph reference lanczos {prec} ncycles={ncycles}", flush=True)
    model = L.triangular(4, 4, Ch=0.2, J2=0.3)
    o = P.OracleBasis(model, 8, 1, 2)
    ev = np.linalg.eigvalsh(o.make_h())[:3]
    basis = ed.Basis(ed.Lattice.of(model), 8, 1, 2)
    op = ed.Operator(basis, model)
    pr = ed.ED_C128 if prec == "c128" else ed.ED_C64
    kw = dict(nev=3, iterations=140 if ncycles == 1 else 60, ncycles=ncycles, tol=1e-12, seed=4, precision=pr,
              mode=ed.ED_LZ_REFERENCE)
    eager, _ = op.lanczos(**kw)
    runs = [op.lanczos(graph=True, **kw) for _ in range(2)]
    for g, st in runs:
        assert np.abs(g - ev).max() < _bar(prec) * np.abs(ev).max(), (g, ev)
        assert np.abs(g - eager).max() < _bar(prec) * np.abs(ev).max(), (g, eager)
        _within_bounds(g, ev, st)
    assert np.array_equal(runs[0][0], runs[1][0])


def test_graph_captured_full_lanczos_exhausts_small_sector(gpu_ready):
    """A 245-state sector with more iterations than states: the Krylov space
    becomes invariant inside a replay batch; the host truncates at the first
    vanishing beta and every eigenvalue matches dense diagonalisation."""
    ed = _ed()
    model = L.chain(14)
    o = P.OracleBasis(model, 7, 0, 0)
    ev = np.linalg.eigvalsh(o.make_h())[:2]
    basis = ed.Basis(ed.Lattice.of(model), 7, 0, 0)
    assert basis.D < 400
    op = ed.Operator(basis, model)
    g, st = op.lanczos(nev=2, iterations=400, ncycles=1, tol=0.0, seed=3, precision=ed.ED_C128,
                       mode=ed.ED_LZ_FULL, graph=True)
    assert np.abs(g - ev).max() < 1e-10 * np.abs(ev).max(), (g, ev)
    _within_bounds(g, ev, st)


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_thick_restart_lanczos_vs_dense(gpu_ready, prec):
    """ED_LZ_THICK with a Krylov space too small to hold the answer without
    restarts (m = 16 for 4 eigenpairs): the 4 lowest eigenvalues against dense
    diagonalisation, unit Ritz vectors with small FP64 residuals."""
    import ed_code_amd as ed
    model, n, k = L.triangular(4, 4, Ch=0.25, J2=0.1), 8, (1, 3)
    print(f"[case] thick restart {prec}", flush=True)
    o = P.OracleBasis(model, n, *k)
    ref = np.linalg.eigvalsh(o.make_h())[:4]
    basis = ed.Basis(ed.Lattice.of(model), n, *k)
    op = ed.Operator(basis, model)
    ev, vecs, st = op.lanczos(nev=4, iterations=16, ncycles=400, tol=1e-11, seed=3, mode=ed.ED_LZ_THICK,
                              precision=ed.ED_C128 if prec == "c128" else ed.ED_C64, want_vectors=True)
    assert np.abs(ev - ref).max() < _bar(prec) * max(1.0, np.abs(ref).max())
    _within_bounds(ev, ref, st)
    for i in range(4):
        x = vecs[i]
        assert abs(np.linalg.norm(x) - 1) < 1e-9
        tol = 1e-6 if prec == "c128" else 1e-4
        assert np.linalg.norm(o.hv(x) - ev[i] * x) < tol
    assert st["hv_calls"] > 16                     # it did restart
