"""The drop-in boundary end to end: the reference's own main (ED.cpp and its
input parser, compiled from the reference sources) linked against
libed_amd.so, driven by reference-format input files, against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import helpers as Hp
import pyoracle as P
from ed_code_amd import inputfile as I
from ed_code_amd import lattices as L

pytestmark = pytest.mark.gpu
ED2 = os.path.join(Hp.ROOT, "oracle", "_ref", "ED2_gpu")
needs_ed2 = pytest.mark.skipif(not os.path.exists(ED2), reason="reference main not built here")


def _run(tmp_path, model, Sz, kx, ky, **kw):
    f = str(tmp_path / "in.txt")
    log = str(tmp_path / "log.txt")
    I.write_input(f, model, Sz, kx, ky, **kw)
    r = subprocess.run([ED2, f, log], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, open(log).read()[-2000:]
    return I.parse_energies(log), log


@needs_ed2
@pytest.mark.parametrize("name,model,Sz,k", [
    ("chain12_k3", L.chain(12), 0.0, (3, 0)),
    ("tri4x3_chch_k11", L.triangular(4, 3, J1=1.0, Ch=0.3, ChCh=0.2), 0.0, (1, 1)),
    ("sq4x4_j2_sz1_k12", L.square(4, 4, J2=0.5, lam=0.8), 1.0, (1, 2)),
])
def test_reference_main_dense_path(gpu_ready, tmp_path, name, model, Sz, k):
    print(f"[case] {name}", flush=True)
    e, _ = _run(tmp_path, model, Sz, *k)
    n = L.n_ones_for_sz(model.n_sites, Sz)
    o = P.OracleBasis(model, n, *k)
    ev = np.linalg.eigvalsh(o.make_h())
    assert len(e) == o.D                     # exact_evecs returns every eigenvalue
    assert np.abs(np.array(e) - ev).max() < 1e-10


@needs_ed2
def test_reference_main_lanczos_path_and_vectors(gpu_ready, tmp_path):
    """D > 10000 takes the Lanczos branch (Lanczos_sector.cpp:105)."""
    model = L.chain(22)
    vp = str(tmp_path / "vec")
    e, log = _run(tmp_path, model, 0.0, 0, 0, num_eigenvalues=2, find_vectors=True, unwrap_pvector=False,
                  vector_file_path=vp, lanczos_iterations=80, lanczos_ncycles=4, tol=1e-12)
    o = P.OracleBasis(model, 11, 0, 0)
    assert o.D > 10000
    lo = o.lanczos(nev=2, iterations=80, ncycles=4, tol=1e-12, seed=3)
    assert np.abs(np.array(e) - lo).max() < 1e-10 * abs(lo[0])
    # momentum-basis eigenvector file (Lanczos_sector.cpp:182-190): D rows, unit norm, H x = E x
    lines = open(vp + "_pEV0.txt").read().splitlines()
    start = next(i for i, l in enumerate(lines) if l.startswith("pblock_states"))
    rows = [l.split() for l in lines[start + 1:] if l.strip()]
    assert len(rows) == o.D
    x = np.array([complex(*map(float, r[2].strip("()").split(","))) for r in rows])
    assert abs(np.linalg.norm(x) - 1) < 1e-9
    r = o.hv(x) - e[0] * x
    assert np.linalg.norm(r) < 1e-6


@needs_ed2
def test_reference_main_unwrapped_vector(gpu_ready, tmp_path):
    model = L.chain(12)
    vp = str(tmp_path / "vec")
    e, _ = _run(tmp_path, model, 0.0, 2, 0, num_eigenvalues=1, find_vectors=True, unwrap_pvector=True,
                vector_file_path=vp)
    rows = [l.split() for l in open(vp + "_EV0.txt") if l.strip() and l.split()[0].isdigit() and len(l.split()) == 2]
    ids = [int(r[0]) for r in rows]
    c = np.array([complex(*map(float, r[1].strip("()").split(","))) for r in rows])
    assert ids == sorted(ids) and len(set(ids)) == len(ids)
    assert abs(np.linalg.norm(c) - 1) < 1e-9
    # the unwrapped vector is an eigenvector of the full-space H with momentum k
    H, st = Hp.full_space_h(model, 6)
    idx = {s: i for i, s in enumerate(st)}
    psi = np.zeros(len(st), complex)
    for s, a in zip(ids, c):
        psi[idx[s]] = a
    assert np.linalg.norm(H @ psi - e[0] * psi) < 1e-8


# ------------------------------------------------- three generators (3-D) ----
def _driver(*args):
    r = subprocess.run([Hp.refapi_driver(), *args], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


@pytest.mark.parametrize("model,n,k", [(L.cubic(2, 3, 4, J2=0.4, lam=0.8), 12, (1, 2, 3)),
                                       (L.cubic(2, 2, 4), 8, (0, 0, 0)), (L.cubic(3, 2, 3, J2=0.3), 9, (2, 1, 1))],
                         ids=["cubic2x3x4_k123", "cubic2x2x4_k000", "cubic3x2x3_k211"])
def test_three_generator_shv_vs_oracle(gpu_ready, tmp_path, model, n, k):
    """pConstruct_States + sHv (H_v.h:59-83, Momentum_Basis.h:84-100) through
    the C++ interface against the oracle's three-generator basis and H*v."""
    print(f"[case] sHv3 {model.name} n={n} k={k}", flush=True)
    o = P.OracleBasis(model, n, *k)
    v = np.random.default_rng(31).normal(size=o.D) + 1j * np.random.default_rng(32).normal(size=o.D)
    spec, fv, fw, fr = (str(tmp_path / x) for x in ("spec", "v.bin", "w.bin", "r.bin"))
    Hp.write_spec(spec, model, n, k)
    v.astype(np.complex128).tofile(fv)
    out = _driver("hv", spec, fv, fw, fr)
    assert f"D {o.D}" in out
    assert np.array_equal(np.fromfile(fr, np.int64), o.tables()["reps"])
    w = np.fromfile(fw, np.complex128)
    wo = o.hv(v)
    assert np.linalg.norm(w - wo) <= 1e-12 * np.linalg.norm(wo)


def test_three_generator_sector_lanczos_and_vectors(gpu_ready, tmp_path):
    """sector_lanczos (Lanczos_sector.cpp:201-367): energies against the dense
    spectrum, the momentum-basis and unwrapped eigenvector files."""
    model, n, k = L.cubic(2, 2, 4, J2=0.3), 8, (1, 0, 2)
    print(f"[case] sector3 {model.name} n={n} k={k}", flush=True)
    o = P.OracleBasis(model, n, *k)
    ev = np.linalg.eigvalsh(o.make_h())
    spec, log, vp = str(tmp_path / "spec"), str(tmp_path / "log.txt"), str(tmp_path / "vec")
    Hp.write_spec(spec, model, n, k, nev=2, iterations=120, ncycles=8, tol=1e-12)
    _driver("sector", spec, log, vp, "0")
    text = open(log).read()
    assert f"\t{o.D}" in text and "kz" in text
    e = I.parse_energies(log)
    assert len(e) == 2 and np.abs(np.array(e) - ev[:2]).max() < 1e-10 * abs(ev[0])
    lines = open(vp + "_pEV0.txt").read().splitlines()
    assert "(1,0,2)" in lines[3]
    start = next(i for i, l in enumerate(lines) if l.startswith("pblock_states"))
    rows = [l.split() for l in lines[start + 1:] if l.strip()]
    x = np.array([complex(*map(float, r[2].strip("()").split(","))) for r in rows])
    assert len(x) == o.D and abs(np.linalg.norm(x) - 1) < 1e-9
    assert np.linalg.norm(o.hv(x) - e[0] * x) < 1e-6
    # unwrapped (Momentum_Basis.cpp:901-975): an eigenvector of the full-space H
    _driver("sector", spec, str(tmp_path / "log2.txt"), vp, "1")
    rows = [l.split() for l in open(vp + "_EV0.txt") if l.strip() and l.split()[0].isdigit() and len(l.split()) == 2]
    ids = [int(r[0]) for r in rows]
    c = np.array([complex(*map(float, r[1].strip("()").split(","))) for r in rows])
    assert ids == sorted(ids) and len(set(ids)) == len(ids)
    assert abs(np.linalg.norm(c) - 1) < 1e-9
    H, st = Hp.full_space_h(model, n)
    idx = {s: i for i, s in enumerate(st)}
    psi = np.zeros(len(st), complex)
    for s, a in zip(ids, c):
        psi[idx[s]] = a
    assert np.linalg.norm(H @ psi - e[0] * psi) < 1e-8


# --------------------------- makeH, sparse sHv, ARPACK_Iteration_EV (2-D) ----
# bond terms only: the driver's input carries bonds
MAKEH = [(L.tilted_square((3, 1), (-1, 3), J2=0.5), 5, (1, 7)), (L.square(4, 4, J2=0.5, lam=0.8), 8, (1, 2)),
         (L.chain(14, phases=[0.2] * 14), 7, (3, 0))]


@pytest.mark.parametrize("model,n,k", MAKEH, ids=[f"{m.name}_k{k[0]}{k[1]}" for m, n, k in MAKEH])
def test_makeh_and_sparse_shv(gpu_ready, tmp_path, model, n, k):
    """makeH (H_v.cpp:21-190) through the C++ interface against the oracle's
    dense matrix; the sparse sHv (H_v.cpp:3-19) of the same matrix against
    H v."""
    print(f"[case] makeH {model.name} n={n} k={k}", flush=True)
    o = P.OracleBasis(model, n, *k)
    spec, fh, fv, fw = (str(tmp_path / x) for x in ("spec", "H.bin", "v.bin", "w.bin"))
    Hp.write_spec(spec, model, n, k)
    assert f"D {o.D}" in _driver("makeh", spec, fh)
    H = np.fromfile(fh, np.complex128).reshape(o.D, o.D, order="F")
    Ho = o.make_h()
    assert np.abs(H - Ho).max() <= 1e-12 * max(1.0, np.abs(Ho).max())
    v = np.random.default_rng(41).normal(size=o.D) + 1j * np.random.default_rng(42).normal(size=o.D)
    v.astype(np.complex128).tofile(fv)
    _driver("sparse", spec, fv, fw)
    w = np.fromfile(fw, np.complex128)
    assert np.linalg.norm(w - H @ v) <= 1e-12 * np.linalg.norm(H @ v)


def test_arpack_iteration_ev(gpu_ready, tmp_path):
    """ARPACK_Iteration_EV (ARPACK_sector.cpp:4-187): the n_it lowest
    eigenvalues largest first, Evecs(i, j) the matching Ritz vectors."""
    model, n, k = L.square(4, 4, J2=0.3), 8, (0, 0)
    print("[case] ARPACK_Iteration_EV sq4x4", flush=True)
    o = P.OracleBasis(model, n, *k)
    ev = np.linalg.eigvalsh(o.make_h())
    spec, fe, fv = str(tmp_path / "spec"), str(tmp_path / "e.bin"), str(tmp_path / "V.bin")
    Hp.write_spec(spec, model, n, k, nev=3, iterations=300, tol=1e-12)
    out = _driver("arpack", spec, fe, fv)
    assert f"D {o.D} nev 3 rows 3" in out
    e = np.fromfile(fe, np.complex128)
    assert np.abs(e.real - ev[:3][::-1]).max() < 1e-10 * abs(ev[0]) and np.abs(e.imag).max() == 0
    V = np.fromfile(fv, np.complex128).reshape(3, o.D, order="F")
    for i in range(3):
        x = V[i]
        assert abs(np.linalg.norm(x) - 1) < 1e-9
        assert np.linalg.norm(o.hv(x) - e[i].real * x) < 1e-6


# ------------------------------------- ARPACK_sector, sector_eigs_eigvecs ----
ARPSEC = [("sq4x4_nev3", L.square(4, 4, J2=0.3), 8, (0, 0), 3),      # n_p > n_ev + 1: Arnoldi branch
          ("chain6_n2_nev2", L.chain(6, J=0.7), 2, (0, 0), 2),        # n_p = 3 = n_ev + 1: dense branch
          ("chain6_n2_nev5", L.chain(6, J=0.7), 2, (0, 0), 5)]        # n_ev > n_p - 1: dense, it = n_p


@pytest.mark.parametrize("name,model,n,k,nev", ARPSEC, ids=[c[0] for c in ARPSEC])
def test_arpack_sector_branches(gpu_ready, tmp_path, name, model, n, k, nev):
    """ARPACK_sector (ARPACK_sector.cpp:186-259): with n_p > n_ev + 1 the n_ev
    lowest eigenvalues largest first and the vectors as rows of
    pblock_eigvecs; otherwise (:232-251) it = n_p, all n_p eigenpairs of the
    dense matrix, the vectors as its columns, the matrix in the log."""
    print(f"[case] ARPACK_sector {name}", flush=True)
    o = P.OracleBasis(model, n, *k)
    Ho = o.make_h()
    ev = np.linalg.eigvalsh(Ho)
    spec, fe, fv, log = (str(tmp_path / x) for x in ("spec", "e.bin", "V.bin", "log.txt"))
    Hp.write_spec(spec, model, n, k, nev=nev, ncycles=300, tol=1e-12)
    out = _driver("arpsec", spec, fe, fv, log)
    e = np.fromfile(fe, np.complex128)
    V = np.fromfile(fv, np.complex128)
    dense = o.D <= nev + 1
    it = o.D if dense else nev
    assert f"D {o.D} nev {it} " in out
    assert np.abs(e.imag).max() == 0
    assert np.abs(e.real - ev[:it][::-1]).max() < 1e-10
    text = open(log).read()
    if dense:
        assert " Here " in text and f"rows {o.D} cols {o.D}" in out
        V = V.reshape(o.D, o.D, order="F")
        vecs = [V[:, q] for q in range(it)]          # columns, as zgeev's VR
        mat = [l for l in text.split(" Here ")[1].split("Eigenvalues")[0].splitlines() if l.startswith("(")]
        assert len(mat) == o.D
        for ii, l in enumerate(mat):
            vals = [complex(*map(float, t.strip("()").split(","))) for t in l.split()]
            assert np.abs(np.array(vals) - Ho[ii]).max() < 1e-5
    else:
        V = V.reshape(it, o.D, order="F")
        vecs = [V[q] for q in range(it)]
    assert "Eigenvalues" in text
    for q in range(it):
        x = vecs[q]
        assert abs(np.linalg.norm(x) - 1) < 1e-9
        assert np.linalg.norm(o.hv(x) - e[q].real * x) < 1e-6


@pytest.mark.parametrize("unwrap", [0, 1], ids=["pEV", "EV"])
def test_sector_eigs_eigvecs(gpu_ready, tmp_path, unwrap):
    """sector_eigs_eigvecs (ARPACK_sector.cpp:262-401): the energy table
    ascending, eigenvector files numbered from '1', file c holding the vector
    of the c-th returned value (largest of the wanted first)."""
    model, n, k, nev = L.square(4, 4, J2=0.4, lam=0.9), 9, (1, 2), 2
    print(f"[case] sector_eigs_eigvecs unwrap={unwrap}", flush=True)
    o = P.OracleBasis(model, n, *k)
    ev = np.linalg.eigvalsh(o.make_h())
    spec, log, vp = str(tmp_path / "spec"), str(tmp_path / "log.txt"), str(tmp_path / "vec")
    Hp.write_spec(spec, model, n, k, nev=nev, ncycles=300, tol=1e-12)
    _driver("eigs", spec, log, vp, str(unwrap))
    e = I.parse_energies(log)
    assert len(e) == nev and np.abs(np.array(e) - ev[:nev]).max() < 1e-10
    for c in range(nev):
        f = vp + ("_EV" if unwrap else "_pEV") + str(c + 1) + ".txt"
        lines = open(f).read().splitlines()
        E = float(next(l for l in lines if l.startswith("State Energy")).split()[2].strip("()").split(",")[0])
        # the file prints the energy with the stream's default 6 digits
        assert abs(E - ev[nev - 1 - c]) < 1e-5 * max(1.0, abs(E))
        E = ev[nev - 1 - c]
        if unwrap:
            rows = [l.split() for l in lines if l.strip() and l.split()[0].isdigit() and len(l.split()) == 2]
            ids = [int(r[0]) for r in rows]
            amp = np.array([complex(*map(float, r[1].strip("()").split(","))) for r in rows])
            H, st = Hp.full_space_h(model, n)
            idx = {s: i for i, s in enumerate(st)}
            psi = np.zeros(len(st), complex)
            for s, a in zip(ids, amp):
                psi[idx[s]] = a
            assert abs(np.linalg.norm(psi) - 1) < 1e-9
            assert np.linalg.norm(H @ psi - E * psi) < 1e-8
        else:
            start = next(i for i, l in enumerate(lines) if l.startswith("pblock_states"))
            rows = [l.split() for l in lines[start + 1:] if l.strip()]
            x = np.array([complex(*map(float, r[2].strip("()").split(","))) for r in rows])
            assert len(x) == o.D and abs(np.linalg.norm(x) - 1) < 1e-9
            assert np.linalg.norm(o.hv(x) - E * x) < 1e-6
    assert not os.path.exists(vp + ("_EV0.txt" if unwrap else "_pEV0.txt"))


@needs_ed2
def test_reference_main_complex64_krylov(gpu_ready, tmp_path):
    """ED_PRECISION=c64: the reference's own main on the Lanczos branch with
    complex64 Krylov storage, FP64 alpha/beta and the FP64 Rayleigh-quotient
    correction; energies within 1e-10 of the complex128 run, FP64 residuals
    in the log."""
    model = L.chain(22)
    out = {}
    for prec in ("c128", "c64"):
        f, log = str(tmp_path / f"in_{prec}.txt"), str(tmp_path / f"log_{prec}.txt")
        I.write_input(f, model, 0.0, 0, 0, num_eigenvalues=2, lanczos_iterations=80, lanczos_ncycles=4, tol=1e-12)
        r = subprocess.run([ED2, f, log], capture_output=True, text=True, timeout=180,
                           env=dict(os.environ, ED_PRECISION=prec))
        assert r.returncode == 0, open(log).read()[-2000:]
        out[prec] = I.parse_energies(log)
        text = open(log).read()
        assert ("complex64" if prec == "c64" else "complex128") in text
        res = [float(l.split()[-1]) for l in text.splitlines() if l.startswith(" FP64 residual")]
        assert len(res) == 2 and max(res) < 1e-4
    assert np.abs(np.array(out["c64"]) - np.array(out["c128"])).max() < 1e-10 * abs(out["c128"][0])


@needs_ed2
def test_reference_main_graph_and_eager_steps(gpu_ready, tmp_path):
    """The drop-in runs the reference schedule's steps as replayed hipGraphs
    by default; ED_GRAPH=0 launches them one by one.  The reference's own
    main on both: the same energies (within 1e-12 relative), each log saying
    which ran (22 sites: above the dense-path size)."""
    model = L.chain(22)
    out = {}
    for mode in ("1", "0"):
        f, log = str(tmp_path / f"in_{mode}.txt"), str(tmp_path / f"log_{mode}.txt")
        I.write_input(f, model, 0.0, 0, 0, num_eigenvalues=3, lanczos_iterations=60, lanczos_ncycles=5, tol=1e-12)
        r = subprocess.run([ED2, f, log], capture_output=True, text=True, timeout=180,
                           env=dict(os.environ, ED_GRAPH=mode))
        assert r.returncode == 0, open(log).read()[-2000:]
        out[mode] = I.parse_energies(log)
        assert ("hipGraph replays" if mode == "1" else "eager launches") in open(log).read()
    assert len(out["1"]) == 3
    assert np.abs(np.array(out["1"]) - np.array(out["0"])).max() < 1e-12 * abs(out["0"][0])
