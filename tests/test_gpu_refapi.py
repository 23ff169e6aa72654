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
    assert np.abs(np.array(e) - lo).max() < 1e-9
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
