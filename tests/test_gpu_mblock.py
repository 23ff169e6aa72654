"""The multi-sector block operator: every momentum sector of one lattice and
Sz, M coupling instances each, in one H*V launch over a queue of (sector,
row-tile) items, and the Lanczos that advances all (sector, instance) pairs
together.  Against the single-sector block operator, the oracle and dense
diagonalisation.  Bar: H*V within 1e-12 relative (complex128) / 1e-6
(complex64) of the oracle; eigenvalues within 1e-10 of dense, each within
its reported FP64 error bound."""
import numpy as np
import pytest

import pyoracle as P
from ed_code_amd import lattices as L

pytestmark = pytest.mark.gpu


def _ed():
    import ed_code_amd as ed
    ed.init(0)
    return ed


def _tri4x4(M, seed=21):
    rng = np.random.default_rng(seed)
    return [L.triangular(4, 4, J1=1.0, lam=rng.uniform(0.5, 1.5), Ch=rng.uniform(0.05, 0.3), J2=rng.uniform(0.05, 0.6))
            for _ in range(M)]


def _stack(mb, per_sector):
    import torch
    return torch.cat(per_sector, dim=0).contiguous()


@pytest.mark.parametrize("M", [2, 4])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_mblock_hv_all_sectors(gpu_ready, M, dtype):
    """All 16 sectors of the 4x4 triangular torus (k = 0 among them) at Sz=0
    with chiral terms: one launch against each sector's block operator and
    against the oracle's H v of every instance."""
    import torch
    ed = _ed()
    models = _tri4x4(M)
    lat = ed.Lattice.of(models[0])
    keys = [(kx, ky) for kx in range(lat.L1) for ky in range(lat.L2)]
    bases = [ed.Basis(lat, 8, kx, ky) for kx, ky in keys]
    mb = ed.MultiBlockOperator(bases, models)
    assert mb.fast
    assert mb.rows == sum(b.D for b in bases)
    cdt = torch.complex128 if dtype == "c128" else torch.complex64
    g = torch.Generator(device="cuda").manual_seed(M)
    V = torch.randn(mb.rows, M, dtype=torch.complex128, device="cuda", generator=g)
    W = mb.hv(V.to(cdt)).to(torch.complex128)
    bar = 1e-12 if dtype == "c128" else 1e-6
    for s, (b, key) in enumerate(zip(bases, keys)):
        print(f"[case] mblock M={M} {dtype} sector {key}", flush=True)
        Vs, Ws = mb.split(V)[s], mb.split(W)[s]
        blk = ed.BlockOperator(b, models)
        Wb = blk.hv(Vs.to(cdt).contiguous()).to(torch.complex128)
        assert torch.linalg.vector_norm(Ws - Wb).item() <= bar * torch.linalg.vector_norm(Wb).item()
        o = P.OracleBasis(models[0], 8, *key)
        for m, mod in enumerate(models):
            o.set_model(mod)
            wo = o.hv(Vs[:, m].cpu().numpy())
            assert np.linalg.norm(Ws[:, m].cpu().numpy() - wo) <= bar * np.linalg.norm(wo), (key, m)


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_mblock_lanczos_vs_dense(gpu_ready, prec):
    """64 solves (16 sectors x 4 instances), lowest 2 each, in one lock-step
    Lanczos: every eigenvalue against dense diagonalisation and within its
    FP64 error bound; the same eigenvalues as the single-sector block
    Lanczos of each sector."""
    ed = _ed()
    models = _tri4x4(4, seed=33)
    lat = ed.Lattice.of(models[0])
    keys = [(kx, ky) for kx in range(lat.L1) for ky in range(lat.L2)]
    bases = [ed.Basis(lat, 8, kx, ky) for kx, ky in keys]
    mb = ed.MultiBlockOperator(bases, models)
    precision = ed.ED_C128 if prec == "c128" else ed.ED_C64
    print(f"[case] mblock lanczos {prec}", flush=True)
    ev, st = mb.lanczos(nev=2, iterations=300, tol=1e-10, seed=3, precision=precision)
    assert ev.shape == (16, 4, 2)
    for s, key in enumerate(keys):
        o = P.OracleBasis(models[0], 8, *key)
        for m, mod in enumerate(models):
            o.set_model(mod)
            dense = np.sort(np.linalg.eigvalsh(o.make_h()))[:2]
            err = np.abs(ev[s, m] - dense)
            assert err.max() <= 1e-10 * max(1.0, np.abs(dense).max()), (key, m, ev[s, m], dense)
            assert np.all(err <= st["eval_error"][s, m] + 1e-12), (key, m, err, st["eval_error"][s, m])
    for s in (0, 5):
        blk = ed.BlockOperator(bases[s], models)
        eb, _ = blk.lanczos(nev=2, iterations=300, tol=1e-10, seed=3, precision=precision)
        assert np.abs(eb - ev[s]).max() <= 1e-12 * np.abs(eb).max()


def test_mblock_fallback_one_launch_per_sector(gpu_ready):
    """A lattice without the compiled rotation kernels (the tilted 10-site
    cluster) runs its sectors one after the other, with the same results."""
    import torch
    ed = _ed()
    rng = np.random.default_rng(4)
    models = [L.tilted_square((3, 1), (-1, 3), J2=rng.uniform(0.1, 0.5)) for _ in range(2)]
    lat = ed.Lattice.of(models[0])
    bases = [b for b in (ed.Basis(lat, 5, kx, ky) for kx in range(lat.L1) for ky in range(lat.L2)) if b.D > 0][:6]
    mb = ed.MultiBlockOperator(bases, models)
    assert not mb.fast
    V = torch.randn(mb.rows, 2, dtype=torch.complex128, device="cuda")
    W = mb.hv(V)
    for s, b in enumerate(bases):
        Wb = ed.BlockOperator(b, models).hv(mb.split(V)[s].contiguous())
        assert torch.equal(mb.split(W)[s], Wb)
    ev, st = mb.lanczos(nev=1, iterations=200, tol=1e-11, precision=ed.ED_C128)
    for s, b in enumerate(bases):
        for m, mod in enumerate(models):
            o = P.OracleBasis(mod, 5, b.kx, b.ky)
            assert abs(ev[s, m, 0] - np.linalg.eigvalsh(o.make_h())[0]) < 1e-10
