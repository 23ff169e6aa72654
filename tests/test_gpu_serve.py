"""GPU check of the serving loop on small lattices: every request answered,
the answers equal to the one-at-a-time loop's and to dense diagonalisation,
batching and the basis cache exercised."""
import numpy as np
import pytest

import pyoracle as P


@pytest.mark.gpu
def test_serving_loop_matches_plain_loop_and_dense(gpu_ready):
    import ed_code_amd as ed
    from ed_code_amd import serve as S
    ed.init(0)
    print("[case] serve chain16/tri4x4 k=0 burst", flush=True)
    # a burst: all 12 requests arrive within microseconds, so the ones the
    # first dispatch does not take are pending together, and 12 requests over
    # at most 8 batch keys (lattice, Sz, nev at k=0) must share one
    reqs = S.make_requests(12, rate=1e6, seed=3, lattices=["chain16", "tri4x4"], k_zero=True)
    srv = S.Server(workers=3, iterations=200)
    t0, wall = srv.run(reqs)
    assert not srv.errors
    out = S.summarize(reqs, wall, t0)
    assert out["answered"] == len(reqs)
    assert max(out["batch_sizes"]) >= 2          # a burst at k=0 batches
    assert srv.cache.hits >= 1
    base = S.make_requests(12, rate=1e6, seed=3, lattices=["chain16", "tri4x4"], k_zero=True)
    S.run_baseline(base, iterations=200)
    for r, b in zip(reqs, base):
        assert r.rid == b.rid
        assert np.abs(np.array(r.evals) - np.array(b.evals)).max() < 1e-9
        m = S.model_for(r.lattice, r.J2, r.lam, r.Ch)
        n_up = S.sector_key(r)[1]
        dense = np.sort(np.linalg.eigvalsh(P.OracleBasis(m, n_up, r.kx, r.ky).make_h()))
        assert np.abs(np.array(r.evals) - dense[:r.nev]).max() < 1e-9, (r.rid, r.evals, dense[:r.nev])


def test_footprint_admission_is_monotone():
    from ed_code_amd import serve as S
    a = S.footprint(S.estimate_rows("tri6x5", 15), 4, 300, 2)
    b = S.footprint(S.estimate_rows("tri6x5", 15), 8, 300, 2)
    assert 0 < a < b
    assert S.estimate_rows("sq4x8", 16) == 601080390 // 32 + 1


@pytest.mark.gpu
def test_serving_random_momenta_and_error_bounds(gpu_ready):
    """Random allowed momenta (k != 0 sectors, complex matrix elements):
    every answer equals dense diagonalisation and lies within the FP64
    error bound the solver reports for it; the device-derived HBM budget
    is positive."""
    import ed_code_amd as ed
    from ed_code_amd import serve as S
    ed.init(0)
    print("[case] serve chain16/tri4x4 random k", flush=True)
    reqs = S.make_requests(10, rate=50.0, seed=5, lattices=["chain16", "tri4x4"], k_zero=False)
    assert any((r.kx, r.ky) != (0, 0) for r in reqs)
    srv = S.Server(workers=3, iterations=200)
    assert srv.budget > 0
    t0, wall = srv.run(reqs)
    assert not srv.errors
    out = S.summarize(reqs, wall, t0)
    assert out["answered"] == len(reqs) and out["max_eval_error_bound"] is not None
    for r in reqs:
        m = S.model_for(r.lattice, r.J2, r.lam, r.Ch)
        dense = np.sort(np.linalg.eigvalsh(P.OracleBasis(m, S.sector_key(r)[1], r.kx, r.ky).make_h()))
        err = np.abs(np.array(r.evals) - dense[:r.nev])
        assert err.max() < 1e-9, (r.rid, r.evals, dense[:r.nev])
        assert len(r.errors) == r.nev
        assert np.all(err <= np.array(r.errors) + 1e-12), (r.rid, err, r.errors)
