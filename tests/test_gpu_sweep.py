"""GPU check of the sector sweep (every momentum sector of one Sz sector, M
coupling instances per sector on a block operator): the thick-restart block
Lanczos against the deflated full-reorthogonalisation runs on every sector,
and against dense diagonalisation of sampled sectors, instance by instance."""
import numpy as np
import pytest

import pyoracle as P

pytestmark = pytest.mark.gpu


def test_sweep_thick_restart_vs_full_and_dense(gpu_ready):
    import ed_code_amd as ed
    from ed_code_amd import sweep as S
    ed.init(0)
    models = S.coupling_instances("tri4x4", 4)
    n_up = 8
    print("[case] sweep tri4x4 M=4 all sectors", flush=True)
    thick = S.sweep_sectors(models, n_up, nev=2, tol=1e-10, iterations=200, thick=24)
    full = S.sweep_sectors(models, n_up, nev=2, tol=1e-10, iterations=300)
    assert len(thick["sectors"]) == 16 and thick["solves"] == 64
    for a, b in zip(thick["sectors"], full["sectors"]):
        assert (a["kx"], a["ky"]) == (b["kx"], b["ky"])
        assert np.abs(np.array(a["evals"]) - np.array(b["evals"])).max() < 1e-9, (a["kx"], a["ky"])
    for r in thick["sectors"][::5]:
        for m, model in enumerate(models):
            dense = np.sort(np.linalg.eigvalsh(P.OracleBasis(model, n_up, r["kx"], r["ky"]).make_h()))
            assert np.abs(np.array(r["evals"][m]) - dense[:2]).max() < 1e-9, (r["kx"], r["ky"], m)
