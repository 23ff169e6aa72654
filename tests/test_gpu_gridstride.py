"""The grid-stride loops of the basis build and the gather-image passes
(the form every launch takes past 2^22 blocks, as at 40 sites, where a
single pass over 2^32 work-items silently classified a fraction of the
states and a 32-bit scan length gave D = 0) on small sectors: the launches
capped at a few blocks through ED_GRID_CAP (set before the library loads, so
in a child process), tables and H*v against the oracle."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import helpers as Hp

pytestmark = pytest.mark.gpu

CHILD = r"""
import json, sys
import numpy as np
import torch
sys.path.insert(0, sys.argv[1])
import ed_code_amd as ed
from ed_code_amd import lattices as L
ed.init(0)
out = {}
for name, model, n, k in [("sq4x4_k12", L.square(4, 4, J2=0.5), 8, (1, 2)),
                          ("tri4x4_ch_k13", L.triangular(4, 4, Ch=0.2), 8, (1, 3)),
                          ("tilted20_two_row", L.regular_relabel(L.tilted_square((4, 2), (-2, 4), J2=0.5))[0], 10, (0, 0))]:
    b = ed.Basis(ed.Lattice.of(model), n, *k)
    reps, norm, _ = b.export()
    full = b.export_full()
    op = ed.Operator(b, model)
    v = np.random.default_rng(4).normal(size=b.D) + 1j * np.random.default_rng(5).normal(size=b.D)
    w = op.hv(torch.from_numpy(v).to("cuda")).cpu().numpy()
    np.savez(sys.argv[2] + "_" + name + ".npz", reps=reps, norm=norm, rep_loc=full["rep_loc"],
             state_ph=full["state_ph"], v=v, w=w)
    out[name] = int(b.D)
print(json.dumps(out))
"""


def test_capped_grid_matches_oracle(gpu_ready, tmp_path):
    import pyoracle as P
    from ed_code_amd import lattices as L
    print("[case] grid-stride basis build, ED_GRID_CAP=3", flush=True)
    script = tmp_path / "child.py"
    script.write_text(CHILD)
    prefix = str(tmp_path / "t")
    r = subprocess.run([sys.executable, str(script), Hp.ROOT, prefix], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, ED_GRID_CAP="3"))
    assert r.returncode == 0, r.stderr[-3000:]
    dims = json.loads(r.stdout.strip().splitlines()[-1])
    cases = {"sq4x4_k12": (L.square(4, 4, J2=0.5), 8, (1, 2)),
             "tri4x4_ch_k13": (L.triangular(4, 4, Ch=0.2), 8, (1, 3)),
             "tilted20_two_row": (L.regular_relabel(L.tilted_square((4, 2), (-2, 4), J2=0.5))[0], 10, (0, 0))}
    for name, (model, n, k) in cases.items():
        d = np.load(prefix + "_" + name + ".npz")
        o = P.OracleBasis(model, n, *k)
        t = o.tables()
        assert dims[name] == o.D
        assert np.array_equal(d["reps"], t["reps"]) and np.array_equal(d["norm"], t["norm"])
        assert np.array_equal(d["rep_loc"], t["rep_loc"])
        assert np.array_equal(d["state_ph"].view(np.float64), t["state_ph"].view(np.float64))
        wo = o.hv(d["v"])
        assert np.linalg.norm(d["w"] - wo) <= 1e-12 * np.linalg.norm(wo), name
