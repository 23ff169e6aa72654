"""Memory-lean mode at 40 sites: the tilted square cluster with periods
(6,2), (-2,6) (translation group Z20 x Z2, the reference's walk grid 20 x 20),
J1 on 80 bonds, J2 on 80 diagonals, Sz=0, k=0: D = 3,446,177,112.

The reference cannot hold this sector (rep_loc + state_ph over C(40,20) =
1.4e11 states would take 3.3 TB, and its Lanczos indexes vectors with int,
lanczos.cpp:188,313).  Here the cluster is relabelled so its translations are
row rotations (lattices.regular_relabel), the basis keeps 8 B per
representative plus 1.5 bit per state of the directory (53.4 GB), H*v runs
the two-row 64-bit kernel, and the eigensolver keeps three complex64 Krylov
vectors (ED_LZ_LEAN).

    python -m ed_code_amd.lean                    # basis, H*v timing, Hermiticity
    python -m ed_code_amd.lean --iterations 90    # + lean Lanczos, FP64 residual
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

from . import ED_C64, ED_LZ_LEAN, Basis, Lattice, Operator, init
from . import lattices as L

HBM_PEAK_GBS = 8000.0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--J2", type=float, default=0.5)
    ap.add_argument("--iterations", type=int, default=0, help="lean Lanczos steps (0 = H*v only)")
    ap.add_argument("--tol", type=float, default=1e-9)
    ap.add_argument("--hv", type=int, default=2, help="timed H*v launches")
    ap.add_argument("--no-check", action="store_true", help="skip the Hermiticity check")
    a = ap.parse_args(argv)
    import torch
    init(0)
    log = lambda s: print(s, file=sys.stderr, flush=True)
    m0 = L.tilted_square((6, 2), (-2, 6), J2=a.J2)
    m, info = L.regular_relabel(m0)
    k = info.momentum(0, 0)
    t0 = time.perf_counter()
    basis = Basis(Lattice.of(m), 20, *k)
    torch.cuda.synchronize()
    t_basis = time.perf_counter() - t0
    log(f"[lean] basis D={basis.D} in {t_basis:.1f}s")
    op = Operator(basis, m)
    G, T = op.count_targets()
    D = basis.D
    out = dict(cluster="tilted40 (6,2),(-2,6)", J1=1.0, J2=a.J2, n_up=20, k=[0, 0], D=D, distinct_targets=G,
               basis_build_s=t_basis, relabelled_rows=[info.n2, info.n1])
    g = torch.Generator(device="cuda").manual_seed(1)
    x = torch.randn(D, dtype=torch.complex64, device="cuda", generator=g)
    y = torch.empty_like(x)
    op.hv(x, y)
    torch.cuda.synchronize()
    ms = [op.hv_profile(x, y) for _ in range(a.hv)]
    kms = sum(ms) / len(ms)
    B = D * (8 + 2 * 8) + G * 8
    out.update(hv_kernel_ms=kms, alg_bytes=B, achieved_gbs=B / (kms * 1e-3) / 1e9,
               frac=B / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS)
    log(f"[lean] H*v {kms:.1f} ms, {out['achieved_gbs']:.0f} GB/s algorithmic")
    if not a.no_check:
        def cdot(p, q):
            s, step = 0j, 1 << 30
            for i in range(0, p.numel(), step):
                s += torch.vdot(p[i:i + step].to(torch.complex128), q[i:i + step].to(torch.complex128)).item()
            return s
        z = torch.randn(D, dtype=torch.complex64, device="cuda", generator=g)
        hz = torch.empty_like(z)
        op.hv(z, hz)
        torch.cuda.synchronize()
        lhs, rhs = cdot(z, y), cdot(hz, x)
        out["hermiticity_rel"] = abs(lhs - rhs) / abs(lhs)
        log(f"[lean] <z,Hx> vs <Hz,x>: {out['hermiticity_rel']:.2e}")
        del z, hz
    del x, y
    torch.cuda.empty_cache()
    if a.iterations:
        os.environ["ED_PROGRESS"] = "1"
        t0 = time.perf_counter()
        ev, st = op.lanczos(nev=1, iterations=a.iterations, ncycles=1, tol=a.tol, seed=5, precision=ED_C64,
                            mode=ED_LZ_LEAN)
        out.update(E0=float(ev[0]), E0_per_site=float(ev[0]) / 40, lanczos_seconds=time.perf_counter() - t0,
                   hv_calls=st["hv_calls"], hv_seconds=st["hv_seconds"], fp64_residual=st["residual"][0])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
