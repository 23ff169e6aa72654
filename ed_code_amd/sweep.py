"""Sector sweeps with block operators: every momentum sector of one Sz
sector, M coupling instances per sector solved together.

The reference computes one coupling set per run and sweeps sectors one at a
time (ARPACK_full_eigs.cpp:169-277 loops n_ones and every (kx, ky), calling
ARPACK per sector).  Here each sector's basis is built once on the device and
the M instances share every H*V pass (ed_block_hv) and one Lanczos loop
(ed_block_lanczos).

    python -m ed_code_amd.sweep --lattice tri6x5 --instances 4 --nev 2 --tol 1e-8

By default each sector runs the block thick-restart Lanczos (16 Krylov
vectors per instance, at least 4 nev + 8; the nev pairs from one space);
--thick 0 runs one full-reorthogonalisation Lanczos per eigenvalue with
deflation instead (30-site triangular torus, 30 sectors x 4 couplings,
nev = 2: 23 s with 16 vectors).
"""
from __future__ import annotations

import argparse
import json
import time
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import (ED_C64, ED_C128, ED_LZ_FULL, ED_LZ_THICK, Basis, BlockOperator, Lattice, MultiBlockOperator, Operator,
               init)
from . import lattices as L
from .lattices import Model


def coupling_instances(lattice: str, M: int, seed: int = 7) -> List[Model]:
    """M seeded coupling instances with one term structure: triangular tori
    draw Ch ~ U[0, 0.3], square tori J2 ~ U[0, 0.6], chains Delta ~
    U[0.5, 1.5]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(M):
        if lattice.startswith("tri"):
            a, b = (int(t) for t in lattice[3:].split("x"))
            out.append(L.triangular(a, b, Ch=float(rng.uniform(0.0, 0.3))))
        elif lattice.startswith("sq"):
            a, b = (int(t) for t in lattice[2:].split("x"))
            out.append(L.square(a, b, J2=float(rng.uniform(1e-3, 0.6))))
        elif lattice.startswith("chain"):
            out.append(L.chain(int(lattice[5:]), lam=float(rng.uniform(0.5, 1.5))))
        else:
            raise ValueError(f"unknown lattice {lattice}")
    return out


def sweep_sectors(models: Sequence[Model], n_up: int, nev: int = 2, tol: float = 1e-8, iterations: int = 300,
                  precision: int = ED_C64, sectors: Optional[Sequence[tuple]] = None, seed: int = 1234,
                  log: Optional[Callable[[str], None]] = None, thick: int = 0) -> dict:
    """Lowest nev eigenvalues of every instance in every momentum sector
    (or the listed (kx, ky) sectors).  thick > 0: the block thick-restart
    Lanczos with that many Krylov vectors per instance, at least 4 nev + 8
    (all nev pairs from one space); 0: one full-reorthogonalisation run of up to `iterations`
    steps per eigenvalue.  Returns per-sector results and totals."""
    lat = Lattice.of(models[0])
    if sectors is None:
        sectors = [(kx, ky) for kx in range(lat.L1) for ky in range(lat.L2)]
    res, t_all = [], time.perf_counter()
    hv_calls = 0
    for (kx, ky) in sectors:
        t0 = time.perf_counter()
        basis = Basis(lat, n_up, kx, ky)
        if basis.D == 0:
            continue
        blk = BlockOperator(basis, models)
        if thick > 0:
            ev, st = blk.lanczos(nev=nev, iterations=max(thick, 4 * nev + 8), ncycles=max(1, iterations), tol=tol,
                                 seed=seed, precision=precision, mode=ED_LZ_THICK)
        else:
            ev, st = blk.lanczos(nev=nev, iterations=iterations, tol=tol, seed=seed, precision=precision)
        dt = time.perf_counter() - t0
        hv_calls += st["hv_calls"]
        res.append(dict(kx=kx, ky=ky, D=basis.D, evals=ev.tolist(), residual=st["residual"][:len(models)],
                        seconds=dt, hv_calls=st["hv_calls"], hv_seconds=st["hv_seconds"], fast=blk.fast))
        if log:
            log(f"[sweep] k=({kx},{ky}) D={basis.D} {dt:.2f}s hv={st['hv_calls']} "
                f"E0={[round(e[0], 10) for e in ev.tolist()]} res={max(st['residual'][:len(models)]):.1e}")
        del blk, basis
    total = time.perf_counter() - t_all
    solves = len(res) * len(models)
    return dict(sectors=res, seconds=total, solves=solves, solves_per_s=solves / total if total else 0.0,
                hv_block_calls=hv_calls)


def krylov_bytes(D: int, M: int, iterations: int, precision: int) -> int:
    """Device bytes one sector adds to a multi-sector Lanczos: Krylov vectors
    (the complex128 stage that follows a complex64 one when tol asks for it
    holds as many vectors at twice the size, so complex128 decides), two work
    vectors, FP64 Ritz vectors and the gather images (at most 6 elements per
    row) of both precisions."""
    rows = D * M
    krylov = (iterations + 1) * rows * 16
    return krylov + rows * (2 * 16 + 4 * 16 + 6 * (8 + 16)) + D * 20


def sweep_sectors_multi(models: Sequence[Model], n_up: int, nev: int = 2, tol: float = 1e-8, iterations: int = 120,
                        precision: int = ED_C64, sectors: Optional[Sequence[tuple]] = None, seed: int = 1234,
                        budget: Optional[int] = None, log: Optional[Callable[[str], None]] = None) -> dict:
    """sweep_sectors with the sectors in as few multi-sector block operators
    as device memory allows: each group's every Lanczos step is one H*V
    launch over all its sectors and instances (ed_mblock_lanczos).  budget:
    device bytes for the groups' Krylov storage (default 60 % of the free
    memory)."""
    import torch
    lat = Lattice.of(models[0])
    if sectors is None:
        sectors = [(kx, ky) for kx in range(lat.L1) for ky in range(lat.L2)]
    if budget is None:
        budget = int(0.6 * torch.cuda.mem_get_info()[0])
    t_all = time.perf_counter()
    bases = []
    for (kx, ky) in sectors:
        b = Basis(lat, n_up, kx, ky)
        if b.D > 0:
            bases.append(((kx, ky), b))
    groups, cur, used = [], [], 0
    for item in bases:
        need = krylov_bytes(item[1].D, len(models), iterations, precision)
        if cur and used + need > budget:
            groups.append(cur)
            cur, used = [], 0
        cur.append(item)
        used += need
    if cur:
        groups.append(cur)
    res, hv_calls, launches = [], 0, 0
    for g in groups:
        t0 = time.perf_counter()
        mb = MultiBlockOperator([b for _, b in g], models)
        ev, st = mb.lanczos(nev=nev, iterations=iterations, tol=tol, seed=seed, precision=precision)
        dt = time.perf_counter() - t0
        hv_calls += st["hv_calls"]
        for s, ((kx, ky), b) in enumerate(g):
            res.append(dict(kx=kx, ky=ky, D=b.D, evals=ev[s].tolist(), eval_error=st["eval_error"][s].tolist(),
                            seconds=dt / len(g), group=len(groups) and groups.index(g), fast=mb.fast))
        if log:
            log(f"[sweep] group of {len(g)} sectors ({sum(b.D for _, b in g)} rows) {dt:.2f}s "
                f"hv={st['hv_calls']} fast={mb.fast} max err bound={st['eval_error'].max():.1e}")
        del mb
    total = time.perf_counter() - t_all
    solves = len(res) * len(models)
    return dict(sectors=res, seconds=total, solves=solves, solves_per_s=solves / total if total else 0.0,
                hv_calls=hv_calls, groups=[len(g) for g in groups])


def one_at_a_time(models: Sequence[Model], n_up: int, sectors: Sequence[tuple], nev: int = 2, tol: float = 1e-8,
                  iterations: int = 300, precision: int = ED_C64, seed: int = 1234) -> dict:
    """The plain loop the block sweep replaces: one single-instance Lanczos
    (full reorthogonalisation with the same conditional second pass, same
    precision) per (sector, instance)."""
    lat = Lattice.of(models[0])
    t_all = time.perf_counter()
    out = []
    for (kx, ky) in sectors:
        basis = Basis(lat, n_up, kx, ky)
        if basis.D == 0:
            continue
        for m in models:
            op = Operator(basis, m)
            ev, st = op.lanczos(nev=nev, iterations=iterations, ncycles=1, tol=tol * 1e-3, seed=seed,
                                precision=precision, mode=ED_LZ_FULL, selective=True)
            out.append(dict(kx=kx, ky=ky, evals=list(ev)))
            del op
    total = time.perf_counter() - t_all
    return dict(results=out, seconds=total, solves=len(out))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lattice", default="tri6x5")
    ap.add_argument("--instances", type=int, default=4)
    ap.add_argument("--sz", type=float, default=0.0)
    ap.add_argument("--nev", type=int, default=2)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--iterations", type=int, default=300)
    ap.add_argument("--precision", choices=["c64", "c128"], default="c64")
    ap.add_argument("--sectors", type=int, default=0, help="only the first n sectors (0 = all)")
    ap.add_argument("--compare", type=int, default=0, help="also run the one-at-a-time loop on n sectors")
    ap.add_argument("--thick", type=int, default=16,
                    help="block thick-restart Lanczos with this many Krylov vectors per instance (at least "
                         "4 nev + 8), at most "
                         "--iterations restarts (0: one full-reorthogonalisation run of up to --iterations "
                         "steps per eigenvalue)")
    ap.add_argument("--multi", action="store_true",
                    help="multi-sector block operators: every Lanczos step one H*V launch over all sectors of a group")
    a = ap.parse_args(argv)
    init(0)
    models = coupling_instances(a.lattice, a.instances)
    n_up = L.n_ones_for_sz(models[0].n_sites, a.sz)
    lat = Lattice.of(models[0])
    secs = [(kx, ky) for kx in range(lat.L1) for ky in range(lat.L2)]
    if a.sectors:
        secs = secs[:a.sectors]
    prec = ED_C64 if a.precision == "c64" else ED_C128
    import sys
    logf = lambda s: print(s, file=sys.stderr, flush=True)
    if a.multi:
        out = sweep_sectors_multi(models, n_up, a.nev, a.tol, a.iterations, prec, secs, log=logf)
    else:
        out = sweep_sectors(models, n_up, a.nev, a.tol, a.iterations, prec, secs, log=logf, thick=a.thick)
    out["lattice"], out["instances"], out["n_up"] = a.lattice, a.instances, n_up
    if a.compare:
        cmp_secs = [(r["kx"], r["ky"]) for r in out["sectors"][:a.compare]]
        t_blk = sum(r["seconds"] for r in out["sectors"][:a.compare])
        base = one_at_a_time(models, n_up, cmp_secs, a.nev, a.tol, a.iterations, prec)
        worst = 0.0
        for r in out["sectors"][:a.compare]:
            refs = [x["evals"] for x in base["results"] if x["kx"] == r["kx"] and x["ky"] == r["ky"]]
            for mi, ref in enumerate(refs):
                worst = max(worst, max(abs(p - q) for p, q in zip(r["evals"][mi], ref)))
        out["compare"] = dict(sectors=len(cmp_secs), block_seconds=t_blk, one_at_a_time_seconds=base["seconds"],
                              speedup=base["seconds"] / t_blk if t_blk else None, max_abs_eval_diff=worst)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
