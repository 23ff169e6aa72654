"""Writes the golden fixtures in tests/golden/ from the reference's own code.

Runs the reference's pConstruct_States / sHv / makeH, compiled from the
reference sources by `make -C oracle ref` into oracle/_ref/libedref.so, on
small seeded cases and stores inputs and outputs as .npz data.  The reference
ships no tests or recorded outputs for this path, so these vectors are the
fixtures the oracle and the GPU path are checked against.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from ed_code_amd import lattices as L  # noqa: E402
import pyoracle as P  # noqa: E402


def cases():
    """(name, model, n_ones, kx, ky) -- every case keeps C(N, n) small."""
    out = []
    ch12 = L.chain(12)
    out += [("chain12_n6_k0", ch12, 6, 0, 0), ("chain12_n6_k3", ch12, 6, 3, 0), ("chain12_n7_k5", ch12, 7, 5, 0)]
    ch15 = L.chain(15, phases=[0.3] * 15)
    out += [("chain15ph_n7_k2", ch15, L.n_ones_for_sz(15, 0.0), 2, 0), ("chain15ph_n8_k0", ch15, L.n_ones_for_sz(15, 1.0), 0, 0)]
    sq = L.square(4, 4, J1=1.0, J2=0.5)
    out += [("sq4x4_n8_k00", sq, 8, 0, 0), ("sq4x4_n8_k12", sq, 8, 1, 2), ("sq4x4_n9_k20", sq, 9, 2, 0)]
    tri = L.triangular(4, 3, J1=1.0, Ch=0.3, ChCh=0.2)
    out += [("tri4x3chch_n6_k00", tri, 6, 0, 0), ("tri4x3chch_n6_k11", tri, 6, 1, 1), ("tri4x3chch_n7_k32", tri, 7, 3, 2)]
    tri3 = L.triangular(3, 3, J1=1.0, Ch=0.25)
    out += [("tri3x3ch_n4_k12", tri3, 4, 1, 2), ("tri3x3ch_n5_k00", tri3, 5, 0, 0)]
    tl = L.tilted_square((3, 1), (-1, 3), J1=1.0, J2=0.5, lam=0.7)
    out += [("tilted10_n5_k00", tl, 5, 0, 0), ("tilted10_n5_k17", tl, 5, 1, 7), ("tilted10_n4_k55", tl, 4, 5, 5),
            ("tilted10_n5_k11_empty", tl, 5, 1, 1)]
    # three translation generators (pConstruct_States / sHv 3-D overloads)
    cu = L.cubic(2, 2, 3, J1=1.0, J2=0.3, lam=0.8)
    out += [("cubic2x2x3_n6_k000", cu, 6, (0, 0, 0)), ("cubic2x2x3_n6_k102", cu, 6, (1, 0, 2)),
            ("cubic2x2x3_n5_k011", cu, 5, (0, 1, 1))]
    cu4 = L.cubic(2, 2, 4, J1=1.0, lam=1.1)
    out += [("cubic2x2x4_n8_k013", cu4, 8, (0, 1, 3))]
    return out


def _case(c):
    """(name, model, n, (kx, ky, kz)) from either case form."""
    if len(c) == 5:
        return c[0], c[1], c[2], (c[3], c[4], 0)
    return c[0], c[1], c[2], c[3]


def main():
    if not P.ref_available():
        raise SystemExit("build the reference slice first: make -C oracle ref")
    force = "--force" in sys.argv
    rng = np.random.default_rng(20261203)
    for c in cases():
        name, m, n, (kx, ky, kz) = _case(c)
        path = os.path.join(HERE, f"{name}.npz")
        if os.path.exists(path) and not force:
            # keep the recorded file; advance the generator as writing it did
            d = dict(np.load(path, allow_pickle=False))
            D = int(d["reps"].shape[0])
            rng.normal(size=D); rng.normal(size=D)
            if "unwrap_states" not in d:
                # the reference's Unwrap_pState of the recorded v, added to the file
                st, co = P.RefBasis(m, n, kx, ky, kz).unwrap(d["v"])
                np.savez_compressed(path, **d, unwrap_states=st, unwrap_coef=co)
                print(f"{name}: kept, unwrap added ({len(st)} entries)")
            else:
                print(f"{name}: kept")
            continue
        if P.OracleBasis(m, n, kx, ky, kz).D == 0:
            # pConstruct_States reads pBasis[0] of an empty list in findstate
            # (Momentum_Basis.cpp:228-230,776); the reference exits on empty
            # sectors anyway (Lanczos_sector.cpp:93-97), so none is recorded.
            print(f"{name}: empty sector, skipped")
            continue
        R = P.RefBasis(m, n, kx, ky, kz)
        t = R.tables()
        D = R.D
        v = rng.normal(size=D) + 1j * rng.normal(size=D)
        w = R.hv(v) if D else np.zeros(0, complex)
        H = R.make_h() if D else np.zeros((0, 0), complex)
        ev = np.linalg.eigvalsh(H) if D else np.zeros(0)
        ust, uco = R.unwrap(v)
        a = m.term_arrays()
        extra = {} if m.T3 is None else dict(T3=m.T3, kz=kz)
        np.savez_compressed(
            path, **extra,
            n_sites=m.n_sites, n_ones=n, kx=kx, ky=ky, T1=m.T1, T2=m.T2,
            lam=a["lam"], bond_xy=a["bond_xy"][: a["nbonds"]], bond_J=a["bond_J"][: a["nbonds"]],
            bond_phase=a["bond_phase"][: a["nbonds"]], Ch=a["Ch"], chiral_ijk=a["chiral_ijk"][: a["nchiral"]],
            chiral_phase=a["chiral_phase"][: a["nchiral"]], ChCh=a["ChCh"], bowtie=a["bowtie"][: a["nbowtie"]],
            reps=t["reps"], norm=t["norm"], norm_inv=t["norm_inv"], rep_loc=t["rep_loc"], state_ph=t["state_ph"],
            Ia=t["Ia"], Ib=t["Ib"], bits_right=R.bits_right, v=v, w=w, eigvals=ev, unwrap_states=ust,
            unwrap_coef=uco,
        )
        print(f"{name}: D={D} C={R.C}")


if __name__ == "__main__":
    main()
