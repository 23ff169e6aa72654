"""Lattice and Hamiltonian-term generators.

A `Model` carries what the reference's input parser produces
(read_input_file.cpp:110-430): the translation permutations T1/T2 (bit i of a
state moves to bit T[i]), the XXZ bond list with per-bond J and phase
(`coordinates`, global.h:25-30), the cyclic chiral entries (`triangles`,
global.h:32-35, three per input triangle, read_input_file.cpp:262-273) and
the chirality-chirality expansion (`bowties`, global.h:37-40,
read_input_file.cpp:286-319).  Sites are labelled x + Lx*y on rectangular
tori; the tilted 40-site cluster uses the canonical labelling of
`tilted_square`.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np


@dataclass
class Model:
    name: str
    n_sites: int
    T1: np.ndarray
    T2: np.ndarray
    lam: float = 1.0                      # Jz (Delta): scales only S^z S^z
    bonds: List[Tuple[int, int, float, complex]] = field(default_factory=list)
    Ch: float = 0.0
    chiral: List[Tuple[int, int, int, complex]] = field(default_factory=list)
    ChCh: float = 0.0
    bowties: List[Tuple[int, int, int, int, int, int]] = field(default_factory=list)
    triangles: List[Tuple[int, int, int]] = field(default_factory=list)
    T3: Optional[np.ndarray] = None       # third translation (3-D clusters), else None

    # ---------------------------------------------------------------- arrays
    def term_arrays(self):
        """Struct-of-arrays form used by the C ABI (ed_terms)."""
        nb = len(self.bonds)
        bond_xy = np.zeros((max(nb, 1), 2), np.int32)
        bond_J = np.zeros(max(nb, 1), np.float64)
        bond_ph = np.zeros((max(nb, 1), 2), np.float64)
        for k, (x, y, J, ph) in enumerate(self.bonds):
            bond_xy[k] = (x, y)
            bond_J[k] = J
            bond_ph[k] = (ph.real, ph.imag)
        nc = len(self.chiral)
        ch_ijk = np.zeros((max(nc, 1), 3), np.int32)
        ch_ph = np.zeros((max(nc, 1), 2), np.float64)
        for k, (i, j, kk, ph) in enumerate(self.chiral):
            ch_ijk[k] = (i, j, kk)
            ch_ph[k] = (ph.real, ph.imag)
        nw = len(self.bowties)
        bw = np.zeros((max(nw, 1), 6), np.int32)
        for k, b in enumerate(self.bowties):
            bw[k] = b
        return dict(lam=float(self.lam), nbonds=nb, bond_xy=bond_xy, bond_J=bond_J, bond_phase=bond_ph,
                    Ch=float(self.Ch), nchiral=nc, chiral_ijk=ch_ijk, chiral_phase=ch_ph,
                    ChCh=float(self.ChCh), nbowtie=nw, bowtie=bw)

    def orders(self) -> Tuple[int, int]:
        """(L1, L2): cycle length of site 0 under T1 and T2 (Momentum_Basis.cpp:19-36)."""
        def cyc(T):
            s, n = 0, 0
            while True:
                s = int(T[s]); n += 1
                if s == 0:
                    return n
        return cyc(self.T1), cyc(self.T2)

    def order3(self) -> int:
        """L3 (LxLyLz, Momentum_Basis.cpp:38-64); 1 without a third translation."""
        if self.T3 is None:
            return 1
        s, n = 0, 0
        while True:
            s = int(self.T3[s]); n += 1
            if s == 0:
                return n

    def momentum(self, kx: int, ky: int) -> Tuple[float, float]:
        """(px, py) exactly as Lanczos_sector.cpp:57 forms them."""
        L1, L2 = self.orders()
        px = float(2 * kx) * math.pi / L1
        py = float(2 * ky) * math.pi / L2
        return px, py

    def with_couplings(self, **kw) -> "Model":
        import copy
        m = copy.deepcopy(self)
        for k, v in kw.items():
            setattr(m, k, v)
        return m


def n_ones_for_sz(n_sites: int, cSz: float) -> int:
    """Lanczos_sector.cpp:45-46: int(cSz + n_spins/2 + 1e-6) with integer n_spins/2."""
    return int(cSz + (n_sites // 2) + 1.0e-6)


def bond_phase_of(bonds, i: int, j: int) -> complex:
    """get_expphase (read_input_file.cpp:3-14): phase of bond (i,j), conj if
    stored as (j,i), 1 if absent.  First match in list order wins."""
    for (x, y, _J, ph) in bonds:
        if i == x and j == y:
            return ph
        if i == y and j == x:
            return ph.conjugate()
    return complex(1.0, 0.0)


def chiral_entries(triangles, bonds):
    """Three cyclic entries per triangle (read_input_file.cpp:262-273)."""
    out = []
    for (x, y, z) in triangles:
        out.append((x, y, z, bond_phase_of(bonds, x, y)))
        out.append((y, z, x, bond_phase_of(bonds, y, z)))
        out.append((z, x, y, bond_phase_of(bonds, z, x)))
    return out


def _neighbor_tri(t1, t2, bonds) -> bool:
    """neighbor_tri (read_input_file.cpp:16-45): some bond joins a vertex of t1 to one of t2."""
    s1, s2 = set(t1), set(t2)
    for (x, y, _J, _ph) in bonds:
        if (x in s1 and y in s2) or (y in s1 and x in s2):
            return True
    return False


def bowtie_entries(triangles, bonds):
    """Bowtie expansion (read_input_file.cpp:286-319): 9 entries per ordered
    pair of distinct neighbouring triangles."""
    out = []
    for m, t1 in enumerate(triangles):
        for n, t2 in enumerate(triangles):
            if m == n or not _neighbor_tri(t1, t2, bonds):
                continue
            i1, j1, k1 = t1
            i2, j2, k2 = t2
            for (x, y, z) in ((i1, j1, k1), (k1, i1, j1), (j1, k1, i1)):
                for (p, q, r) in ((i2, j2, k2), (k2, i2, j2), (j2, k2, i2)):
                    out.append((x, y, z, p, q, r))
    return out


# -------------------------------------------------------------- relabelling
@dataclass
class Relabel:
    """Site relabelling of a cluster whose translation group acts regularly:
    new label of old site s = i + n2*j where s = a^i b^j (site 0), with a of
    order n2 (the group's exponent) and b of order n1 = N/n2 generating a
    complement.  a = T1^xa T2^ya, b = T1^xb T2^yb in the original generators."""
    sigma: np.ndarray            # old site -> new site
    n2: int
    n1: int
    xa: int
    ya: int
    xb: int
    yb: int
    L1: int                      # orders of the original T1, T2
    L2: int
    t1: Tuple[int, int]          # T1 = a^i b^j, T2 = a^i' b^j'
    t2: Tuple[int, int]

    def momentum(self, kx: int, ky: int):
        """The relabelled model's (kx', ky') carrying the same character,
        exp(-i p'.(i, j)) = exp(-i p.(x, y)) for a^i b^j = T1^x T2^y; None
        when (kx, ky) is not a character of the group (the reference's sector
        is then empty, Lanczos_sector.cpp:93-97)."""
        fa = self.n2 * (kx * self.xa * self.L2 + ky * self.ya * self.L1)
        fb = self.n1 * (kx * self.xb * self.L2 + ky * self.yb * self.L1)
        den = self.L1 * self.L2
        if fa % den or fb % den:
            return None
        k2 = ((fa // den) % self.n2, (fb // den) % self.n1)
        # the character must also reproduce exp(-i px), exp(-i py) on T1, T2:
        # k/L == k2[0]*i/n2 + k2[1]*j/n1 (mod 1)
        for k, Lk, (i, j) in ((kx, self.L1, self.t1), (ky, self.L2, self.t2)):
            lhs = k * self.n2 * self.n1
            rhs = (k2[0] * i * self.n1 + k2[1] * j * self.n2) * Lk
            if (lhs - rhs) % (Lk * self.n2 * self.n1):
                return None
        return k2


def _perm_pow(T: np.ndarray, e: int) -> np.ndarray:
    g = np.arange(len(T))
    for _ in range(e):
        g = T[g]
    return g


def regular_relabel(model: Model) -> Tuple[Model, Relabel]:
    """Relabels the sites so that T1' is a rotation inside rows of n2 sites
    and T2' moves whole rows (the layout of a rectangular torus).  On a
    cluster whose translation group acts regularly (|G| = N, e.g. the tilted
    40-site square cluster, group Z2 x Z20) every translation then is a bit
    rotation, which the rotation fast path needs.  Energies are unchanged;
    representatives and vectors are those of the relabelled model."""
    N = model.n_sites
    T1 = np.asarray(model.T1, np.int64)
    T2 = np.asarray(model.T2, np.int64)
    L1, L2 = model.orders()
    elems = {}
    for x in range(L1):
        for y in range(L2):
            g = _perm_pow(T1, x)[_perm_pow(T2, y)]
            elems.setdefault(tuple(g.tolist()), (x, y))
    if len(elems) != N:
        raise ValueError(f"translation group of order {len(elems)} does not act regularly on {N} sites")
    perms = {k: np.array(k) for k in elems}

    def order(g):
        e, h = 1, g.copy()
        ident = np.arange(N)
        while not np.array_equal(h, ident):
            h = g[h]
            e += 1
        return e

    ords = {k: order(v) for k, v in perms.items()}
    n2 = max(ords.values())
    ka = min((k for k in perms if ords[k] == n2), key=lambda k: elems[k])
    a = perms[ka]
    cyc_a = set()
    h = np.arange(N)
    for _ in range(n2):
        cyc_a.add(tuple(h.tolist()))
        h = a[h]
    n1 = N // n2
    kb = None
    for k in sorted(perms, key=lambda k: elems[k]):
        if ords[k] != n1:
            continue
        g, ok = np.arange(N), True
        for e in range(1, n1):
            g = perms[k][g]
            if tuple(g.tolist()) in cyc_a:
                ok = False
                break
        if ok:
            kb = k
            break
    if kb is None:
        raise ValueError("no complement to a maximal cyclic subgroup")
    b = perms[kb]
    sigma = np.full(N, -1, np.int64)
    gj = np.arange(N)
    for j in range(n1):
        gi = gj.copy()
        for i in range(n2):
            sigma[gi[0]] = i + n2 * j
            gi = a[gi]
        gj = b[gj]
    if (sigma < 0).any():
        raise ValueError("relabelling does not cover every site")
    T1n = np.zeros(N, np.int32)
    T2n = np.zeros(N, np.int32)
    for j in range(n1):
        for i in range(n2):
            T1n[i + n2 * j] = (i + 1) % n2 + n2 * j
            T2n[i + n2 * j] = i + n2 * ((j + 1) % n1)
    S = lambda q: int(sigma[q])
    m = Model(model.name + "_rl", N, T1n, T2n, lam=model.lam,
              bonds=[(S(x), S(y), J, ph) for (x, y, J, ph) in model.bonds], Ch=model.Ch,
              chiral=[(S(i), S(j), S(k), ph) for (i, j, k, ph) in model.chiral], ChCh=model.ChCh,
              bowties=[tuple(S(q) for q in bw) for bw in model.bowties],
              triangles=[(S(i), S(j), S(k)) for (i, j, k) in model.triangles])
    (xa, ya), (xb, yb) = elems[ka], elems[kb]
    # coordinates of T1 and T2 in the new generators: where they send site 0
    def coords(T):
        q = int(sigma[int(T[0])])
        return q % n2, q // n2
    return m, Relabel(sigma, n2, n1, xa, ya, xb, yb, L1, L2, coords(T1), coords(T2))


# ----------------------------------------------------------------- lattices
def chain(L: int, J: float = 1.0, lam: float = 1.0, phases: Sequence[float] | None = None) -> Model:
    """Periodic chain; T1: i -> i+1, T2 = identity (so L2 = 1)."""
    T1 = np.array([(i + 1) % L for i in range(L)], np.int32)
    T2 = np.arange(L, dtype=np.int32)
    bonds = []
    for i in range(L):
        ph = 0.0 if phases is None else phases[i]
        bonds.append((i, (i + 1) % L, J, complex(math.cos(ph), math.sin(ph)) if ph else complex(1.0, 0.0)))
    return Model(f"chain{L}", L, T1, T2, lam=lam, bonds=bonds)


def _torus_perms(Lx: int, Ly: int):
    N = Lx * Ly
    T1 = np.zeros(N, np.int32)
    T2 = np.zeros(N, np.int32)
    for y in range(Ly):
        for x in range(Lx):
            s = x + Lx * y
            T1[s] = (x + 1) % Lx + Lx * y
            T2[s] = x + Lx * ((y + 1) % Ly)
    return T1, T2


def square(Lx: int, Ly: int, J1: float = 1.0, J2: float = 0.0, lam: float = 1.0) -> Model:
    """Lx x Ly square torus; J1 on x/y bonds, J2 on both diagonals (phase 1)."""
    N = Lx * Ly
    T1, T2 = _torus_perms(Lx, Ly)
    site = lambda x, y: (x % Lx) + Lx * (y % Ly)
    bonds = []
    for y in range(Ly):
        for x in range(Lx):
            bonds.append((site(x, y), site(x + 1, y), J1, complex(1.0, 0.0)))
            bonds.append((site(x, y), site(x, y + 1), J1, complex(1.0, 0.0)))
    if J2 != 0.0:
        for y in range(Ly):
            for x in range(Lx):
                bonds.append((site(x, y), site(x + 1, y + 1), J2, complex(1.0, 0.0)))
                bonds.append((site(x, y), site(x + 1, y - 1), J2, complex(1.0, 0.0)))
    return Model(f"sq{Lx}x{Ly}", N, T1, T2, lam=lam, bonds=bonds)


def triangular(Lx: int, Ly: int, J1: float = 1.0, lam: float = 1.0, Ch: float = 0.0,
               ChCh: float = 0.0, J2: float = 0.0) -> Model:
    """Lx x Ly triangular torus with a1=(1,0), a2=(0,1) at 60 degrees.
    Nearest-neighbour bonds along a1, a2, a2-a1; up (r, r+a1, r+a2) and down
    (r+a1, r+a1+a2, r+a2) triangles, both counter-clockwise.  Optional J2 on
    the second-neighbour directions a1+a2, 2a2-a1, a2-2a1."""
    N = Lx * Ly
    T1, T2 = _torus_perms(Lx, Ly)
    site = lambda x, y: (x % Lx) + Lx * (y % Ly)
    bonds = []
    for y in range(Ly):
        for x in range(Lx):
            s = site(x, y)
            bonds.append((s, site(x + 1, y), J1, complex(1.0, 0.0)))
            bonds.append((s, site(x, y + 1), J1, complex(1.0, 0.0)))
            bonds.append((s, site(x - 1, y + 1), J1, complex(1.0, 0.0)))
    if J2 != 0.0:
        for y in range(Ly):
            for x in range(Lx):
                s = site(x, y)
                bonds.append((s, site(x + 1, y + 1), J2, complex(1.0, 0.0)))
                bonds.append((s, site(x - 1, y + 2), J2, complex(1.0, 0.0)))
                bonds.append((s, site(x - 2, y + 1), J2, complex(1.0, 0.0)))
    tris = []
    for y in range(Ly):
        for x in range(Lx):
            tris.append((site(x, y), site(x + 1, y), site(x, y + 1)))
            tris.append((site(x + 1, y), site(x + 1, y + 1), site(x, y + 1)))
    m = Model(f"tri{Lx}x{Ly}", N, T1, T2, lam=lam, bonds=bonds, triangles=tris)
    m.Ch = Ch
    m.chiral = chiral_entries(tris, bonds)
    m.ChCh = ChCh
    if ChCh != 0.0:
        m.bowties = bowtie_entries(tris, bonds)
    return m


def cubic(Lx: int, Ly: int, Lz: int, J1: float = 1.0, J2: float = 0.0, lam: float = 1.0) -> Model:
    """Lx x Ly x Lz simple-cubic torus, site x + Lx*(y + Ly*z); T1, T2, T3
    translate by one along x, y, z.  J1 on nearest neighbours, J2 on the
    face diagonals of the xy planes."""
    N = Lx * Ly * Lz
    site = lambda x, y, z: (x % Lx) + Lx * ((y % Ly) + Ly * (z % Lz))
    T1 = np.zeros(N, np.int32)
    T2 = np.zeros(N, np.int32)
    T3 = np.zeros(N, np.int32)
    bonds = []
    for z in range(Lz):
        for y in range(Ly):
            for x in range(Lx):
                s = site(x, y, z)
                T1[s], T2[s], T3[s] = site(x + 1, y, z), site(x, y + 1, z), site(x, y, z + 1)
                bonds.append((s, site(x + 1, y, z), J1, complex(1.0, 0.0)))
                bonds.append((s, site(x, y + 1, z), J1, complex(1.0, 0.0)))
                bonds.append((s, site(x, y, z + 1), J1, complex(1.0, 0.0)))
                if J2 != 0.0:
                    bonds.append((s, site(x + 1, y + 1, z), J2, complex(1.0, 0.0)))
                    bonds.append((s, site(x + 1, y - 1, z), J2, complex(1.0, 0.0)))
    return Model(f"cubic{Lx}x{Ly}x{Lz}", N, T1, T2, lam=lam, bonds=bonds, T3=T3)


def tilted_square(P1: Tuple[int, int], P2: Tuple[int, int], J1: float = 1.0, J2: float = 0.0,
                  lam: float = 1.0) -> Model:
    """Square-lattice cluster with periodic vectors P1, P2 (N = |det|).
    Sites are the classes of Z^2 modulo the period lattice; each class is
    labelled by the first point met scanning y = 0, 1, ... and x = 0, 1, ...
    T1/T2 translate by (1,0)/(0,1).  J1 on nearest neighbours, J2 on both
    diagonals."""
    (a, b), (c, d) = P1, P2
    N = abs(a * d - b * c)
    det = a * d - b * c

    def key(x, y):
        # coordinates in the period basis, scaled by det, reduced mod det
        u = (d * x - c * y) % abs(det)
        v = (-b * x + a * y) % abs(det)
        return (u, v)

    labels = {}
    R = 4 * (abs(a) + abs(b) + abs(c) + abs(d)) + 4
    for y in range(0, R):
        for x in range(0, R):
            k = key(x, y)
            if k not in labels:
                labels[k] = len(labels)
            if len(labels) == N:
                break
        if len(labels) == N:
            break
    assert len(labels) == N
    coords = {}
    for y in range(0, R):
        for x in range(0, R):
            k = key(x, y)
            if labels[k] not in coords:
                coords[labels[k]] = (x, y)
    site = lambda x, y: labels[key(x, y)]
    T1 = np.array([site(coords[s][0] + 1, coords[s][1]) for s in range(N)], np.int32)
    T2 = np.array([site(coords[s][0], coords[s][1] + 1) for s in range(N)], np.int32)
    bonds = []
    for s in range(N):
        x, y = coords[s]
        bonds.append((s, site(x + 1, y), J1, complex(1.0, 0.0)))
        bonds.append((s, site(x, y + 1), J1, complex(1.0, 0.0)))
    if J2 != 0.0:
        for s in range(N):
            x, y = coords[s]
            bonds.append((s, site(x + 1, y + 1), J2, complex(1.0, 0.0)))
            bonds.append((s, site(x + 1, y - 1), J2, complex(1.0, 0.0)))
    return Model(f"tilted{N}", N, T1, T2, lam=lam, bonds=bonds)


def by_name(name: str, **kw) -> Model:
    """Named lattices of the benchmark workloads."""
    if name == "chain24":
        return chain(24, **kw)
    if name == "sq4x8":
        return square(4, 8, J1=kw.get("J1", 1.0), J2=kw.get("J2", 0.5), lam=kw.get("lam", 1.0))
    if name == "tri6x5":
        return triangular(6, 5, Ch=kw.get("Ch", 0.0), lam=kw.get("lam", 1.0), J2=kw.get("J2", 0.0))
    if name == "tri3x9":
        return triangular(3, 9, Ch=kw.get("Ch", 0.0), lam=kw.get("lam", 1.0), J2=kw.get("J2", 0.0))
    if name == "tilted40":
        return tilted_square((6, 2), (-2, 6), J1=kw.get("J1", 1.0), J2=kw.get("J2", 0.5), lam=kw.get("lam", 1.0))
    raise KeyError(name)
