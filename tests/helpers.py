"""Shared test helpers: fixtures, brute-force mathematics, the CPU build of
the kernel arithmetic (tests/hostcore)."""
from __future__ import annotations

import ctypes as C
import glob
import itertools
import math
import os
import subprocess

import numpy as np

from ed_code_amd import lattices as L

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden")


# ------------------------------------------------------------ fixtures ----
def golden_files():
    return sorted(glob.glob(os.path.join(GOLDEN, "*.npz")))


def load_golden(path):
    z = np.load(path, allow_pickle=False)
    d = {k: z[k] for k in z.files}
    m = L.Model(os.path.basename(path)[:-4], int(d["n_sites"]), d["T1"].astype(np.int32), d["T2"].astype(np.int32),
                lam=float(d["lam"]))
    m.bonds = [(int(x), int(y), float(J), complex(p[0], p[1]))
               for (x, y), J, p in zip(d["bond_xy"], d["bond_J"], d["bond_phase"])]
    m.Ch = float(d["Ch"])
    m.chiral = [(int(i), int(j), int(k), complex(p[0], p[1])) for (i, j, k), p in zip(d["chiral_ijk"], d["chiral_phase"])]
    m.ChCh = float(d["ChCh"])
    m.bowties = [tuple(int(q) for q in b) for b in d["bowtie"]]
    if "T3" in d:
        m.T3 = d["T3"].astype(np.int32)
    d.setdefault("kz", np.int64(0))
    return m, d


# ------------------------------------------------- brute-force mathematics ----
def compose(a, b):
    """Permutation a after b: site i -> a[b[i]]."""
    return np.asarray(a)[np.asarray(b)]


def group_elements(T1, T2):
    """All distinct elements T1^x T2^y as site permutations."""
    N = len(T1)
    ident = np.arange(N)
    seen = {}
    g1 = ident.copy()
    for _x in range(N + 1):
        g = g1.copy()
        for _y in range(N + 1):
            seen.setdefault(tuple(g), g.copy())
            g = compose(T2, g)
        g1 = compose(T1, g1)
    return list(seen.values())


def apply_perm(s, perm):
    t = 0
    for i, p in enumerate(perm):
        if (s >> i) & 1:
            t |= 1 << int(p)
    return t


def cns_rank(s):
    """Combinatorial-number-system rank: sum_t C(p_t, t)."""
    r, t, p = 0, 0, 0
    while s:
        if s & 1:
            t += 1
            r += math.comb(p, t)
        s >>= 1
        p += 1
    return r


def states(N, n):
    out = []
    for c in itertools.combinations(range(N), n):
        s = 0
        for p in c:
            s |= 1 << p
        out.append(s)
    return sorted(out)


def burnside_dims(model, n):
    """Sector dimension of every (kx, ky) by the character formula
    D_k = (1/|G|) sum over the L1 x L2 grid of chi_k(g) * #{weight-n states fixed by g}
    (the grid covers the group L1*L2/|G| times, which cancels)."""
    L1, L2 = model.orders()
    N = model.n_sites
    T1, T2 = np.asarray(model.T1), np.asarray(model.T2)

    def fixed_count(perm):
        # cycles of perm; fixed states are constant on cycles
        seen = [False] * N
        lens = []
        for i in range(N):
            if not seen[i]:
                l, j = 0, i
                while not seen[j]:
                    seen[j] = True
                    j = int(perm[j])
                    l += 1
                lens.append(l)
        # number of subsets of cycles with total length n
        cnt = np.zeros(N + 1, dtype=object)
        cnt[0] = 1
        for l in lens:
            new = cnt.copy()
            for t in range(N - l, -1, -1):
                new[t + l] += cnt[t]
            cnt = new
        return int(cnt[n])

    fix = {}
    g1 = np.arange(N)
    for a in range(L1):
        g = g1.copy()
        for b in range(L2):
            fix[(a, b)] = fixed_count(g)
            g = compose(T2, g)
        g1 = compose(T1, g1)
    dims = {}
    for kx in range(L1):
        for ky in range(L2):
            tot = 0j
            for (a, b), f in fix.items():
                tot += f * np.exp(-2j * np.pi * (kx * a / L1 + ky * b / L2))
            dims[(kx, ky)] = int(round((tot / (L1 * L2)).real))
    return dims


def full_space_h(model, n):
    """Dense H on all weight-n states, built directly from the term lists
    with no symmetry machinery, in the gather convention of H_v.cpp (row s,
    column t = s with two bits flipped).  Bowtie terms are not included.
    Returns (H, state list)."""
    N = model.n_sites
    st = states(N, n)
    idx = {s: i for i, s in enumerate(st)}
    H = np.zeros((len(st), len(st)), complex)
    for i, s in enumerate(st):
        for (x, y, J, ph) in model.bonds:
            bx, by = (s >> x) & 1, (s >> y) & 1
            H[i, i] += model.lam * J * (0.25 if bx == by else -0.25)
            if bx != by:
                t = s ^ (1 << x) ^ (1 << y)
                H[i, idx[t]] += 0.5 * J * (np.conj(ph) if bx else ph)
        if abs(model.Ch) > 1e-16:
            for (a, b, k, ep) in model.chiral:
                ba, bb = (s >> a) & 1, (s >> b) & 1
                if ba != bb:
                    t = s ^ (1 << a) ^ (1 << b)
                    sg = 1j * ep if bb else -1j * np.conj(ep)
                    if not (s >> k) & 1:
                        sg = -sg
                    H[i, idx[t]] += model.Ch * 0.25 * sg
    return H, st


def scalar_chirality_matrix(N, st, tri):
    """S_i . (S_j x S_k) = sum_abc eps_abc S^a_i S^b_j S^c_k on the weight-n
    states st, built from the spin-1/2 matrices (bit 1 = up) with no use of
    the Hamiltonian's term lists: column t is the operator applied to |t>."""
    import itertools
    sx = np.array([[0, 1], [1, 0]]) / 2
    sy = np.array([[0, 1j], [-1j, 0]]) / 2
    sz = np.array([[-1, 0], [0, 1]]) / 2
    S = [sx, sy, sz]
    M = np.zeros((8, 8), complex)
    for a, b, c in itertools.permutations(range(3)):
        M += np.linalg.det(np.eye(3)[[a, b, c]]) * np.kron(np.kron(S[a], S[b]), S[c])
    M = M.reshape((2,) * 6)
    ax = [N - 1 - tri[0], N - 1 - tri[1], N - 1 - tri[2]]   # axis of site s in the (2,)*N view
    X = np.zeros((len(st), len(st)), complex)
    for col, t in enumerate(st):
        e = np.zeros((2,) * N, complex)
        e.reshape(-1)[t] = 1
        y = np.moveaxis(np.tensordot(M, np.moveaxis(e, ax, [0, 1, 2]), axes=([3, 4, 5], [0, 1, 2])), [0, 1, 2], ax)
        X[:, col] = y.reshape(-1)[st]
    return X


def momentum_projector(model, n, kx, ky, st):
    """Orthonormal momentum states as columns over the full weight-n space:
    |r(k)> ~ sum_g exp(-i k.g) g|r> over the translation grid."""
    L1, L2 = model.orders()
    px, py = model.momentum(kx, ky)
    T1, T2 = np.asarray(model.T1), np.asarray(model.T2)
    idx = {s: i for i, s in enumerate(st)}
    grid = []
    g1 = np.arange(model.n_sites)
    for a in range(L1):
        g = g1.copy()
        for b in range(L2):
            grid.append((a, b, g.copy()))
            g = compose(T2, g)
        g1 = compose(T1, g1)
    cols, done = [], set()
    for s in st:
        orb = {apply_perm(s, g) for (_, _, g) in grid}
        r = min(orb)
        if r in done:
            continue
        done.add(r)
        vec = np.zeros(len(st), complex)
        for (a, b, g) in grid:
            vec[idx[apply_perm(r, g)]] += np.exp(-1j * (px * a + py * b))
        nrm = np.linalg.norm(vec)
        if nrm > 1e-8:
            cols.append(vec / nrm)
    return np.array(cols).T if cols else np.zeros((len(st), 0), complex)


# --------------------------------------------------- CPU build of ed_core ----
_hc = None


def _build_locked(out, deps, cmd):
    """Builds `out` with `cmd + ["-o", tmp]` when it is missing or older than
    a dependency: under a file lock (parallel test workers build it once),
    into a temporary file renamed into place (no worker loads a partial
    file)."""
    import fcntl
    os.makedirs(os.path.dirname(out), exist_ok=True)
    stale = lambda: not os.path.exists(out) or any(os.path.getmtime(f) > os.path.getmtime(out) for f in deps)
    if not stale():
        return
    with open(out + ".lock", "w") as lk:
        fcntl.flock(lk, fcntl.LOCK_EX)
        if stale():
            tmp = f"{out}.{os.getpid()}.tmp"
            subprocess.run(cmd + ["-o", tmp], check=True)
            os.replace(tmp, out)


def refapi_driver():
    """tests/refapi/driver.cpp (a caller of the entry points of
    include/ed_refapi.h that the reference's main does not reach) linked
    against ed_code_amd/libed_amd.so."""
    out = os.path.join(HERE, "refapi", "_build", "driver")
    src = os.path.join(HERE, "refapi", "driver.cpp")
    lib = os.path.join(ROOT, "ed_code_amd", "libed_amd.so")
    hdr = os.path.join(ROOT, "include", "ed_refapi.h")
    _build_locked(out, [src, hdr], ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), src,
                                    "-L", os.path.dirname(lib), "-led_amd", "-Wl,-rpath,$ORIGIN/../../../ed_code_amd"])
    return out


def write_spec(path, model, n, k, nev=1, iterations=100, ncycles=10, tol=1e-10):
    """Input of the refapi driver: sizes and solver settings, T1, T2, T3 (the
    identity for two generators), bonds."""
    k = tuple(k) + (0,) * (3 - len(k))
    T3 = model.T3 if model.T3 is not None else np.arange(model.n_sites)
    with open(path, "w") as f:
        f.write(f"{model.n_sites} {n} {k[0]} {k[1]} {k[2]} {model.lam!r} {iterations} {ncycles} {nev} {tol!r}\n")
        for T in (model.T1, model.T2, T3):
            f.write(" ".join(str(int(t)) for t in T) + "\n")
        f.write(f"{len(model.bonds)}\n")
        for x, y, J, ph in model.bonds:
            f.write(f"{x} {y} {J!r} {ph.real!r} {ph.imag!r}\n")


def hostcore():
    """tests/hostcore/hostcore.cpp + ed_core.h compiled with g++ (CPU)."""
    global _hc
    if _hc is None:
        out = os.path.join(HERE, "hostcore", "_build", "libhostcore.so")
        src = os.path.join(HERE, "hostcore", "hostcore.cpp")
        hdrs = glob.glob(os.path.join(ROOT, "ed_code_amd", "csrc", "*.h"))
        _build_locked(out, hdrs + [src], ["g++", "-O2", "-fopenmp", "-fPIC", "-shared", "-std=c++17",
                                          "-I", os.path.join(ROOT, "ed_code_amd", "csrc"), src])
        H = C.CDLL(out)
        vp = C.c_void_p
        H.hc_build.restype = vp
        H.hc_build.argtypes = [C.c_int, C.c_int, vp, vp, C.c_int, C.c_int]
        H.hc_build3.restype = vp
        H.hc_build3.argtypes = [C.c_int, C.c_int, vp, vp, vp, C.c_int, C.c_int, C.c_int]
        H.hc_free.argtypes = [vp]
        H.hc_dim.restype = C.c_int64
        H.hc_dim.argtypes = [vp]
        H.hc_export.argtypes = [vp] * 4
        H.hc_full_tables.argtypes = [vp] * 3
        H.hc_set_terms.argtypes = [vp, vp]
        H.hc_hv.argtypes = [vp] * 3
        H.hc_rep64_check.restype = C.c_int
        H.hc_rep64_check.argtypes = [C.c_int64, C.c_uint64]
        H.hc_ranker_new.restype = vp
        H.hc_ranker_new.argtypes = [C.c_int, C.c_int]
        H.hc_ranker_free.argtypes = [vp]
        H.hc_gosper_next.restype = C.c_uint64
        H.hc_gosper_next.argtypes = [C.c_uint64]
        H.hc_rank.restype = C.c_uint64
        H.hc_rank.argtypes = [vp, C.c_uint64]
        H.hc_unrank.restype = C.c_uint64
        H.hc_unrank.argtypes = [vp, C.c_uint64]
        H.hc_hv_torus.restype = C.c_int
        H.hc_hv_torus.argtypes = [vp] * 3
        H.hc_symeig.restype = C.c_int
        H.hc_symeig.argtypes = [C.c_int, vp, vp, vp]
        H.hc_torus_mode.restype = C.c_int
        H.hc_torus_mode.argtypes = [vp]
        H.hc_hv_torus_p.restype = C.c_int
        H.hc_hv_torus_p.argtypes = [vp] * 3
        H.hc_hv_torus_s.restype = C.c_int
        H.hc_hv_torus_s.argtypes = [vp, vp, vp, C.c_int]
        H.hc_torus_rep.restype = C.c_int
        H.hc_torus_rep.argtypes = [vp, C.c_uint32, vp, vp]
        H.hc_generic_rep.argtypes = [vp, C.c_uint64, vp, vp]
        H.hc_hv_torus_c.restype = C.c_int
        H.hc_hv_torus_c.argtypes = [vp] * 3
        H.hc_torus_rep_c.restype = C.c_int
        H.hc_torus_rep_c.argtypes = [vp, C.c_uint32, vp, vp]
        H.hc_torus_rep_c_any.restype = C.c_int
        H.hc_torus_rep_c_any.argtypes = [vp, C.c_uint32, vp, vp]
        H.hc_phase.argtypes = [vp, C.c_int, vp, vp]
        H.hc_rank_split.restype = C.c_uint64
        H.hc_rank_split.argtypes = [vp, C.c_uint64]
        H.hc_hv_rows2.restype = C.c_int
        H.hc_hv_rows2.argtypes = [vp] * 3
        H.hc_hv_rows2w.restype = C.c_int
        H.hc_hv_rows2w.argtypes = [vp, vp, vp, C.c_int]
        H.hc_tridiag.restype = C.c_int
        H.hc_tridiag.argtypes = [C.c_int, vp, vp, vp]
        _hc = H
    return _hc


class HostCore:
    def __init__(self, model, n, kx=0, ky=0, kz=0):
        H = hostcore()
        self.T1 = np.ascontiguousarray(model.T1, np.int32)
        self.T2 = np.ascontiguousarray(model.T2, np.int32)
        self.T3 = None if model.T3 is None else np.ascontiguousarray(model.T3, np.int32)
        self.h = H.hc_build3(model.n_sites, n, self.T1.ctypes.data, self.T2.ctypes.data,
                             None if self.T3 is None else self.T3.ctypes.data, kx, ky, kz)
        self.D = H.hc_dim(self.h)
        import pyoracle
        self.terms, self._keep = pyoracle.make_terms(model)
        H.hc_set_terms(self.h, C.addressof(self.terms))

    def __del__(self):
        try:
            hostcore().hc_free(self.h)
        except Exception:
            pass

    def export(self):
        reps = np.zeros(self.D, np.int64); nm = np.zeros(self.D); ni = np.zeros(self.D)
        hostcore().hc_export(self.h, reps.ctypes.data, nm.ctypes.data, ni.ctypes.data)
        return reps, nm, ni

    def full_tables(self, C_):
        rl = np.zeros(C_, np.int64); sp = np.zeros(2 * C_)
        hostcore().hc_full_tables(self.h, rl.ctypes.data, sp.ctypes.data)
        return rl, sp.view(np.complex128)

    def hv(self, v):
        v = np.ascontiguousarray(v, np.complex128)
        w = np.zeros(self.D, np.complex128)
        hostcore().hc_hv(self.h, v.ctypes.data, w.ctypes.data)
        return w

    def hv_torus(self, v):
        """Rotation fast path; None if the lattice does not fit it."""
        v = np.ascontiguousarray(v, np.complex128)
        w = np.zeros(self.D, np.complex128)
        if hostcore().hc_hv_torus(self.h, v.ctypes.data, w.ctypes.data):
            return None
        return w

    def hv_torus_c(self, v, pipelined=False):
        """Compiled-shape batched row function (or its software-pipelined
        form); None if no compiled shape fits."""
        v = np.ascontiguousarray(v, np.complex128)
        w = np.zeros(self.D, np.complex128)
        fn = hostcore().hc_hv_torus_p if pipelined else hostcore().hc_hv_torus_c
        if fn(self.h, v.ctypes.data, w.ctypes.data):
            return None
        return w

    def hv_torus_split(self, v, lanes):
        """Compiled-shape row split over `lanes` lanes, summed in the split
        kernels' tree order; None if no compiled shape fits."""
        v = np.ascontiguousarray(v, np.complex128)
        w = np.zeros(self.D, np.complex128)
        if hostcore().hc_hv_torus_s(self.h, v.ctypes.data, w.ctypes.data, lanes):
            return None
        return w

    def hv_rows2(self, v, wide=True):
        """Two-row 64-bit row function (ed_torus64.h), gathering from the wide
        image (or by row through the directory); None if the lattice is not a
        compiled two-row shape."""
        v = np.ascontiguousarray(v, np.complex128)
        w = np.zeros(self.D, np.complex128)
        if hostcore().hc_hv_rows2w(self.h, v.ctypes.data, w.ctypes.data, int(wide)):
            return None
        return w

    def torus_rep(self, t):
        rep, code = C.c_uint32(), C.c_int()
        if hostcore().hc_torus_rep(self.h, t, C.addressof(rep), C.addressof(code)):
            return None
        return rep.value, code.value

    def torus_rep_c(self, t):
        rep, code = C.c_uint32(), C.c_int()
        if hostcore().hc_torus_rep_c(self.h, t, C.addressof(rep), C.addressof(code)):
            return None
        return rep.value, code.value

    def torus_rep_c_any(self, t):
        """The kernels' search: representative and the code of some
        minimising translation."""
        rep, code = C.c_uint32(), C.c_int()
        if hostcore().hc_torus_rep_c_any(self.h, t, C.addressof(rep), C.addressof(code)):
            return None
        return rep.value, code.value

    def phase(self, code):
        re, im = C.c_double(), C.c_double()
        hostcore().hc_phase(self.h, code, C.addressof(re), C.addressof(im))
        return complex(re.value, im.value)

    def generic_rep(self, t):
        rep, code = C.c_uint64(), C.c_int()
        hostcore().hc_generic_rep(self.h, t, C.addressof(rep), C.addressof(code))
        return rep.value, code.value
