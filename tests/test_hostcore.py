"""The kernel arithmetic (ed_code_amd/csrc/ed_core.h) compiled for the CPU,
checked against the oracle and the golden fixtures.  These run the same
functions the gfx950 kernels run, so logic errors surface without a GPU."""
import itertools
import math
import os

import numpy as np
import pytest

import helpers as Hp
import pyoracle as P
from ed_code_amd import lattices as L

GOLDEN = Hp.golden_files()


@pytest.mark.parametrize("path", GOLDEN, ids=[p.split("/")[-1][:-4] for p in GOLDEN])
def test_core_tables_and_hv_vs_golden(path):
    m, d = Hp.load_golden(path)
    n = int(d["n_ones"])
    h = Hp.HostCore(m, n, int(d["kx"]), int(d["ky"]), int(d["kz"]))
    reps, nm, ni = h.export()
    assert np.array_equal(reps, d["reps"])
    assert np.array_equal(nm, d["norm"]) and np.array_equal(ni, d["norm_inv"])
    rl, sp = h.full_tables(len(d["rep_loc"]))
    assert np.array_equal(rl, d["rep_loc"])
    assert np.array_equal(sp.view(np.float64), d["state_ph"].view(np.float64))
    w = h.hv(d["v"])
    assert np.linalg.norm(w - d["w"]) <= 1e-12 * np.linalg.norm(d["w"])


@pytest.mark.parametrize("model,n,k", [(L.chain(20), 10, (0, 0)), (L.chain(19), 9, (4, 0)),
                                       (L.triangular(4, 4, Ch=0.2), 8, (1, 3)), (L.square(4, 4, J2=0.5), 8, (2, 2)),
                                       (L.tilted_square((3, 1), (-1, 3), J2=0.5), 4, (5, 5))])
def test_core_vs_oracle(model, n, k):
    o = P.OracleBasis(model, n, *k)
    h = Hp.HostCore(model, n, *k)
    t = o.tables()
    reps, nm, ni = h.export()
    assert np.array_equal(reps, t["reps"]) and np.array_equal(nm, t["norm"])
    rl, sp = h.full_tables(o.C)
    assert np.array_equal(rl, t["rep_loc"])
    assert np.array_equal(sp.view(np.float64), t["state_ph"].view(np.float64))
    v = np.random.default_rng(2).normal(size=o.D) + 1j * np.random.default_rng(3).normal(size=o.D)
    assert np.linalg.norm(h.hv(v) - o.hv(v)) <= 1e-12 * np.linalg.norm(o.hv(v))


@pytest.mark.parametrize("N,n", [(24, 12), (32, 16), (40, 20), (27, 13), (30, 15), (48, 40), (63, 31)])
def test_rank_unrank_closed_form(N, n):
    """Byte-sliced rank == sum_t C(p_t, t); unrank inverts it; Gosper's next
    word has rank + 1."""
    h = Hp.hostcore()
    rk = h.hc_ranker_new(N, n)
    rng = np.random.default_rng(N)
    try:
        for _ in range(400):
            bits = rng.choice(N, size=n, replace=False)
            s = int(sum(1 << int(b) for b in bits))
            r = h.hc_rank(rk, s)
            assert r == Hp.cns_rank(s)
            assert h.hc_unrank(rk, r) == s
            if r + 1 < math.comb(N, n):
                assert h.hc_rank(rk, h.hc_gosper_next(s)) == r + 1
        assert h.hc_rank(rk, (1 << n) - 1) == 0
        top = ((1 << n) - 1) << (N - n)
        assert h.hc_rank(rk, top) == math.comb(N, n) - 1
    finally:
        h.hc_ranker_free(rk)


def test_tridiagonal_ql_vs_dense():
    import ctypes as C
    rng = np.random.default_rng(5)
    for m in (1, 2, 7, 60, 200):
        d = rng.normal(size=m)
        e = np.abs(rng.normal(size=max(m - 1, 0))) + 0.1
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        dd = d.copy()
        z = np.zeros(m * m)
        assert Hp.hostcore().hc_tridiag(m, dd.ctypes.data, np.ascontiguousarray(e).ctypes.data, z.ctypes.data) == 0
        ref = np.linalg.eigvalsh(T)
        assert np.abs(dd - ref).max() < 1e-12 * max(1, np.abs(ref).max())
        Z = z.reshape(m, m, order="F")
        assert np.abs(T @ Z - Z * dd).max() < 1e-11 * max(1, np.abs(ref).max())


def test_dense_symmetric_jacobi_vs_numpy():
    """The thick-restart Lanczos's Rayleigh-Ritz solver (ed_tridiag.h
    symeig_jacobi) on random symmetric, arrowhead and degenerate matrices."""
    rng = np.random.default_rng(7)
    cases = []
    for m in (1, 2, 5, 40, 120):
        a = rng.normal(size=(m, m))
        cases.append((a + a.T) / 2)
    m = 30                                   # arrowhead + tridiagonal, as after a restart
    a = np.diag(np.sort(rng.normal(size=m)))
    k = 12
    a[:k, k] = a[k, :k] = rng.normal(size=k)
    for j in range(k, m - 1):
        a[j, j + 1] = a[j + 1, j] = abs(rng.normal()) + 0.1
    cases.append(a)
    cases.append(np.diag([1.0, 1.0, 2.0, 2.0, 2.0, -3.0]))   # degenerate
    for A in cases:
        m = A.shape[0]
        d = np.zeros(m)
        z = np.zeros(m * m)
        assert Hp.hostcore().hc_symeig(m, np.ascontiguousarray(A).ctypes.data, d.ctypes.data, z.ctypes.data) == 0
        ref = np.linalg.eigvalsh(A)
        scale = max(1.0, np.abs(ref).max())
        assert np.abs(d - ref).max() < 1e-12 * scale
        Z = z.reshape(m, m, order="F")
        assert np.abs(A @ Z - Z * d).max() < 1e-11 * scale
        assert np.abs(Z.T @ Z - np.eye(m)).max() < 1e-12


TORUS = [
    (L.chain(16), 8, (0, 0)), (L.chain(16, phases=[0.3] * 16), 8, (5, 0)), (L.chain(13), 6, (2, 0)),
    (L.square(4, 4, J2=0.5), 8, (0, 0)), (L.square(4, 4, J2=0.5, lam=0.7), 8, (1, 3)),
    (L.square(4, 5, J2=0.5), 9, (2, 1)),
    (L.triangular(4, 4, Ch=0.3), 8, (0, 0)), (L.triangular(4, 4, Ch=0.3), 8, (1, 2)),
    (L.triangular(3, 6, Ch=0.2, J2=0.1), 9, (2, 5)), (L.triangular(6, 3, Ch=0.2), 9, (3, 1)),
]


@pytest.mark.parametrize("model,n,k", TORUS, ids=[f"{m.name}_n{n}_k{k[0]}{k[1]}" for m, n, k in TORUS])
def test_rotation_fast_path_vs_oracle(model, n, k):
    """The rotation fast path (ed_torus.h) against the oracle; its
    representative search against the generic walk, phase code included."""
    o = P.OracleBasis(model, n, *k)
    h = Hp.HostCore(model, n, *k)
    v = np.random.default_rng(4).normal(size=o.D) + 1j * np.random.default_rng(5).normal(size=o.D)
    w = h.hv_torus(v)
    assert w is not None, "lattice should take the rotation fast path"
    wo = o.hv(v)
    assert np.linalg.norm(w - wo) <= 1e-12 * np.linalg.norm(wo)
    rng = np.random.default_rng(6)
    for _ in range(300):
        bits = rng.choice(model.n_sites, size=n, replace=False)
        t = int(sum(1 << int(b) for b in bits))
        assert h.torus_rep(t) == h.generic_rep(t)


def test_rotation_fast_path_declines_other_lattices():
    h = Hp.HostCore(L.tilted_square((3, 1), (-1, 3)), 5, 0, 0)
    assert h.hv_torus(np.zeros(h.D, complex)) is None
    h = Hp.HostCore(L.triangular(4, 3, ChCh=0.1), 6, 0, 0)
    assert h.hv_torus(np.zeros(h.D, complex)) is None


SHAPES = [(L.square(4, 8), 16, (1, 3)), (L.triangular(6, 5), 15, (2, 4)), (L.triangular(3, 9), 13, (1, 7)),
          (L.chain(24), 12, (5, 0)), (L.square(4, 4), 8, (1, 2)), (L.square(4, 5), 10, (3, 4)),
          (L.triangular(6, 3), 9, (5, 2)), (L.chain(16), 8, (3, 0))]


@pytest.mark.parametrize("model,n,k", SHAPES, ids=[m.name for m, n, k in SHAPES])
def test_compiled_shape_search_matches_generic_walk(model, n, k):
    """Representative and first-minimum phase code of the compile-time shapes
    against the generic translation walk, on random states of the full
    benchmark sizes (no basis enumeration needed)."""
    # a tiny-weight build gives the handle (n_ones=1 enumerates N states only)
    h = Hp.HostCore(model, 1, *k)
    rng = np.random.default_rng(model.n_sites)
    for _ in range(2000):
        bits = rng.choice(model.n_sites, size=n, replace=False)
        t = int(sum(1 << int(b) for b in bits))
        assert h.torus_rep_c(t) == h.generic_rep(t)
    # states with nontrivial stabilisers (periodic patterns) exercise the tie-break
    L1, L2 = model.orders()
    if L2 > 1:
        w = L1
        row = sum(1 << q for q in range(w // 2))
        t = sum(row << (w * y) for y in range(L2))
        assert h.torus_rep_c(t) == h.generic_rep(t)
    else:   # rings: period-2 and period-4 patterns
        for t in (sum(1 << q for q in range(0, model.n_sites, 2)), sum(3 << q for q in range(0, model.n_sites, 4))):
            assert h.torus_rep_c(t) == h.generic_rep(t)
    # states invariant under a subgroup (ties between several minimising
    # translations): the union of the images of a random pattern under the
    # powers of one translation T1^p T2^q
    T1, T2 = np.asarray(model.T1), np.asarray(model.T2)
    for _ in range(300):
        p, q = int(rng.integers(0, L1)), int(rng.integers(0, max(L2, 1)))
        g = np.arange(model.n_sites)
        for _ in range(p):
            g = T1[g]
        for _ in range(q):
            g = T2[g]
        sites = set(int(x) for x in rng.choice(model.n_sites, size=int(rng.integers(1, 5)), replace=False))
        orbit, cur = set(sites), set(sites)
        for _ in range(2 * model.n_sites):
            cur = set(int(g[x]) for x in cur)
            orbit |= cur
        t = sum(1 << x for x in orbit)
        assert h.torus_rep_c(t) == h.generic_rep(t), (p, q, sorted(orbit))


ANYMIN = [(L.square(4, 4), 8, (1, 2)), (L.square(4, 4), 8, (2, 2)), (L.triangular(6, 3), 9, (2, 1)),
          (L.triangular(6, 3), 9, (3, 0)), (L.square(4, 5), 10, (2, 1)), (L.triangular(4, 4), 8, (2, 0))]


@pytest.mark.parametrize("model,n,k", ANYMIN, ids=[f"{m.name}_k{k[0]}{k[1]}" for m, n, k in ANYMIN])
def test_any_minimising_translation_gives_the_reference_phase(model, n, k):
    """The kernels take the code of some minimising translation instead of the
    reference's first minimum.  On every state of the sector's weight: the
    same representative, and wherever the representative belongs to the
    sector (the only targets whose phase enters H*v) the same phase as the
    reference's state_ph from the oracle; the two codes differ on some states
    (representatives with nontrivial stabilisers)."""
    o = P.OracleBasis(model, n, *k)
    tab = o.tables(full=True)
    rep_loc, sph = tab["rep_loc"], tab["state_ph"]
    h = Hp.HostCore(model, 1, *k)
    differ = 0
    for bits in itertools.combinations(range(model.n_sites), n):
        t = sum(1 << b for b in bits)
        r0, c0 = h.torus_rep_c(t)
        r1, c1 = h.torus_rep_c_any(t)
        assert r1 == r0
        differ += c1 != c0
        q = Hp.cns_rank(t)
        if rep_loc[q] >= 0:
            assert abs(h.phase(c1) - sph[q]) <= 1e-14, (t, c0, c1)
            assert abs(h.phase(c0) - sph[q]) == 0.0, (t, c0)
    assert differ > 0


BATCHED = [(L.square(4, 4, J2=0.5, lam=0.8), 8, (1, 3)), (L.square(4, 4, J2=0.5), 8, (0, 0)),
           (L.square(4, 5, J2=0.3), 10, (2, 1)), (L.triangular(6, 3, Ch=0.2, J2=0.1), 9, (3, 1)),
           (L.chain(16, phases=[0.4] * 16), 8, (5, 0)), (L.triangular(4, 4, Ch=0.3), 7, (2, 3)),
           (L.triangular(4, 4, Ch=0.3), 8, (0, 0)), (L.triangular(4, 4, J1=0.0, Ch=0.3), 8, (1, 1))]


@pytest.mark.parametrize("model,n,k", BATCHED, ids=[f"{m.name}_n{n}_k{k[0]}{k[1]}" for m, n, k in BATCHED])
def test_batched_compiled_row_vs_oracle(model, n, k):
    o = P.OracleBasis(model, n, *k)
    h = Hp.HostCore(model, n, *k)
    v = np.random.default_rng(8).normal(size=o.D) + 1j * np.random.default_rng(9).normal(size=o.D)
    w = h.hv_torus_c(v)
    assert w is not None
    wo = o.hv(v)
    assert np.linalg.norm(w - wo) <= 1e-12 * np.linalg.norm(wo)
    # the software-pipelined row accumulates in the same order: bit-identical
    assert np.array_equal(h.hv_torus_c(v, pipelined=True), w)
    # lattices with chiral terms take the merged class (one element per pair)
    mode = Hp.hostcore().hc_torus_mode(h.h)
    assert mode == (3 if model.chiral and model.Ch else mode)
    # rows split over 2, 4 and 8 lanes: the shares cover every element once
    for lanes in (2, 4, 8):
        ws = h.hv_torus_split(v, lanes)
        assert ws is not None
        assert np.linalg.norm(ws - wo) <= 1e-12 * np.linalg.norm(wo), lanes
        assert np.linalg.norm(ws - w) <= 1e-14 * np.linalg.norm(w), lanes


def _nonuniform_square():
    m = L.square(4, 4, J2=0.5)
    x, y, J, ph = m.bonds[3]
    m.bonds[3] = (x, y, 0.7, ph)          # one J1 bond differs: its family is not uniform
    return m


UNIFORM = [(L.square(4, 4, J2=0.5), 8, (0, 0), 4), (L.chain(16), 8, (0, 0), 4), (L.square(4, 4, J2=0.5), 8, (1, 2), 1),
           (_nonuniform_square(), 8, (0, 0), 2), (L.triangular(6, 3, J2=0.1, lam=0.7), 9, (0, 0), 4)]


@pytest.mark.parametrize("model,n,k,mode", UNIFORM, ids=[f"{m.name}_n{n}_k{k[0]}{k[1]}_m{c}" for m, n, k, c in UNIFORM])
def test_uniform_family_class(model, n, k, mode):
    """Real bond families whose members share J/2 and Delta J/2 take the
    uniform class at k = 0 (coefficient per family, diagonal from the
    families' antiparallel counts); a family with one different J keeps the
    per-member class.  Every layout matches the oracle."""
    o = P.OracleBasis(model, n, *k)
    h = Hp.HostCore(model, n, *k)
    assert Hp.hostcore().hc_torus_mode(h.h) == mode
    v = np.random.default_rng(21).normal(size=o.D) + 1j * np.random.default_rng(22).normal(size=o.D)
    wo = o.hv(v)
    w = h.hv_torus_c(v)
    assert np.linalg.norm(w - wo) <= 1e-12 * np.linalg.norm(wo)
    assert np.array_equal(h.hv_torus_c(v, pipelined=True), w)
    for lanes in (2, 4, 8):
        ws = h.hv_torus_split(v, lanes)
        assert np.linalg.norm(ws - wo) <= 1e-12 * np.linalg.norm(wo), lanes


def _rows2_cases():
    t20, info = L.regular_relabel(L.tilted_square((4, 2), (-2, 4), J2=0.5, lam=0.9))
    tri, _ = L.regular_relabel(L.triangular(6, 2, Ch=0.25))
    return [(t20, 10, info.momentum(0, 0)), (t20, 10, (7, 1)), (t20, 9, (3, 1)),
            (L.square(8, 2, J2=0.4), 8, (3, 1)), (tri, 6, (0, 0)), (tri, 5, (2, 1)),
            (L.chain(12, phases=[0.3] * 12), 6, (0, 0))]


ROWS2 = _rows2_cases()


@pytest.mark.parametrize("model,n,k", ROWS2, ids=[f"{m.name}_n{n}_k{k[0]}{k[1]}" for m, n, k in ROWS2])
def test_two_row_64bit_row_vs_oracle(model, n, k):
    """The two-row 64-bit row function (exact double orbit keys, split rank
    tables, directory index through a double) against the oracle."""
    o = P.OracleBasis(model, n, *k)
    h = Hp.HostCore(model, n, *k)
    assert h.D == o.D and o.D > 0
    v = np.random.default_rng(18).normal(size=o.D) + 1j * np.random.default_rng(19).normal(size=o.D)
    w = h.hv_rows2(v)
    if model.name.startswith("chain"):
        assert w is None                      # a ring is not a two-row lattice
        return
    assert w is not None
    wo = o.hv(v)
    assert np.linalg.norm(w - wo) <= 1e-12 * np.linalg.norm(wo)
    # the wide gather image and the directory give the same sums
    assert np.array_equal(w, h.hv_rows2(v, wide=False))


def test_pruned_orbit_minimum_matches_full_scan():
    """The two-row orbit minimum from the zero-run candidates gives the
    representative and translation code of the scan over all 2W
    translations: every state of the 20-site cluster, 400k random, sparse,
    dense and row-periodic 40-site states, and the all-zero/all-one rows."""
    H = Hp.hostcore()
    assert H.hc_rep64_check(400_000, 7) == 0


def test_split_rank_tables_at_40_sites():
    """N=40, n=20: the split 32/64-bit rank tables give the combinatorial
    number system rank (closed form) on random and extreme states."""
    import math
    H = Hp.hostcore()
    r = H.hc_ranker_new(40, 20)
    rng = np.random.default_rng(4)
    states = [sum(1 << int(b) for b in rng.choice(40, 20, replace=False)) for _ in range(2000)]
    states += [(1 << 20) - 1, ((1 << 20) - 1) << 20]
    for s in states:
        assert H.hc_rank_split(r, s) == Hp.cns_rank(s)
    assert H.hc_rank_split(r, ((1 << 20) - 1) << 20) == math.comb(40, 20) - 1
    H.hc_ranker_free(r)


def test_kernel_arithmetic_under_address_and_ub_sanitizers(tmp_path):
    """tests/hostcore/sancheck.cpp: the host build of the kernel arithmetic
    compiled with -fsanitize=address,undefined (aborting on the first report)
    and run on small lattices, every code path checked against the others."""
    import subprocess
    src = os.path.join(Hp.HERE, "hostcore", "sancheck.cpp")
    exe = str(tmp_path / "sancheck")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-fno-omit-frame-pointer", "-I", os.path.join(Hp.ROOT, "ed_code_amd", "csrc"), "-I",
                    os.path.join(Hp.HERE, "hostcore"), src, "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
                                UBSAN_OPTIONS="print_stacktrace=1"))
    assert r.returncode == 0, r.stderr[-4000:]
    assert "sancheck ok" in r.stdout
