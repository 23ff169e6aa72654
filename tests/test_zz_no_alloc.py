"""ed_hv_noalloc (the allocation-free H*v) through the C ABI against ed_hv:
bit-identical results on two calls in a row, the second on a smaller sector
(the shared scratch keeps the first call's larger contents), both
precisions, every kernel family (compiled rotation kernels with the gather
image, the run-time-shape and generic kernels gathering u by row, the
two-row 64-bit kernel with its wide image); sectors beyond the 64 MiB
scratch are rejected before anything runs."""
import ctypes as C

import numpy as np
import pytest

from ed_code_amd import lattices as L

pytestmark = pytest.mark.gpu


def _ed():
    import ed_code_amd as ed
    ed.init(0)
    return ed


def _noalloc(op, v, w):
    import torch
    from ed_code_amd._lib import ED_C64, ED_C128, check, lib
    dt = ED_C128 if v.dtype == torch.complex128 else ED_C64
    st = torch.cuda.current_stream().cuda_stream
    check(lib().ed_hv_noalloc(op.h, dt, C.c_void_p(v.data_ptr()), C.c_void_p(w.data_ptr()), C.c_void_p(st)))
    return w


CASES = [  # (name, model, n_ones, k, tune path); the larger sector of each pair first
    ("tri3x9_ch_k14", L.triangular(3, 9, Ch=0.15), 13, (1, 4), 0),
    ("chain20_k3", L.chain(20, phases=[0.1] * 20), 10, (3, 0), 0),
    ("sq4x5_k00_runtime_shape", L.square(4, 5, J2=0.5), 10, (0, 0), 3),
    ("tri4x4_ch_k13_generic", L.triangular(4, 4, Ch=0.2), 8, (1, 3), 1),
    ("tilted20_two_row", L.regular_relabel(L.tilted_square((4, 2), (-2, 4), J2=0.5))[0], 10, (0, 0), 6),
    ("tri3x6_chch_k02", L.triangular(3, 6, Ch=0.2, ChCh=0.1), 9, (0, 2), 0),
]


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_noalloc_matches_ed_hv(gpu_ready, dtype):
    import torch
    ed = _ed()
    cdt = torch.complex128 if dtype == "c128" else torch.complex64
    ops = []
    for name, model, n, k, path in CASES:
        basis = ed.Basis(ed.Lattice.of(model), n, *k)
        op = ed.Operator(basis, model)
        if path:
            op.tune(path)
        ops.append((name, basis, op))
    # sizes descend pairwise, so each second call runs on a smaller sector
    for q in range(0, len(ops), 2):
        big, small = ops[q], ops[q + 1]
        assert big[1].D > small[1].D
        for name, basis, op in (big, small):
            print(f"[case] noalloc {dtype} {name} D={basis.D}", flush=True)
            g = torch.Generator(device="cuda").manual_seed(basis.D)
            v = torch.randn(basis.D, dtype=cdt, device="cuda", generator=g)
            w_ref = op.hv(v)
            w = _noalloc(op, v, torch.full_like(v, float("nan")))
            torch.cuda.synchronize()
            assert torch.equal(w, w_ref), name
    # again, in the opposite order, the same results
    for name, basis, op in reversed(ops):
        g = torch.Generator(device="cuda").manual_seed(basis.D)
        v = torch.randn(basis.D, dtype=cdt, device="cuda", generator=g)
        assert torch.equal(_noalloc(op, v, torch.empty_like(v)), op.hv(v)), name


def test_noalloc_rejects_sectors_beyond_the_scratch(gpu_ready):
    import torch
    from ed_code_amd._lib import ED_ENOMEM, ED_C128, lib
    ed = _ed()
    model = L.triangular(6, 5, Ch=0.1)          # D = 5.2e6: gather image of ~13e6 elements
    basis = ed.Basis(ed.Lattice.of(model), 15, 0, 0)
    op = ed.Operator(basis, model)
    v = torch.zeros(basis.D, dtype=torch.complex128, device="cuda")
    w = torch.full_like(v, 7.0)
    rc = lib().ed_hv_noalloc(op.h, ED_C128, C.c_void_p(v.data_ptr()), C.c_void_p(w.data_ptr()), None)
    assert rc == ED_ENOMEM
    assert b"scratch" in lib().ed_last_error()
    torch.cuda.synchronize()
    assert bool((w == 7.0).all())               # nothing ran
