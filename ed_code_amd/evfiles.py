"""Eigenvector files: the reference's text formats and a binary twin.

The reference writes one text file per eigenvector (Lanczos_sector.cpp:136-197,
three generators :306-366): `<prefix>_pEV<i>.txt` holds the momentum-basis
vector (representative, orbit norm, coefficient per row) and
`<prefix>_EV<i>.txt` the vector unwrapped to the full Sz basis (state,
coefficient).  Its analysis programs read them back line by line
(analysis/src/read_inputfile.cpp:297-345, read_pEV / readEV).  The text is
formatted as a C++ stream formats it: numbers in the default format (6
significant digits, like printf's %g) until the first coefficient, which
switches the stream to fixed notation with 15 decimals for the rest of the
file (the norm column of the first row is therefore short, all later ones
long).

At 32 sites an unwrapped vector has up to 6e8 entries, tens of GB of text, so
each file also gets a binary twin `<name>.bin` (little-endian):

    8 bytes   magic b"EDAMDV1\\0"
    int64     kind (0 = unwrapped, 1 = momentum basis)
    int64     n (entries)
    float64   px, py, pz
    int64[n]  states (representatives for kind 1)
    float64[n] norms (kind 1 only)
    complex128[n] coefficients

`unwrap_to_files` unwraps on the GPU (Basis.unwrap) and writes both forms.
"""
from __future__ import annotations

from typing import Optional, Sequence, Tuple

import numpy as np

MAGIC = b"EDAMDV1\0"


def _g(x: float) -> str:
    """A double in the C++ stream's default format (%g, 6 significant digits)."""
    return "%g" % x


def _cg(z: complex) -> str:
    return "(" + _g(z.real) + "," + _g(z.imag) + ")"


def _cf(z: complex) -> str:
    return "(%.15f,%.15f)" % (z.real, z.imag)


def _header(f, Sz, k, lam, energy):
    f.write("\n Eigenvector for state \n")
    if len(k) == 3:
        f.write("Sz\t(kx,ky,kz)\tJz/lambda\n")
        f.write(f"{_g(Sz)}\t({k[0]},{k[1]},{k[2]})\t{_g(lam)}\n")
    else:
        f.write("Sz\t(kx,ky)\tJz/lambda\n")
        f.write(f"{_g(Sz)}\t({k[0]},{k[1]})\t{_g(lam)}\n")
    f.write("\n\n-------------------------------------------")
    f.write("\nState Energy " + _cg(complex(energy)) + "\n")


def write_pev(path: str, Sz: float, k: Sequence[int], lam: float, energy: float, p: Sequence[float],
              reps: np.ndarray, norm: np.ndarray, coeffs: np.ndarray, binary: bool = True) -> None:
    """Momentum-basis eigenvector file (Lanczos_sector.cpp:182-190), p = (px, py[, pz])."""
    D = len(reps)
    with open(path, "w") as f:
        _header(f, Sz, k, lam, energy)
        f.write("\nMomentum EIGENVECTOR\n")
        if len(p) == 3:
            f.write("\nnp \t\t px \t\tpy \t\tpz\n")
            f.write(f"{D}\t\t{_g(p[0])}\t\t{_g(p[1])}\t\t{_g(p[2])}\n")
        else:
            f.write("\nnp \t\t px \t\tpy\n")
            f.write(f"{D}\t\t{_g(p[0])}\t\t{_g(p[1])}\n")
        f.write("\n\npblock_states \t\t Norm \t\t\t v_coeff\n\n")
        for jb in range(D):
            nrm = _g(norm[jb]) if jb == 0 else "%.15f" % norm[jb]
            f.write(f"{int(reps[jb])}\t\t\t{nrm}\t\t\t{_cf(complex(coeffs[jb]))}\n")
    if binary:
        write_bin(path[:-4] + ".bin" if path.endswith(".txt") else path + ".bin", 1, reps, coeffs, norm, p)


def write_ev(path: str, Sz: float, k: Sequence[int], lam: float, energy: float, states: np.ndarray,
             coeffs: np.ndarray, p: Sequence[float] = (0.0, 0.0), binary: bool = True) -> None:
    """Unwrapped eigenvector file (Lanczos_sector.cpp:173-180)."""
    with open(path, "w") as f:
        _header(f, Sz, k, lam, energy)
        f.write("\nEIGENVECTOR\n")
        for s, c in zip(states, coeffs):
            f.write(f"{int(s)}\t{_cf(complex(c))}\n")
    if binary:
        write_bin(path[:-4] + ".bin" if path.endswith(".txt") else path + ".bin", 0, states, coeffs, None, p)


def write_bin(path: str, kind: int, states: np.ndarray, coeffs: np.ndarray, norm: Optional[np.ndarray],
              p: Sequence[float]) -> None:
    n = len(states)
    pz = p[2] if len(p) == 3 else 0.0
    with open(path, "wb") as f:
        f.write(MAGIC)
        np.array([kind, n], "<i8").tofile(f)
        np.array([p[0], p[1], pz], "<f8").tofile(f)
        np.ascontiguousarray(states, "<i8").tofile(f)
        if kind == 1:
            np.ascontiguousarray(norm, "<f8").tofile(f)
        np.ascontiguousarray(coeffs, "<c16").tofile(f)


def read_bin(path: str) -> dict:
    """Inverse of write_bin: dict(kind, p, states, norm (kind 1), coeffs)."""
    with open(path, "rb") as f:
        if f.read(8) != MAGIC:
            raise ValueError(f"{path}: not an eigenvector binary file")
        kind, n = (int(x) for x in np.fromfile(f, "<i8", 2))
        p = np.fromfile(f, "<f8", 3)
        states = np.fromfile(f, "<i8", n)
        norm = np.fromfile(f, "<f8", n) if kind == 1 else None
        coeffs = np.fromfile(f, "<c16", n)
    if len(coeffs) != n:
        raise ValueError(f"{path}: truncated")
    return dict(kind=kind, p=tuple(p), states=states, norm=norm, coeffs=coeffs)


def _parse_complex(tok: str) -> complex:
    re, im = tok.strip("()").split(",")
    return complex(float(re), float(im))


def read_pev(path: str) -> Tuple[int, Tuple[float, ...], np.ndarray, np.ndarray, np.ndarray]:
    """read_pEV (analysis/src/read_inputfile.cpp:297-336): the line after the
    'np px py' header gives n_p and the momenta, the rows after the
    'pblock_states Norm v_coeff' header give the vector."""
    lines = open(path).read().split("\n")
    n_p, p = 0, ()
    for i, line in enumerate(lines):
        if line.startswith("np \t\t px"):
            vals = lines[i + 1].split()
            n_p, p = int(vals[0]), tuple(float(x) for x in vals[1:])
            break
    start = next(i for i, line in enumerate(lines) if line.startswith("pblock_states \t\t Norm"))
    rows = [l.split() for l in lines[start + 1:] if l.strip()][:n_p]
    reps = np.array([int(r[0]) for r in rows], np.int64)
    norm = np.array([float(r[1]) for r in rows])
    co = np.array([_parse_complex(r[2]) for r in rows], np.complex128)
    return n_p, p, reps, norm, co


def read_ev(path: str) -> Tuple[np.ndarray, np.ndarray]:
    """readEV (analysis/src/read_inputfile.cpp:338-360): (state, coefficient)
    rows after the 'EIGENVECTOR' line."""
    lines = open(path).read().split("\n")
    start = next(i for i, line in enumerate(lines) if line == "EIGENVECTOR")
    rows = [l.split() for l in lines[start + 1:] if l.strip()]
    return (np.array([int(r[0]) for r in rows], np.int64),
            np.array([_parse_complex(r[1]) for r in rows], np.complex128))


def unwrap_to_files(basis, vec, prefix: str, index: int, Sz: float, k: Sequence[int], lam: float, energy: float,
                    momentum: bool = True, unwrapped: bool = True, binary: bool = True) -> None:
    """Writes <prefix>_pEV<index>.txt and/or <prefix>_EV<index>.txt (and their
    binary twins) for the momentum-basis vector vec (complex128 device tensor
    of basis); the unwrapping runs on the GPU."""
    import torch
    p = basis.momenta()
    if momentum:
        reps, norm, _ = basis.export()
        write_pev(f"{prefix}_pEV{index}.txt", Sz, k, lam, energy, p, reps, norm,
                  vec.detach().cpu().numpy() if torch.is_tensor(vec) else np.asarray(vec), binary)
    if unwrapped:
        st, co = basis.unwrap(vec)
        write_ev(f"{prefix}_EV{index}.txt", Sz, k, lam, energy, st, co, p, binary)
