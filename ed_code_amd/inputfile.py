"""The reference's key=value input format (read_input_file.cpp:59-430).

`write_input` emits a file the reference's own parser (and main, ED.cpp)
reads; `read_input` restates ED_2D_Lattice_Geometry / routine_parameters with
their parse-order semantics:

* J1 is captured when a Pairs1 entry is read and J2 when a Pairs2 entry is
  read (read_input_file.cpp:207,233), so each must precede its pair list;
* Pairs1 entries carry exp(i*phase), Pairs2 entries phase 1;
* every Triangles entry becomes three cyclic chiral entries whose phases are
  looked up in the bonds read so far (:262-273, get_expphase :3-14);
* the bowtie list is the 9-fold expansion of every ordered pair of
  neighbouring triangles over the final bond list (:286-319);
* a line whose first non-blank character is '[' is a comment (:178-179).
"""
from __future__ import annotations

import cmath
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .lattices import Model, bowtie_entries, chiral_entries


@dataclass
class InputParams:
    model: Model
    Sz: float = -1.0
    kx: int = -1
    ky: int = -1
    pSector: bool = False
    num_eigenvalues: int = 2
    find_vectors: bool = False
    unwrap_pvector: bool = True
    vector_file_path: str = ""
    tol: float = 10e-12
    lanczos_iterations: int = 100
    lanczos_ncycles: int = 10
    use_arpack: bool = False


def _fmt(x) -> str:
    return repr(float(x))


def write_input(path: str, model: Model, Sz: float, kx: int, ky: int, num_eigenvalues: int = 2,
                find_vectors: bool = False, unwrap_pvector: bool = True, vector_file_path: str = "",
                tol: float = 1e-11, lanczos_iterations: int = 100, lanczos_ncycles: int = 10) -> None:
    """Writes `model` and one sector request.  Bonds whose J equals the first
    bond's J go to Pairs1 (with phases), the rest (one other J, phase 1) to
    Pairs2."""
    if not model.bonds:
        raise ValueError("model without bonds")
    J1 = model.bonds[0][2]
    p1 = [(x, y, cmath.phase(ph)) for (x, y, J, ph) in model.bonds if J == J1]
    rest = [(x, y, J, ph) for (x, y, J, ph) in model.bonds if J != J1]
    J2s = {J for (_, _, J, _) in rest}
    if len(J2s) > 1 or any(abs(ph - 1) > 1e-15 for (_, _, _, ph) in rest):
        raise ValueError("the input format has one J1 family with phases and one J2 family without")
    lines = [f"tol={tol!r}", f"Lanczos_iterations={lanczos_iterations}", f"Lanczos_ncycles={lanczos_ncycles}",
             f"Jz={_fmt(model.lam)}", f"J1={_fmt(J1)}",
             "Pairs1=[" + ",".join(f"[{x},{y},{_fmt(ph)}]" for (x, y, ph) in p1) + "]"]
    if rest:
        lines.append(f"J2={_fmt(rest[0][2])}")
        lines.append("Pairs2=[" + ",".join(f"[{x},{y}]" for (x, y, _, _) in rest) + "]")
    if model.triangles:
        lines.append(f"Ch={_fmt(model.Ch)}")
        lines.append(f"ChCh={_fmt(model.ChCh)}")
        lines.append("Triangles=[" + ",".join(f"[{i},{j},{k}]" for (i, j, k) in model.triangles) + "]")
    lines.append("T1=[" + ",".join(str(int(t)) for t in model.T1) + "]")
    lines.append("T2=[" + ",".join(str(int(t)) for t in model.T2) + "]")
    lines += ["Specific_pSector=1", f"Sz={_fmt(Sz)}", f"kx={kx}", f"ky={ky}", f"Num_Eigenvalues={num_eigenvalues}",
              f"Find_Vectors={int(find_vectors)}", f"Unwrap_pVector={int(unwrap_pvector)}"]
    if vector_file_path:
        lines.append(f"Vector_File_Path={vector_file_path}")
    with open(path, "w") as f:
        f.write("[ generated input: one momentum sector\n")
        f.write("\n".join(lines) + "\n")


def _int_list(s: str) -> List[int]:
    s = s.strip()
    return [int(t) for t in s.strip("[]").split(",") if t.strip()]


def read_input(path: str, name: str = "input") -> InputParams:
    J1, J2, Jz, Ch, ChCh = 1.0, 0.0, 1.0, 0.0, 0.0
    bonds, tris = [], []
    T1: List[int] = []
    T2: List[int] = []
    p = dict(Sz=-1.0, kx=-1, ky=-1, pSector=False, num_eigenvalues=2, find_vectors=False, unwrap_pvector=True,
             vector_file_path="", tol=10e-12, lanczos_iterations=100, lanczos_ncycles=10, use_arpack=False)
    chiral = []
    for raw in open(path):
        line = "".join(raw.split())          # the parser reads non-blank characters
        if not line or line.startswith("["):
            continue
        key, _, val = line.partition("=")
        if key == "Jz":
            Jz = float(val)
        elif key == "J1":
            J1 = float(val)
        elif key == "J2":
            J2 = float(val)
        elif key == "Ch":
            Ch = float(val)
        elif key == "ChCh":
            ChCh = float(val)
        elif key == "Pairs1":
            for item in val.strip("[]").split("],["):
                x, y, ph = item.split(",")
                bonds.append((int(x), int(y), J1, cmath.exp(1j * float(ph))))
        elif key == "Pairs2":
            for item in val.strip("[]").split("],["):
                x, y = item.split(",")
                bonds.append((int(x), int(y), J2, complex(1.0, 0.0)))
        elif key == "Triangles":
            new = []
            for item in val.strip("[]").split("],["):
                i, j, k = (int(t) for t in item.split(","))
                new.append((i, j, k))
            tris += new
            chiral += chiral_entries(new, bonds)
        elif key == "T1":
            T1 = _int_list(val)
        elif key == "T2":
            T2 = _int_list(val)
        elif key == "Specific_pSector":
            p["pSector"] = bool(int(val))
        elif key == "Sz":
            p["Sz"] = float(val)
        elif key == "kx":
            p["kx"] = int(val)
        elif key == "ky":
            p["ky"] = int(val)
        elif key == "Num_Eigenvalues":
            p["num_eigenvalues"] = int(val)
        elif key == "Find_Vectors":
            p["find_vectors"] = bool(int(val))
        elif key == "Unwrap_pVector":
            p["unwrap_pvector"] = bool(int(val))
        elif key == "Vector_File_Path":
            p["vector_file_path"] = val
        elif key == "tol":
            p["tol"] = float(val)
        elif key == "Lanczos_iterations":
            p["lanczos_iterations"] = int(val)
        elif key == "Lanczos_ncycles":
            p["lanczos_ncycles"] = int(val)
        elif key == "Use_ARPACK":
            p["use_arpack"] = bool(int(val))
    if len(T1) != len(T2):
        raise ValueError("T1 and T2 have different lengths (read_input_file.cpp:423-427)")
    m = Model(name, len(T1), np.array(T1, np.int32), np.array(T2, np.int32), lam=Jz, bonds=bonds, Ch=Ch,
              chiral=chiral, ChCh=ChCh, triangles=tris)
    m.bowties = bowtie_entries(tris, bonds) if tris else []
    return InputParams(model=m, **p)


def parse_energies(log_path: str) -> List[float]:
    """Energies of the table the sector drivers write after
    'Energies\\t\\tEnergies/site...' (Lanczos_sector.cpp:128-132)."""
    out, on = [], False
    for line in open(log_path):
        if line.startswith("Energies\t"):
            on = True
            continue
        if on:
            t = line.split()
            if len(t) >= 2:
                try:
                    out.append(float(t[0]))
                except ValueError:
                    break
    return out
