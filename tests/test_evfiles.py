"""Eigenvector files (ed_code_amd.evfiles): the reference's text layout as its
readers parse it (analysis/src/read_inputfile.cpp:297-360) and the binary
twin, round trips on synthetic vectors."""
import numpy as np

from ed_code_amd import evfiles as E


def _vec(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=n) + 1j * rng.normal(size=n)


def test_pev_text_and_binary_round_trip(tmp_path):
    reps = np.array([15, 23, 27, 29, 39, 43], np.int64)
    norm = np.array([2.0, 1.4142135623730951, 2.0, 2.0, 1.0, 2.0])
    co = _vec(6, 1)
    path = str(tmp_path / "v_pEV0.txt")
    E.write_pev(path, 0.0, (1, 2), 0.8, -7.25, (0.5, 1.5707963267948966), reps, norm, co)
    text = open(path).read().split("\n")
    assert text[1] == " Eigenvector for state " and text[2] == "Sz\t(kx,ky)\tJz/lambda"
    assert text[3] == "0\t(1,2)\t0.8" and text[6] == "-------------------------------------------"
    assert text[7] == "State Energy (-7.25,0)"
    n_p, p, r, nm, c = E.read_pev(path)
    assert n_p == 6 and p == (0.5, 1.5708)      # the momenta line keeps 6 significant digits
    assert np.array_equal(r, reps) and np.abs(c - co).max() < 1e-15
    # the stream switches to fixed notation after the first coefficient
    rows = [l for l in text if l.startswith(("15\t", "23\t"))]
    assert rows[0].split("\t\t\t")[1] == "2" and rows[1].split("\t\t\t")[1] == "1.414213562373095"
    b = E.read_bin(str(tmp_path / "v_pEV0.bin"))
    assert b["kind"] == 1 and np.array_equal(b["states"], reps) and np.array_equal(b["norm"], norm)
    assert np.array_equal(b["coeffs"], co) and b["p"] == (0.5, 1.5707963267948966, 0.0)


def test_ev_text_and_binary_round_trip_three_generators(tmp_path):
    states = np.array([3, 5, 6, 9, 10, 12], np.int64)
    co = _vec(6, 2)
    path = str(tmp_path / "v_EV1.txt")
    E.write_ev(path, 1.0, (0, 1, 2), 1.0, -3.0, states, co, p=(0.0, 1.0, 2.0))
    text = open(path).read().split("\n")
    assert text[2] == "Sz\t(kx,ky,kz)\tJz/lambda" and text[3] == "1\t(0,1,2)\t1"
    st, c = E.read_ev(path)
    assert np.array_equal(st, states) and np.abs(c - co).max() < 1e-15
    assert all(len(l.split("\t")[1].strip("()").split(",")[0].split(".")[1]) == 15
               for l in text[text.index("EIGENVECTOR") + 1:] if l)
    b = E.read_bin(str(tmp_path / "v_EV1.bin"))
    assert b["kind"] == 0 and b["norm"] is None and np.array_equal(b["coeffs"], co)
    assert b["p"] == (0.0, 1.0, 2.0)


def test_binary_rejects_foreign_and_truncated_files(tmp_path):
    import pytest
    bad = tmp_path / "x.bin"
    bad.write_bytes(b"not a vector file")
    with pytest.raises(ValueError):
        E.read_bin(str(bad))
    E.write_bin(str(bad), 0, np.arange(4), _vec(4, 3), None, (0.0, 0.0))
    data = bad.read_bytes()
    bad.write_bytes(data[:-16])
    with pytest.raises(ValueError):
        E.read_bin(str(bad))
