"""CPU: the CRD writer and the PDB -> CRD converter (host code of libpac).

The writer must reproduce the layout of the reference's own data files byte
for byte (data/1grm.crd, data/1k4c.crd: count line, blank line, rows of
four 4-decimal numbers separated by " \\t", "#TER").  The PDB reader is
checked on hand-made records covering the fixed-column format: alternate
locations, several models, hydrogens, waters, HETATM records, elements
from columns 77-78 or from the atom name, radius overrides, and errors.
The reference ships no converter, so these checks pin the documented
behaviour, not a reference output.
"""
import os
import subprocess

import numpy as np
import pytest

import pac

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "parallelac-xmp_amd", "bin")


def _rec(kind, serial, name, alt, res, chain, seq, x, y, z, elem):
    """one ATOM / HETATM record in the PDB's fixed columns; name is the
    4-character atom name field (columns 13-16)"""
    assert len(name) == 4
    line = "%-6s%5d %4s%1s%3s %1s%4d    %8.3f%8.3f%8.3f%6.2f%6.2f          %2s" % (
        kind, serial, name, alt, res, chain, seq, x, y, z, 1.0, 20.0, elem)
    assert line[30:38].strip() == "%.3f" % x and line[76:78] == "%2s" % elem
    return line


def _pdb(tmp_path, lines, name="t.pdb"):
    p = tmp_path / name
    p.write_text("\n".join(lines) + "\n")
    return str(p)


@pytest.mark.parametrize("fname", ["1grm.crd", "1k4c.crd"])
def test_write_crd_reproduces_reference_files(golden_dir, tmp_path, fname):
    src = os.path.join(golden_dir, fname)
    b = pac.read_crd(src)
    out = tmp_path / "w.crd"
    pac.write_crd(str(out), b)
    orig = open(src, "rb").read()
    # the two reference files differ only in a newline after "#TER"
    assert out.read_bytes().rstrip(b"\n") == orig.rstrip(b"\n")
    assert np.array_equal(pac.read_crd(str(out)), b)


def test_read_pdb_records(tmp_path):
    L = [
        "HEADER    TEST",
        "MODEL        1",
        _rec("ATOM", 1, " N  ", " ", "MET", "A", 1, 1.0, 2.0, 3.0, "N"),
        _rec("ATOM", 2, " CA ", " ", "MET", "A", 1, 2.5, 2.0, 3.0, "C"),
        _rec("ATOM", 3, " SD ", "A", "MET", "A", 1, 4.0, 1.0, -1.25, "S"),
        _rec("ATOM", 4, " SD ", "B", "MET", "A", 1, 9.0, 9.0, 9.0, "S"),
        _rec("ATOM", 5, " O  ", " ", "MET", "A", 1, -7.125, 0.5, 100.0, "O"),
        _rec("ATOM", 6, " H  ", " ", "MET", "A", 1, 0.0, 0.0, 0.0, "H"),
        _rec("ATOM", 7, "1HB ", " ", "MET", "A", 1, 0.0, 0.0, 1.0, ""),  # H from name
        _rec("ATOM", 8, " CB ", " ", "MET", "A", 1, 3.0, 3.0, 3.0, ""),  # C from name
        _rec("HETATM", 9, "ZN  ", " ", " ZN", "A", 101, 5.0, 5.0, 5.0, "ZN"),
        _rec("HETATM", 10, " O  ", " ", "HOH", "A", 201, 6.0, 6.0, 6.0, "O"),
        "TER",
        "ENDMDL",
        "MODEL        2",
        _rec("ATOM", 1, " N  ", " ", "MET", "A", 1, 11.0, 12.0, 13.0, "N"),
        "ENDMDL",
        "END",
    ]
    path = _pdb(tmp_path, L)
    b = pac.read_pdb(path)
    want = np.array([[1.0, 2.0, 3.0, 1.400], [2.5, 2.0, 3.0, 1.548],
                     [4.0, 1.0, -1.25, 1.808], [-7.125, 0.5, 100.0, 1.348],
                     [3.0, 3.0, 3.0, 1.548]], np.float32)
    assert np.array_equal(b, want)
    # hydrogens (one from columns 77-78, one from the atom name)
    bh = pac.read_pdb(path, hydrogens=True)
    assert bh.shape == (7, 4)
    assert np.array_equal(bh[4], np.float32([0.0, 0.0, 0.0, 1.2]))
    assert np.array_equal(bh[5], np.float32([0.0, 0.0, 1.0, 1.2]))
    assert np.array_equal(bh[6], want[4])
    # HETATM: the zinc needs a radius; waters stay out unless asked for
    with pytest.raises(pac.PacError) as e:
        pac.read_pdb(path, hetatm=True)
    assert e.value.code == pac.PAC_ERR_INVALID and "ZN" in str(e.value)
    bz = pac.read_pdb(path, hetatm=True, radii={"ZN": 1.39})
    assert bz.shape == (6, 4) and np.array_equal(bz[5], np.float32([5, 5, 5, 1.39]))
    bw = pac.read_pdb(path, hetatm=True, water=True, radii="ZN=1.39,O=1.5")
    assert bw.shape == (7, 4)
    assert bw[3, 3] == np.float32(1.5) and np.array_equal(
        bw[6], np.float32([6, 6, 6, 1.5]))
    bs = pac.read_pdb(path, hetatm=True, radii="*=2.0")
    assert bs[5, 3] == np.float32(2.0) and bs[0, 3] == np.float32(1.4)


def test_pdb_to_crd_round_trip_and_cli(tmp_path):
    rng = np.random.default_rng(4)
    L = []
    els = ["N", "C", "C", "O", "S"]
    names = [" N  ", " CA ", " C  ", " O  ", " SG "]
    want = []
    for i in range(200):
        x, y, z = np.round(rng.uniform(-99, 99, 3), 3)
        k = i % 5
        L.append(_rec("ATOM", i + 1, names[k], " ", "CYS", "B", i // 5 + 1,
                      x, y, z, els[k]))
        want.append([x, y, z, {"N": 1.4, "C": 1.548, "O": 1.348, "S": 1.808}[els[k]]])
    path = _pdb(tmp_path, L + ["END"])
    b = pac.read_pdb(path)
    assert np.array_equal(b, np.array(want, np.float32))
    crd = tmp_path / "x.crd"
    pac.write_crd(str(crd), b)
    assert np.array_equal(pac.read_crd(str(crd)), b)
    crd2 = tmp_path / "y.crd"
    p = subprocess.run([os.path.join(BIN, "pdb2crd"), path, str(crd2)],
                       capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stderr
    assert crd2.read_bytes() == crd.read_bytes()
    assert "Wrote 200 balls" in p.stdout


def test_pdb_errors(tmp_path):
    with pytest.raises(pac.PacError) as e:
        pac.read_pdb(str(tmp_path / "missing.pdb"))
    assert e.value.code == pac.PAC_ERR_IO
    bad = _rec("ATOM", 1, " CA ", " ", "GLY", "A", 1, 1.0, 2.0, 3.0, "C")
    bad = bad[:30] + "  1.0x0 " + bad[38:]
    with pytest.raises(pac.PacError) as e:
        pac.read_pdb(_pdb(tmp_path, [bad], "bad.pdb"))
    assert e.value.code == pac.PAC_ERR_IO and "line 1" in str(e.value)
    ok = _pdb(tmp_path, [_rec("ATOM", 1, " CA ", " ", "GLY", "A", 1, 1, 2, 3, "C")],
              "ok.pdb")
    for spec in ("C", "C=", "=1.5", "C=-1", "C=1.5x"):
        with pytest.raises(pac.PacError) as e:
            pac.read_pdb(ok, radii=spec)
        assert e.value.code == pac.PAC_ERR_INVALID, spec
    assert pac.read_pdb(_pdb(tmp_path, ["REMARK nothing"], "e.pdb")).shape == (0, 4)
