"""GPU: the `parallelac` command itself against the reference's own fixture.

output_files/1grm_pAC_LD_alpha1.txt was written by the reference at
solvent 0, alpha 1 with the tet emptiness test not in effect; the command
`parallelac 1grm.crd out 0 1 --skip-tet-ac2` must reproduce it byte for
byte (writer AlphaComplexParallel.cu:1834-1856), and its stdout summary
must have the reference's lines and field widths (:1809-1830)."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "parallelac-xmp_amd", "bin")
GOLD = os.path.join(ROOT, "tests", "golden")
N, E, T, K = 260, 842, 724, 183


def _say(msg):
    print(msg, flush=True)


def summary_pattern(n, e, t, k):
    """The reference's summary printf formats (AlphaComplexParallel.cu:
    1809-1830) with the counts filled in and the times as %W.3f fields."""
    def f(w):  # a %W.3f field: at least W characters, three decimals
        return r"(?=[ \d.\-]{%d,}$) *-?\d+\.\d{3}$" % w
    lines = [
        "Alpha complex computed.", "",
        "Number of vertices: %15d" % n,
        "Number of edges: %18ld" % e,
        "Number of triangles: %14ld" % t,
        "Number of tetrahedra: %13ld" % k,
        "Total simplices: %18ld" % (n + e + t + k), "",
        "Time taken (ms):",
    ]
    pat = [re.escape(x) for x in lines]
    for label, w in (("Data transfer: ", 20), ("Grid computation: ", 17),
                     ("Potential edges: ", 18), ("Potential triangles: ", 14),
                     ("Potential tetrahedra: ", 13), ("AC2 test for tetrahedra: ", 10),
                     ("AC2 test for triangles: ", 11), ("AC2 test for edges: ", 15),
                     ("Total time taken (ms): ", 12)):
        pat.append(re.escape(label) + f(w))
    return re.compile("\n".join(pat), re.M)


@pytest.mark.parametrize("flags", [["--staged"], ["--plain"], ["--staged", "--reference-raw"]])
def test_cli_reproduces_ld_fixture(tmp_path, flags):
    out = tmp_path / "1grm_out.txt"
    cmd = [os.path.join(BIN, "parallelac"), os.path.join(GOLD, "1grm.crd"), str(out), "0", "1",
           "--skip-tet-ac2"] + flags
    _say("case: " + " ".join(cmd[2:]))
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    ref = open(os.path.join(GOLD, "1grm_pAC_LD_alpha1.txt"), "rb").read()
    assert out.read_bytes() == ref
    with open(os.path.join(GOLD, "1grm.crd")) as fh:
        assert int(fh.read().split()[0]) == N
    assert summary_pattern(N, E, T, K).search(p.stdout), p.stdout
    assert "Writing computed alpha complex to the file %s ..." % out in p.stdout


def test_cli_summary_counts_full_complex(tmp_path):
    """Without --skip-tet-ac2 the summary reports the exact complex of 1grm
    at solvent 0, alpha 1 (812 / 650 / 136, the lower-hull count)."""
    out = tmp_path / "full.txt"
    cmd = [os.path.join(BIN, "parallelac"), os.path.join(GOLD, "1grm.crd"), str(out), "0", "1"]
    _say("case: parallelac 1grm 0 1")
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert summary_pattern(N, 812, 650, 136).search(p.stdout), p.stdout
    head = out.read_text().splitlines()[0]
    assert head == "%d %ld %ld %ld" % (N, 812, 650, 136)
