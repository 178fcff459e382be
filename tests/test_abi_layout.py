"""CPU: the ctypes mirrors of the C structs (pac.Options, pac.Stats, and the
minimal stub INTEGRATION.md gives to maintainers) have the size and field
offsets that include/pac.h gives them, as compiled by gcc.
pac_options_init clears sizeof(pac_options) bytes, so a short mirror
would be written past its end."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

import pac

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def c_layout(struct, fields):
    src = ['#include <stddef.h>', '#include <stdio.h>', '#include "pac.h"',
           'int main(void) {', '  printf("size %%zu\\n", sizeof(%s));' % struct]
    for f in fields:
        src.append('  printf("%s %%zu\\n", offsetof(%s, %s));' % (f, struct, f))
    src += ['  return 0;', '}']
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        with open(c, "w") as fh:
            fh.write("\n".join(src) + "\n")
        subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), c, "-o", exe],
                       check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return {k: int(v) for k, v in (line.split() for line in out.splitlines())}


def py_layout(cls):
    d = {"size": ctypes.sizeof(cls)}
    for f, _ in cls._fields_:
        d[f] = getattr(cls, f).offset
    return d


@pytest.mark.parametrize("cls,struct", [(pac.Options, "pac_options"), (pac.Stats, "pac_stats")])
def test_ctypes_mirror_matches_c(cls, struct):
    want = c_layout(struct, [f for f, _ in cls._fields_])
    assert py_layout(cls) == want


def test_integration_stub_matches_c():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"(class Opt\(ctypes\.Structure\):\n(?:    .*\n|\s+\(.*\n)+)", text)
    assert m, "stub struct not found"
    ns = {"ctypes": ctypes}
    exec(m.group(1), ns)
    Opt = ns["Opt"]
    want = c_layout("pac_options", [f for f, _ in Opt._fields_])
    assert py_layout(Opt) == want
    assert ctypes.sizeof(Opt) == ctypes.sizeof(pac.Options)
