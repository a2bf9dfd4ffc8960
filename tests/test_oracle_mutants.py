"""Each pin of the reference-only rules catches a plausible mistake.

For every rule of kde (Library/cudaLibrary.cu:356-422) and of the
Lyapunov-spectrum reference directions (:1025-1040), a copy of
oracle/orc.c with one planted mistake is compiled (same flags as
oracle/Makefile, into pytest's temporary directory) and the pin checks of
tests/test_oracle_pins.py are run against it: at least one must fail.  The
unmodified oracle passes them (tests/test_oracle_pins.py itself).
"""
import concurrent.futures as cf
import os
import subprocess

import pytest

import orclib
import test_oracle_pins as pins

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORC_C = os.path.join(ROOT, "oracle", "orc.c")

KDE_CHECKS = [pins.check_kde_sums_are_the_kde, pins.check_kde_count_is_the_run_rule,
              pins.check_kde_hand_cases]
LS_CHECKS = [pins.check_ls_reference_directions]

# (name, text in orc.c, replacement, checks that must catch it)
MUTANTS = [
    ("kde loop samples I2 too (w < k1)",
     "for (int w = 0; w < k1 - 1; ++w) {", "for (int w = 0; w < k1; ++w) {", KDE_CHECKS),
    ("kde compares from w = 1",
     "        if (w < 2)\n            continue;\n        if (strange)",
     "        if (w < 1)\n            continue;\n        if (strange)", KDE_CHECKS),
    ("kde compares from w = 3",
     "        if (w < 2)\n            continue;\n        if (strange)",
     "        if (w < 3)\n            continue;\n        if (strange)", KDE_CHECKS),
    ("kde without the strange-peak state",
     "            strange = 1;\n", "            strange = 0;\n", KDE_CHECKS),
    ("kde counts a plateau that rises",
     "            else if (p < cur) {\n                strange = 0;\n",
     "            else if (p < cur) {\n                strange = 0;\n                ++result;\n",
     KDE_CHECKS),
    ("kde does not count a plateau that falls",
     "                strange = 0;\n                ++result;\n", "                strange = 0;\n",
     KDE_CHECKS),
    ("kde without the tail rule",
     "    if (p < cur)\n        ++result;\n    return result;",
     "    return result;", KDE_CHECKS),
    ("kde tail rule not strict",
     "    if (p < cur)\n        ++result;\n    return result;",
     "    if (p <= cur)\n        ++result;\n    return result;", KDE_CHECKS),
    ("kde step (I2-I1)/k1 in the counting loop",
     "    double k2 = (I2 - I1) / (k1 - 1);\n    double delt = 0, pp",
     "    double k2 = (I2 - I1) / k1;\n    double delt = 0, pp", KDE_CHECKS),
    ("kde sums step (I2-I1)/k1",
     "    double k2 = (I2 - I1) / (k1 - 1);\n    int n = 0;",
     "    double k2 = (I2 - I1) / k1;\n    int n = 0;", KDE_CHECKS),
    ("kde sums without the /2 of the exponent",
     "            cur += expf((float)(-((t * t) / 2)));\n        }\n        if (n < cap)",
     "            cur += expf((float)(-(t * t)));\n        }\n        if (n < cap)", KDE_CHECKS),
    ("LS directions (i+j+1) instead of (i+1)(j+1)",
     "0.2171828 * (i + 1) * (j + 1) + idx", "0.2171828 * (i + j + 1) + idx", LS_CHECKS),
    ("LS directions i*idx instead of i*j*idx",
     "(0.2171828 + i * j * idx)", "(0.2171828 + i * idx)", LS_CHECKS),
    ("LS directions in double sin",
     "z[j * n + i] = sinf((float)arg) * 0.5;", "z[j * n + i] = sin(arg) * 0.5;", LS_CHECKS),
]


def _build(src_text, out_dir, k):
    c = os.path.join(out_dir, f"orc_m{k}.c")
    so = os.path.join(out_dir, f"liborc_m{k}.so")
    with open(c, "w") as f:
        f.write(src_text)
    subprocess.run(["gcc", "-O3", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-fopenmp",
                    "-std=gnu11", "-w", "-I", os.path.join(ROOT, "oracle"), "-shared", "-o", so,
                    c, "-lm"], check=True)
    return so


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    src = open(ORC_C).read()
    out = str(tmp_path_factory.mktemp("mutants"))
    texts = []
    for name, old, new, _ in MUTANTS:
        assert src.count(old) == 1, f"{name}: pattern found {src.count(old)} times"
        texts.append(src.replace(old, new))
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 2)) as ex:
        return list(ex.map(lambda kt: _build(kt[1], out, kt[0]), enumerate(texts)))


@pytest.mark.parametrize("k", range(len(MUTANTS)), ids=[m[0] for m in MUTANTS])
def test_pins_catch_planted_mistake(mutant_libs, k):
    name, _, _, checks = MUTANTS[k]
    caught = []
    with orclib.use_library(mutant_libs[k]):
        for check in checks:
            try:
                check(orclib)
            except AssertionError:
                caught.append(check.__name__)
    assert caught, f"no pin check fails on the mutant: {name}"
