"""C-ABI and host-side behaviour of libbifhip.so (CPU only, no GPU calls).

The library must load and export every function include/bifhip.h declares;
the host-side pieces (CSV writers, getValueByIdx, argument validation,
no-device errors) are exercised without touching a GPU.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import bifhip as B


def test_library_exports_every_declared_symbol():
    lib = B.load()
    names = B.header_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", B.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert set(names) <= exported


def test_value_by_idx_matches_oracle(orc):
    for n, s, f in [(512, 0.05, 0.45), (1000, 2.0, 7.0), (4096, 2.0, 10.0), (7, -1.0, 1.0)]:
        for idx in range(0, n * n, max(1, n * n // 997)):
            for k in (0, 1):
                assert B.value_by_idx(idx, n, s, f, k) == orc.value_by_idx(idx, n, s, f, k)


def _ref_map_text(nPts, ranges, vals, prec, header=True, fmt=None):
    """The reference's 2D writer (Library/hostLibrary.cu:651-655, :708-725)."""
    fmt = fmt or (lambda v: str(int(v)))
    out = []
    if header:
        out.append(f"{ranges[0]:.{prec}g} {ranges[1]:.{prec}g}\n")
        out.append(f"{ranges[2]:.{prec}g} {ranges[3]:.{prec}g}\n")
    counter = 0
    for v in vals:
        if counter != 0:
            out.append(", ")
        if counter == nPts:
            out.append("\n")
            counter = 0
        out.append(fmt(v))
        counter += 1
    return "".join(out)


def test_map_int_writer_format(tmp_path):
    ranges = [0.05, 0.45, 2, 10]
    vals = np.arange(12, dtype=np.int32) - 3
    p = tmp_path / "m.csv"
    B.write_map_int(p, 4, ranges, vals, header=True, precision=6)
    assert p.read_text() == _ref_map_text(4, ranges, vals, 6)
    assert p.read_text().startswith("0.05 0.45\n2 10\n-3, -2, -1, 0, \n1, 2")
    B.write_map_int(p, 4, [0.123456789012345, 1, 2, 3], vals, header=True, precision=12)
    assert p.read_text().startswith("0.123456789012 1\n")


def test_map_double_writer_format_and_stride(tmp_path):
    ranges = [0.05, 0.45, 2, 10]
    rng = np.random.default_rng(3)
    vals = rng.normal(size=(9, 3)) * 10.0 ** rng.integers(-8, 8, size=(9, 3))
    for k in range(3):
        p = tmp_path / f"ls{k}.csv"
        B.write_map_double(p, 3, ranges, vals.ravel(), precision=6, stride=3, offset=k)
        assert p.read_text() == _ref_map_text(3, ranges, vals[:, k], 6, fmt=lambda v: f"{v:.6g}")


def test_bif1d_writer_format(tmp_path, orc):
    ranges = [2.0, 7.0]
    counts = np.array([2, 0, 3], dtype=np.int32)
    peaks = np.array([[1.5, 2.25, 0], [0, 0, 0], [1 / 3, 2 / 3, 12345.678901234567]])
    p = tmp_path / "b.csv"
    B.write_bif1d(p, 3, ranges, counts, peaks)
    lines = []
    for k in range(3):
        pv = orc.value_by_idx(k, 3, 2.0, 7.0, 0)
        for j in range(counts[k]):
            lines.append(f"{pv:.12g}, {peaks[k, j]:.12g}\n")
    assert p.read_text() == "".join(lines)


def test_1d_values_writer_format(tmp_path, orc):
    ranges = [0.05, 0.35]
    v = np.array([[0.1, -0.02, -5.0], [1e-9, 0, 3.14159265]])
    p = tmp_path / "ls1d.csv"
    B.write_1d_values(p, 2, ranges, v.ravel(), per_cell=3)
    want = ""
    for k in range(2):
        want += f"{orc.value_by_idx(k, 2, 0.05, 0.35, 0):g}" + "".join(f", {x:g}" for x in v[k]) + "\n"
    assert p.read_text() == want


def test_writer_reports_unwritable_path(tmp_path):
    with pytest.raises(B.BifhipError) as e:
        B.write_map_int(tmp_path / "no" / "such" / "dir.csv", 2, [0, 1, 0, 1], [1, 2, 3, 4])
    assert e.value.code == B.E_FILE


def test_context_knob_validation():
    with pytest.raises(B.BifhipError):
        B.set_model(7)
    with pytest.raises(B.BifhipError):
        B.set_precision(5)
    with pytest.raises(B.BifhipError):
        B.set_peak_capacity(1)
    with pytest.raises(B.BifhipError):
        B.set_chaos_confidence(0.4)
    with pytest.raises(B.BifhipError):
        B.set_lle32_margin(float("nan"))
    B.set_lle32_margin(-1.0)          # negative: every regime re-solve in FP64
    B.set_lle32_margin(0.05)
    B.reset_context()
    B.set_output_path("/tmp/x.csv")
    assert B.load().bifhip_get_output_path() == b"/tmp/x.csv"
    B.set_output_path("mat.csv")


MAIN_CPP = r"""
#include "bifhip.hpp"
int main() {
    bifhip_set_output_path(OUT);
    // the reference's own call shape (CudaConsole/CudaConsole/main.cpp templates)
    bifurcation2DDBSCAN(500, 4, 0.01, 3, new double[3]{0.1, 0.1, 0.1},
                        new double[4]{0.05, 0.45, 2, 10}, new int[2]{0, 2}, 0, 0, 32, 1000,
                        new double[3]{0.2, 0.2, 5.7}, 3, 1, 0.01);
    LLE2D(1000, 0.5, 4, 0.01, 1e-8, new double[3]{0.1, 0.1, 0.1}, 3,
          new double[4]{0.05, 0.45, 2, 10}, new int[2]{0, 2}, 0, 0, 500,
          new double[3]{0.2, 0.2, 5.7}, 3);
    return 0;
}
"""


def test_reference_style_caller_compiles_and_links(tmp_path):
    """A main.cpp written against the reference's header compiles against
    include/bifhip.hpp; without a GPU it exits through the reference's
    GPUassert path instead of crashing."""
    src = tmp_path / "main.cpp"
    src.write_text(MAIN_CPP)
    exe = tmp_path / "main"
    inc = os.path.join(os.path.dirname(B.HEADER))
    subprocess.run(["g++", "-std=c++17", f"-DOUT=\"{tmp_path / 'mat.csv'}\"", "-I", inc,
                    str(src), "-o", str(exe), f"-L{os.path.dirname(B.LIB_PATH)}", "-lbifhip",
                    f"-Wl,-rpath,{os.path.dirname(B.LIB_PATH)}"], check=True)
    if B.device_count() > 0:
        return
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 1
    assert "GPUassert: no HIP device visible" in r.stderr


def test_drivers_fail_cleanly_without_device():
    """On a host with no GPU every driver returns BIFHIP_E_NODEV (no crash,
    no CPU fallback)."""
    if B.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(B.BifhipError) as e:
        B.dbscan_map(500, 4, 0.01, [0.1] * 3, [0.05, 0.45, 2, 10], [0, 2], 0, 0.0, 32, 1000,
                     [0.2, 0.2, 5.7], 1, 0.01)
    assert e.value.code == B.E_NODEV


def test_tile_generator_is_seeded_and_in_bounds():
    a = B.tiles_generate(4096, 16, 16, 4096, 0, 1, 200.0, 2000.0, 0.5)
    b = B.tiles_generate(4096, 16, 16, 4096, 0, 1, 200.0, 2000.0, 0.5)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    oi, oj, nb = a
    assert oi.min() >= 0 and oi.max() <= 4096 - 16 and oj.min() >= 0 and oj.max() <= 4096 - 16
    assert nb.min() >= 400 and nb.max() <= 4000            # [200, 2000] t.u. in 0.5 blocks
    assert abs(nb.mean() * 0.5 - 1100) < 40                # uniform mean 1100 t.u.
    # half of the origins cluster: the origin histogram is far from uniform
    h, _ = np.histogram(oj, bins=16, range=(0, 4080))
    assert h.max() > 3 * h.min()
    c = B.tiles_generate(64, 16, 16, 4096, 5, 1, 200.0, 2000.0, 0.5)
    assert not np.array_equal(c[0], oi[:64])


def _header_prototypes():
    """name -> list of parameter declarations, from include/bifhip.h."""
    import re
    txt = open(B.HEADER).read()
    txt = re.sub(r"/\*.*?\*/", " ", txt, flags=re.S)
    protos = {}
    for m in re.finditer(r"\b(bifhip_[A-Za-z0-9_]+)\s*\(([^()]*)\)\s*;", txt):
        params = [p.strip() for p in m.group(2).split(",")]
        protos[m.group(1)] = [] if params == ["void"] or params == [""] else params
    return protos


def test_binding_argtypes_match_header_prototypes():
    """Every function the Python binding declares takes as many arguments as
    its prototype, each of the same kind (pointer, floating, integer)."""
    import ctypes as C
    lib = B.load()
    protos = _header_prototypes()
    checked = 0
    for name, params in protos.items():
        fn = getattr(lib, name)
        if fn.argtypes is None:
            continue
        assert len(fn.argtypes) == len(params), (name, len(fn.argtypes), params)
        for t, p in zip(fn.argtypes, params):
            if "*" in p:
                ok = t in (C.c_void_p, C.c_char_p) or hasattr(t, "_type_") and \
                    issubclass(t, C._Pointer)
            elif p.split()[0] == "double":
                ok = t is C.c_double
            else:
                ok = t in (C.c_int, C.c_long, C.c_longlong, C.c_size_t, C.c_ulonglong,
                           C.c_uint, C.c_ulong)
            assert ok, (name, p, t)
        checked += 1
    assert checked >= 40, checked
