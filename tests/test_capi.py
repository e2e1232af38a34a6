"""The C-ABI library loads and exports every entry point of include/piv_hip.h
(CPU only: no compute call)."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from fluidimage_amd import _native

HEADER = Path(__file__).resolve().parents[1] / "include" / "piv_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(piv_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "piv_correlate_batch" in names and "piv_plan_create" in names
    assert set(names) == set(_native.SIGNATURES), "ctypes signatures drift from the header"


def declared_prototypes():
    """name -> list of parameter declarations, from the header text."""
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    protos = {}
    for m in re.finditer(r"\b(piv_[a-z_0-9]+)\s*\(([^)]*)\)\s*;", text):
        params = [p.strip() for p in m.group(2).split(",") if p.strip()]
        protos[m.group(1)] = [] if params == ["void"] else params
    return protos


def _ctype_of(decl):
    decl = decl.replace("const ", "")
    if "*" in decl:
        return "p"
    kind = decl.split()[0]
    return {"int32_t": "i32", "int64_t": "i64", "double": "f64", "float": "f32"}[kind]


def test_ctypes_argument_lists_match_the_header():
    codes = {ctypes.c_int32: "i32", ctypes.c_int64: "i64", ctypes.c_double: "f64",
             ctypes.c_float: "f32"}
    protos = declared_prototypes()
    assert set(protos) == set(_native.SIGNATURES)
    for name, params in protos.items():
        argtypes = _native.SIGNATURES[name][1]
        assert len(argtypes) == len(params), name
        for decl, at in zip(params, argtypes):
            want = _ctype_of(decl)
            got = codes.get(at, "p")
            assert want == got, (name, decl, at)


def test_library_exports_every_symbol():
    lib = _native.load()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_struct_layouts():
    assert ctypes.sizeof(_native.PivImage) == 40
    assert ctypes.sizeof(_native.PivWindow) == 40
    assert ctypes.sizeof(_native.PivResults) == 80


def test_plan_validation_without_gpu():
    lib = _native.load()
    assert lib.piv_shape_supported(32, 32) == 1
    # shapes without a register-FFT kernel take the DFT kernel of any shape
    for shape in [(33, 32), (7, 9), (12, 13), (21, 32), (96, 96), (256, 64)]:
        assert lib.piv_shape_supported(*shape) == 1, shape
    for shape in [(1, 32), (257, 8), (160, 160)]:
        assert lib.piv_shape_supported(*shape) == 0, shape
    plan = ctypes.c_void_p()
    st = lib.piv_plan_create(ctypes.byref(plan), 33, 33, -1.0, None, None, 0, 1, 3, 1)
    assert st == 0
    assert lib.piv_plan_destroy(plan) == 0
    # unsupported shape: status PIV_EUNSUPPORTED with a message
    st = lib.piv_plan_create(ctypes.byref(plan), 160, 160, -1.0, None, None, 0, 1, 3, 1)
    assert st == 2 and b"not supported" in lib.piv_last_error()
    # displacement_max None needs no device memory: the plan can be made
    st = lib.piv_plan_create(ctypes.byref(plan), 32, 32, -1.0, None, None, 0, 1, 3, 1)
    assert st == 0
    assert lib.piv_plan_destroy(plan) == 0
    st = lib.piv_plan_create(ctypes.byref(plan), 32, 32, -1.0, None, None, 0, 1, 3, 9)
    assert st == 2
    # 2d_gaussian2 takes no nsubpix (subpix.py:76-81): any value but 1 is refused
    for nsub in (2, 10):
        st = lib.piv_plan_create(ctypes.byref(plan), 32, 32, -1.0, None, None, 0, nsub, 3, 1)
        assert st == 1 and b"2d_gaussian2" in lib.piv_last_error()
    # centroid does take it
    st = lib.piv_plan_create(ctypes.byref(plan), 32, 32, -1.0, None, None, 1, 2, 3, 1)
    assert st == 0
    assert lib.piv_plan_destroy(plan) == 0


def test_tuning_interface_without_gpu():
    """Launch tunables: candidates per window shape, get/set, rejected values
    (host-side state of the library; nothing is launched)."""
    from fluidimage_amd import tuning

    for n in tuning.DEFAULT_SHAPES:
        cands = tuning.candidates(tuning.CORRELATE_WPB, n, n)
        assert len(cands) >= 2 and len(set(cands)) == len(cands), (n, cands)
        before = tuning.get(tuning.CORRELATE_WPB, n, n)
        assert before in cands
        for c in cands:
            tuning.set_value(tuning.CORRELATE_WPB, n, n, c)
            assert tuning.get(tuning.CORRELATE_WPB, n, n) == c
        tuning.set_value(tuning.CORRELATE_WPB, n, n, before)
    with pytest.raises(RuntimeError, match="not among the candidates"):
        tuning.set_value(tuning.CORRELATE_WPB, 32, 32, 7)
    with pytest.raises(ValueError, match="not built"):
        tuning.candidates(tuning.CORRELATE_WPB, 33, 33)
    assert tuning.candidates(tuning.TPS_CHOL_THREADS) == [0, 256, 512]
    with pytest.raises(RuntimeError, match="0, 256 or 512"):
        tuning.set_value(tuning.TPS_CHOL_THREADS, 0, 0, 128)
    assert tuning.candidates(tuning.TPS_SHARED_SOLVE) == [0, 1]
    with pytest.raises(RuntimeError, match="0 or 1"):
        tuning.set_value(tuning.TPS_SHARED_SOLVE, 0, 0, 2)
    lib = _native.load()
    assert lib.piv_tune_get(99, 0, 0) < 0


def test_tuning_table_is_valid_for_the_library():
    """The committed tuning table (if any) names only built candidates."""
    import json

    from fluidimage_amd import tuning

    path = _native.TUNING_PATH
    if not path.exists():
        pytest.skip("no tuning table")
    table = json.loads(path.read_text())
    for key, wpb in table["correlate_wpb"].items():
        n0, n1 = (int(v) for v in key.split("x"))
        assert wpb in tuning.candidates(tuning.CORRELATE_WPB, n0, n1), key
    assert table["tps_chol_threads"] in tuning.candidates(tuning.TPS_CHOL_THREADS)


def test_tps_matrix_size_matches_host_offsets():
    """The host's TPS matrix offsets use the library's region size: (n-3) x 3
    then (n-3) rows of (n-3) doubles, each padded to 16 doubles."""
    import numpy as np

    from fluidimage_amd.calcul.tps import matrix_sizes

    lib = _native.load()
    ns = np.arange(0, 2049)
    lib_sizes = np.array([lib.piv_tps_matrix_size(int(n)) for n in ns])
    assert np.array_equal(lib_sizes, matrix_sizes(ns))
    m = np.maximum(ns - 3, 0)
    ld = -(-m // 16) * 16
    assert np.all(lib_sizes % 16 == 0) and np.all(ld >= m) and np.all(ld - m < 16)
    assert np.all(lib_sizes >= 3 * m + m * m)
    assert lib.piv_tps_matrix_size(373) == 1120 + 370 * 384
    assert lib.piv_tps_matrix_size(-1) < 0 and lib.piv_tps_matrix_size(2049) < 0


def test_graph_argument_checks_without_gpu():
    """The capture entry points reject what cannot be captured or replayed
    before any runtime call."""
    lib = _native.load()
    assert lib.piv_graph_begin(None) == 1  # PIV_EINVAL: null stream
    assert "null stream" in lib.piv_last_error().decode()
    assert lib.piv_graph_end(None, None) == 1
    assert lib.piv_graph_launch(None, None) == 1
    assert lib.piv_graph_destroy(None) == 0


def test_displacements_from_indices_plural_name():
    """compute_displacements_from_indices takes (ix, iy), scalars or arrays,
    and gives (ix0 - ix, iy0 - iy) (correl.py:209-211); a 32x32 plan without
    displacement_max needs no GPU."""
    from fluidimage_amd.calcul.correl import CorrelHIP

    c = CorrelHIP((32, 32), method_subpix="2d_gaussian2", nsubpix=None)
    assert (c.iy0, c.ix0) == (15, 15)
    assert c.compute_displacements_from_indices(10, 20) == (5, -5)
    dx, dy = c.compute_displacements_from_indices(np.array([15, 0]), np.array([15, 31]))
    assert dx.tolist() == [0, 15] and dy.tolist() == [0, -16]
    with pytest.raises(ValueError, match="2d_gaussian2"):
        CorrelHIP((32, 32), method_subpix="2d_gaussian2", nsubpix=10)
