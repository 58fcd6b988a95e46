"""CPU-side checks of the C ABI library: it loads, exports every symbol
include/scycle_hip.h declares, and its host-only input utilities (the
reference's .in format and PETSc binary Vec files) behave like the reference.
No GPU work is done here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "scycle_hip.h")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sc_[A-Za-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from scycle_amd import _lib
    L = _lib.load()
    names = _declared()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_petsc_vec_fixture_from_reference_tests():
    """tests/testIO/test_io of the reference: two appended Vec[3] records,
    [1,1,1] then [2,2,2] (big-endian, classid 1211214)."""
    from scycle_amd._lib import check, load
    p = os.path.join(GOLDEN, "petsc_vec_test_io.bin").encode()
    n = C.c_int()
    check(load().sc_vec_count(p, C.byref(n)))
    assert n.value == 2
    buf = np.zeros(8)
    for rec, val in ((0, 1.0), (1, 2.0)):
        check(load().sc_vec_read(p, rec, buf.ctypes.data_as(C.POINTER(C.c_double)), 8, C.byref(n)))
        assert n.value == 3 and np.all(buf[:3] == val)


def test_petsc_vec_roundtrip(tmp_path):
    from scycle_amd._lib import check, load
    p = str(tmp_path / "v").encode()
    rng = np.random.default_rng(0)
    a, b = rng.standard_normal(17), rng.standard_normal(17)
    dp = C.POINTER(C.c_double)
    check(load().sc_vec_write(p, a.ctypes.data_as(dp), 17, 0))
    check(load().sc_vec_write(p, b.ctypes.data_as(dp), 17, 1))
    raw = open(p, "rb").read()
    assert len(raw) == 2 * (8 + 17 * 8)
    assert raw[:4] == (1211214).to_bytes(4, "big") and raw[4:8] == (17).to_bytes(4, "big")
    assert np.frombuffer(raw[8:8 + 136], dtype=">f8").tolist() == a.tolist()
    out = np.zeros(17)
    n = C.c_int()
    check(load().sc_vec_read(p, 1, out.ctypes.data_as(dp), 17, C.byref(n)))
    assert np.array_equal(out, b)


def _write_deck(tmp_path, lines):
    p = tmp_path / "t.in"
    p.write_text("\n".join(lines) + "\n")
    return str(p).encode()


def test_in_parser_quirks(tmp_path):
    from scycle_amd._lib import load
    deck = _write_deck(tmp_path, [
        "# comment line",
        "Ny = 201 # points in y-direction",
        "Ny = 7",
        "aVals = [0.01 0.02 0.03]  # list",
        "bVals = [0.02]",
        "stateLaw = agingLaw  trailing",
        "timeIntInds = [psi slip]",
    ])
    buf = C.create_string_buffer(64)
    assert load().sc_in_get(deck, b"stateLaw", buf, 64) == 0 and buf.value == b"agingLaw"
    assert load().sc_in_get(deck, b"Ny", buf, 64) == 0
    vals = np.zeros(8)
    n = C.c_int()
    dp = C.POINTER(C.c_double)
    assert load().sc_in_get_vector(deck, b"aVals", vals.ctypes.data_as(dp), 8, C.byref(n)) == 0
    assert n.value == 3 and np.allclose(vals[:3], [0.01, 0.02, 0.03])
    # a one-element list parses as empty (genFuncs.cpp:426-434)
    assert load().sc_in_get_vector(deck, b"bVals", vals.ctypes.data_as(dp), 8, C.byref(n)) == 0
    assert n.value == 0
    assert load().sc_in_get(deck, b"missing", buf, 64) != 0


@pytest.mark.parametrize("name", ["ex1.in", "ex2.in"])
def test_reference_example_decks_parse(name):
    src = os.path.join("/root/reference/examples", name)
    if not os.path.exists(src):
        pytest.skip("reference examples not present")
    from scycle_amd._lib import load
    buf = C.create_string_buffer(64)
    assert load().sc_in_get(src.encode(), b"stateLaw", buf, 64) == 0
    assert buf.value in (b"agingLaw", b"slipLaw")
