"""CPU tests of the C ABI: the library loads, exports every declared symbol,
and its host-only parts (options, errors, TSPLIB reader) behave as the
reference's interface.  No GPU compute is called here."""
import gzip
import os
import re
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "f2m.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(f2m_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from f2m_amd import _lib as L
    decl = declared_functions()
    assert decl, "no declarations parsed"
    missing = [s for s in decl if not hasattr(L.lib, s)]
    assert not missing, missing
    assert sorted(L.EXPORTS) == decl


def test_default_opts_match_reference_constants(orc):
    from f2m_amd.solver import default_opts
    o = default_opts()
    r = orc.default_opts()
    for name, _ in type(o)._fields_:
        assert getattr(o, name) == getattr(r, name), name
    assert o.err_grad == 1e-7 and o.err_comb_mult == 1e2 and o.glob_k == 5 and o.factor_max == 5000


def test_strerror_codes():
    from f2m_amd import _lib as L
    assert L.lib.f2m_strerror(0) == b"ok"
    assert b"TSPLIB" in L.lib.f2m_strerror(-3)
    assert L.lib.f2m_version().startswith(b"f2m_hip gfx950")


def _write(tmp, name, text):
    p = os.path.join(tmp, name)
    with open(p, "w", newline="") as f:
        f.write(text)
    return p


def test_tsplib_reader_matches_restatement():
    from f2m_amd.solver import read_tsplib
    from oracle import tsplib
    with tempfile.TemporaryDirectory() as tmp:
        for inst in ["xqf131", "dka1376"]:
            p = os.path.join(tmp, inst + ".tsp")
            with gzip.open(os.path.join(ROOT, "tests", "golden", inst + ".tsp.gz"), "rb") as fi, open(p, "wb") as fo:
                fo.write(fi.read())
            pts = read_tsplib(p)
            lab, x, y, d = tsplib.read(p)
            assert np.array_equal(pts.labels, lab) and np.array_equal(pts.x, x) and np.array_equal(pts.y, y)
        # CRLF, 1e-3 rounding, duplicates, data after EOF ignored
        text = ("NAME : t\r\nNODE_COORD_SECTION\r\n1 0.0004 1.0\r\n2 3.00051 4\r\n3 0.0 1.0\r\n"
                "4 -2.5005 7\r\n5 3.0005 4\r\nEOF\r\n9 9 9\r\n")
        p = _write(tmp, "t.tsp", text)
        pts = read_tsplib(p)
        lab, x, y, d = tsplib.read(p)
        assert pts.labels.tolist() == lab.tolist() == [1, 2, 4]
        assert pts.x.tolist() == x.tolist() and pts.y.tolist() == y.tolist()
        assert pts.n_duplicates == 2
        pts.write_duplicates(os.path.join(tmp, "t.tsp_duplicates"))
        dup = open(os.path.join(tmp, "t.tsp_duplicates")).read()
        assert dup == "node used: 1 duplicates: 3 \nnode used: 2 duplicates: 5 \n"
        # a coordinate rounding to -0.0 equals an earlier +0.0 (and the other way
        # round): the point is a duplicate, as in the reference's hash map
        text = "NODE_COORD_SECTION\n1 0.0 5\n2 -0.0004 5\n3 7 -0.0002\n4 7 0\n5 1 1\nEOF\n"
        p = _write(tmp, "z.tsp", text)
        pts = read_tsplib(p)
        lab, x, y, d = tsplib.read(p)
        assert pts.labels.tolist() == lab.tolist() == [1, 3, 5]
        assert pts.n_duplicates == 2 and d == {1: [2], 3: [4]}


def test_tsplib_reader_errors():
    from f2m_amd import _lib as L
    from f2m_amd.solver import read_tsplib
    with tempfile.TemporaryDirectory() as tmp:
        p = _write(tmp, "bad.tsp", "NAME : x\nTYPE : TSP\n1 2 3\nEOF\n")
        try:
            read_tsplib(p)
            raise AssertionError("no error")
        except L.F2MError as e:
            assert "-3" in str(e)
        try:
            read_tsplib(os.path.join(tmp, "missing.tsp"))
            raise AssertionError("no error")
        except L.F2MError as e:
            assert "-4" in str(e)


def test_no_library_sorts_or_scans_in_the_product():
    """Every sort and scan of the product is written here: no hipCUB,
    rocPRIM or Thrust in the HIP sources."""
    csrc = os.path.join(ROOT, "f2m_amd", "csrc")
    for f in sorted(os.listdir(csrc)):
        text = open(os.path.join(csrc, f)).read()
        for lib in ("hipcub", "rocprim", "thrust"):
            assert not re.search(r"#\s*include\s*[<\"]" + lib, text), (f, lib)
