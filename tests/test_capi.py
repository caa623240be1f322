"""Host-side checks of the C ABI that need no GPU: the library loads, exports every entry point
include/itolbm.h declares, and the Python binding mirrors the header's structs."""
import ctypes as C
import os
import re

import pytest

import _pkg

ROOT = _pkg.ROOT


def header_functions():
    txt = open(os.path.join(ROOT, "include", "itolbm.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_\w+)\s*\(", txt)))


def test_library_exports_header_symbols():
    m = _pkg.load()
    L = m.lib()
    names = header_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(L, n), n
    assert sorted(m.SYMBOLS) == names


def _header_struct(name):
    txt = open(os.path.join(ROOT, "include", "itolbm.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    body = re.search(r"typedef struct \{([^{}]*)\}\s*" + name + ";", txt).group(1)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        typ, names = re.match(r"(long long|double|int)\s+(.*)", decl).groups()
        for n in names.split(","):
            fields.append((n.strip(), typ))
    return fields


@pytest.mark.parametrize("cname,pyname", [("lbm_case_desc", "CaseDesc"), ("lbm_drop_diag", "DropDiag"),
                                          ("lbm_incl_diag", "InclDiag")])
def test_struct_layout_matches_header(cname, pyname):
    m = _pkg.load()
    py = getattr(m, pyname)
    tmap = {"double": C.c_double, "int": C.c_int, "long long": C.c_longlong}
    hdr = _header_struct(cname)
    assert [n for n, _ in hdr] == [n for n, _ in py._fields_]
    for (n, t), (pn, pt) in zip(hdr, py._fields_):
        assert tmap[t] == pt, n


def test_operator_parsing_matches_reference_defaults():
    m = _pkg.load()
    assert m.parse_op("P3") == (3, 0.5, 1.569988)
    assert m.parse_op("P3 -0.4") == (3, -0.4, 1.569988)
    assert m.parse_op("F3 3") == (13, 0.5, 3.0)
    assert m.parse_op("P1 7") == (1, 0.5, 1.569988)
    with pytest.raises(ValueError):
        m.parse_op("P2")


def test_create_rejects_bad_arguments_without_touching_a_device():
    m = _pkg.load()
    L = m.lib()
    h = C.c_void_p()
    bad = m.case_desc(10, 10)
    bad.op = 7
    # argument checks that fail before any device query or allocation
    assert L.lbm_create(None, 1, 0, C.byref(h)) == -1
    assert L.lbm_create((m.CaseDesc * 1)(bad), 0, 0, C.byref(h)) == -1
    assert L.lbm_create((m.CaseDesc * 1)(bad), 1, 5, C.byref(h)) == -1
    assert L.lbm_last_error(None) == b"null context"
    # every entry point reports a null context as an argument error
    assert L.lbm_step(None, 1, 0) == -1
    assert L.lbm_step_case(None, 0, 1, 0) == -1
    assert L.lbm_set_reference(None, 0, 1.0, 1.0) == -1
    assert L.lbm_sync(None) == -1
    assert L.lbm_set_graph(None, 4) == -1
