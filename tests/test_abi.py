"""CPU checks of the C ABI: the library loads, exports every entry point that
include/dsm.h declares, its own DSM parser agrees with the oracle's restatement
of io.cu:122-252 on the fixtures, and malformed inputs fail with error codes
(no process exit, unlike the reference's exit(1), io.cu:135-140)."""
import os
import re
import subprocess

import numpy as np
import pytest

from dsm_amd import engine as E
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "dsm.h")).read()
    return sorted(set(re.findall(r"\b(dsm_[a-z_]+)\s*\(", txt)))


def test_header_matches_binding_list():
    assert header_functions() == sorted(E.EXPORTS)


def test_library_exports_every_header_symbol():
    L = E.lib()
    for name in header_functions():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", E.library_path()], stdout=subprocess.PIPE,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (dsm_\w+)", out))
    assert set(header_functions()) <= exported


@pytest.mark.parametrize("name", ["example", "small"])
def test_parser_agrees_with_oracle(golden_dir, name):
    gname = {"example": "example_g_01.in", "small": "small_g.in"}[name]
    qname = {"example": "example_q_01.in", "small": "small_q.in"}[name]
    dp = os.path.join(golden_dir, gname)
    qp = os.path.join(golden_dir, qname)
    f = E.DsmFile(dp, False)
    o = O.DsmFile(dp, False)
    n, lab, eu, ev = f.graph()
    assert n == o.n and lab.tolist() == o.labels.tolist()
    assert eu.tolist() == o.eu.tolist() and ev.tolist() == o.ev.tolist()
    assert f.num_batches == o.nb
    for b in range(o.nb):
        a1 = f.batch(b)
        a2 = o.batch(b)
        for x, y in zip(a1, a2):
            assert np.asarray(x).tolist() == np.asarray(y).tolist()
    qf = E.DsmFile(qp, True)
    qo = O.DsmFile(qp, True)
    qn, ql, qu, qv = qf.graph()
    assert (qn, ql.tolist(), qu.tolist(), qv.tolist()) == (qo.n, qo.labels.tolist(), qo.eu.tolist(), qo.ev.tolist())


def test_static_format_is_a_graph_without_batches(golden_dir, tmp_path):
    """The static format (t/v/e lines only, io.cu:27-120) parses, in data and
    in query mode, to the dynamic file's initial graph with no batches."""
    src = open(os.path.join(golden_dir, "small_g.in")).read()
    p = tmp_path / "static.in"
    p.write_text(src[:src.index("-")])
    want = E.DsmFile(os.path.join(golden_dir, "small_g.in"), False).graph()
    for is_query in (False, True):
        f = E.DsmFile(str(p), is_query)
        n, lab, eu, ev = f.graph()
        assert n == want[0] and lab.tolist() == want[1].tolist()
        assert eu.tolist() == want[2].tolist() and ev.tolist() == want[3].tolist()
        assert f.num_batches == 0


def test_edge_labels_are_an_optional_third_field(golden_dir, tmp_path):
    """e lines may carry an edge label (the R-MAT workloads' 4 or 8 uniform
    edge labels): on some lines or all, in data, static and query files, the
    graph and the batches are those of the file without them, and the labels
    read back in file order (0 where a line has none)."""
    src = open(os.path.join(golden_dir, "small_g.in")).read().split("\n")
    plain = E.DsmFile(os.path.join(golden_dir, "small_g.in"), False)
    want = plain.graph()
    assert plain.edge_labels() is None
    rng = np.random.default_rng(7)
    elines = [i for i, l in enumerate(src) if l.startswith("e ")]
    for every in (1, 3):
        lab = rng.integers(0, 8, len(elines))
        out = list(src)
        expect = []
        for j, i in enumerate(elines):
            if j % every == 0:
                out[i] = "%s %d" % (src[i], lab[j])
                expect.append(int(lab[j]))
            else:
                expect.append(0)
        text = "\n".join(out)
        for name, body, is_query in (("dyn.in", text, False), ("static.in", text[:text.index("-")], False),
                                     ("query.in", text[:text.index("-")], True)):
            p = tmp_path / name
            p.write_text(body)
            f = E.DsmFile(str(p), is_query)
            g = f.graph()
            assert g[0] == want[0] and all(x.tolist() == y.tolist() for x, y in zip(g[1:], want[1:]))
            assert f.edge_labels().tolist() == expect
            if name == "dyn.in":
                assert f.num_batches == plain.num_batches
                for b in range(f.num_batches):
                    assert all(np.array_equal(x, y) for x, y in zip(f.batch(b), plain.batch(b)))
    # an edge label cannot stand in for a missing endpoint
    p = tmp_path / "short.in"
    p.write_text("t 2 1\nv 0 0 1\nv 1 0 1\ne 0\n-\n")
    with pytest.raises(E.DsmError):
        E.DsmFile(str(p), False)


BAD = {
    "no_header": "v 0 0 0\n",
    "degree_mismatch": "t 2 1\nv 0 0 2\nv 1 0 1\ne 0 1\n-\n",
    "ids_out_of_order": "t 2 1\nv 1 0 1\nv 0 0 1\ne 0 1\n-\n",
    "edge_out_of_range": "t 2 1\nv 0 0 1\nv 1 0 1\ne 0 2\n-\n",
    "missing_edges": "t 2 2\nv 0 0 1\nv 1 0 1\ne 0 1\n-\n",
    "bad_change": "t 2 1\nv 0 0 1\nv 1 0 1\ne 0 1\n-\nb 0 1\nx 0 1 0\n",
    "batch_id_gap": "t 2 1\nv 0 0 1\nv 1 0 1\ne 0 1\n-\nb 1 1\na 0 1 0\n",
    "change_out_of_range": "t 2 1\nv 0 0 1\nv 1 0 1\ne 0 1\n-\nb 0 1\na 0 5 0\n",
    "bad_number": "t 2 1\nv 0 0 1\nv 1 zero 1\ne 0 1\n-\n",
}


@pytest.mark.parametrize("case", sorted(BAD))
def test_malformed_files_fail_with_code(tmp_path, case):
    p = tmp_path / "bad.in"
    p.write_text(BAD[case])
    with pytest.raises(E.DsmError) as ei:
        E.DsmFile(str(p), False)
    assert ei.value.code == E.DSM_EFORMAT


def test_missing_file_is_eio(tmp_path):
    with pytest.raises(E.DsmError) as ei:
        E.DsmFile(str(tmp_path / "nope.in"), False)
    assert ei.value.code == E.DSM_EIO


def test_whitespace_insensitive(tmp_path):
    p = tmp_path / "ws.in"
    p.write_text("t 3 2 v 0 1 1 v 1 2 2\n\n v 2 1 1 e 0 1 e 1 2 - b 0 2 a 0 2 7 d 1 2 0")
    f = E.DsmFile(str(p), False)
    n, lab, eu, ev = f.graph()
    assert n == 3 and lab.tolist() == [1, 2, 1] and f.num_batches == 1
    add, s, d, l = f.batch(0)
    assert add.tolist() == [1, 0] and s.tolist() == [0, 1] and d.tolist() == [2, 2] and l.tolist() == [7, 0]


def test_create_without_device_fails_cleanly():
    import ctypes
    L = E.lib()
    cnt = ctypes.c_int(0)
    # skip where a GPU is visible; elsewhere dsm_create must return DSM_EDEVICE, not abort
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        if hip.hipGetDeviceCount(ctypes.byref(cnt)) == 0 and cnt.value > 0:
            pytest.skip("a GPU is visible")
    except OSError:
        pass
    with pytest.raises(E.DsmError) as ei:
        E.Engine((2, [0, 0], [0], [1]), [(2, [0, 0], [0], [1])])
    assert ei.value.code == E.DSM_EDEVICE


def test_cli_usage_error():
    exe = os.path.join(ROOT, "dsm_amd", "bin", "CDSM")
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert r.returncode == 1 and "usage" in r.stderr


def test_intersect_rejects_unsorted_input():
    # validated on the host before any device call (no GPU needed)
    from dsm_amd import engine as E
    import pytest as _pt
    for a, b in (([3, 2], [1]), ([1, 1], [1]), ([1], [0xFFFFFFFF])):
        with _pt.raises(E.DsmError) as ei:
            E.intersect_sorted(a, b)
        assert ei.value.code == E.DSM_EINVAL
    assert E.intersect_sorted([], [1, 2]).size == 0


STRUCTS = [("Change", "dsm_change"), ("GraphDesc", "dsm_graph_desc"), ("QueryDesc", "dsm_query_desc"),
           ("Options", "dsm_options"), ("TuneReport", "dsm_tune_report"), ("Stats", "dsm_stats"),
           ("ServerOptions", "dsm_server_options"), ("RequestInfo", "dsm_request_info")]


def test_ctypes_structs_match_header_layout(tmp_path):
    """Every ctypes mirror has the C struct's size and its fields' offsets, as
    gcc lays out include/dsm.h."""
    import ctypes
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "dsm.h"', "int main(void) {"]
    for py, c in STRUCTS:
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (py, c))
        for f, _ in getattr(E, py)._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (py, f, c, f))
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True, check=True).stdout.split("\n")
    got = dict((" ".join(l.split()[:-1]), int(l.split()[-1])) for l in out if l)
    for py, _ in STRUCTS:
        cls = getattr(E, py)
        assert got["%s size" % py] == ctypes.sizeof(cls), py
        for f, _ in cls._fields_:
            assert got["%s %s" % (py, f)] == getattr(cls, f).offset, (py, f)
