"""The C-ABI library loads here (no GPU), exports every function that
include/sycldb.h declares, and its host-only entry points (operator-string
parsing, plan parsing and DAG ordering) agree with the oracle."""
import ctypes
import json
import os
import re

import pytest

from oracle import plan_oracle as PO
from sycldb_amd import _lib as L
from sycldb_amd import engine as E
from sycldb_amd import plans as PL

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "sycldb.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(sdb_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 40
    for n in names:
        assert hasattr(L.lib, n), n
    # and every binding in _lib is declared in the header
    assert set(L.SIGNATURES) <= set(names)


def test_op_strings_follow_reference():
    f = lambda s: L.lib.sdb_comp_op_from_string(s.encode())
    assert [f(s) for s in ["==", "!=", "<>", "<", "<=", ">", ">=", "=", "?"]] == [0, 1, 1, 2, 3, 4, 5, 0, 0]
    g = lambda s: L.lib.sdb_logical_op_from_string(s.encode())
    assert [g(s) for s in ["AND", "OR", "", "NONE", "NOT"]] == [1, 2, 0, 0, 0]
    h = lambda s: L.lib.sdb_binop_from_string(s.encode())
    assert [h(s) for s in ["*", "/", "+", "-", "%"]] == [0, 1, 2, 3, -1]


@pytest.mark.parametrize("tmpl", ["t1", "t2", "t3", "t4"])
def test_dag_order_matches_oracle(tmpl):
    plan = PL.build(tmpl)
    assert E.dag_order(plan) == PO.dag_topological_sort(plan["rels"])


def test_bad_plans_raise_value_error():
    with pytest.raises(ValueError):
        E.dag_order("{not json")
    with pytest.raises(ValueError):
        E.dag_order('{"rels": [{"id": 3, "relOp": "TABLE_SCAN", "tables": ["ssb", "ddate"]}]}')
    with pytest.raises(ValueError):
        E.dag_order('{"rels": [{"id": 0, "relOp": "FILTER", "condition": {"exprType": "EXPR", "op": "=", '
                    '"operands": []}}]}')


def test_plan_json_reading():
    """The plan reader: key order is free, a repeated key keeps its last
    value, escaped strings read, deep nesting and malformed text raise."""
    base = PL.build("t1")
    want = PO.dag_topological_sort(base["rels"])
    # members in reverse order, whitespace around every token
    flipped = {"rels": [dict(reversed(list(r.items()))) for r in base["rels"]]}
    assert E.dag_order(json.dumps(flipped, indent=2)) == want
    # a repeated "id" key: the last one counts (the first would break the
    # ids' order, which raises)
    text = json.dumps(base)
    rel1 = json.dumps(base["rels"][1])
    doubled = rel1.replace('"id": 1', '"id": 7, "id": 1', 1)
    assert doubled != rel1
    assert E.dag_order(text.replace(rel1, doubled)) == want
    # escaped strings (an output name: read, not used, by the ordering)
    esc = text.replace('"name": "revenue"', '"name": "rev\\"en\\\\ue\\u0041"', 1)
    assert esc != text and E.dag_order(esc) == want
    for bad in ["[" * 100000, '{"rels": [' + "{" * 300, text[:-1], text + "x", '{"rels": [{"id": 1.5}]}',
                '{"rels": [{"id": 99999999999999999999}]}', '{"rels": "x"}', '{"rels": [{"id": 0, "relOp": 3}]}']:
        with pytest.raises(ValueError):
            E.dag_order(bad)


def test_codegen_shutdown_without_a_compiler():
    """The Python package stops the background compiler at every interpreter
    exit: with none started (no device here) the call does nothing and
    succeeds, also twice."""
    assert L.lib.sdb_codegen_shutdown() == 0
    assert L.lib.sdb_codegen_shutdown() == 0
    assert L.lib.sdb_codegen_pending() == 0


def test_version_and_last_error():
    assert b"sycldb" in L.lib.sdb_version()
    assert isinstance(L.lib.sdb_last_error(), bytes)


def test_date_dimension_matches_oracle_calendar():
    import numpy as np
    from oracle import oracle as O
    dk, yr, ym = E.date_columns(2556)
    odk, oyr, oym = O.gen_dates(2556)
    assert (dk == odk).all() and (yr == oyr).all() and (ym == oym).all()


@pytest.mark.parametrize("order", ["library_first", "torch_first"])
def test_one_hip_runtime_per_process(order):
    """Loading the library and torch in either order maps a single HIP and a
    single HSA runtime (two runtimes in one process cannot both open the GPU)."""
    import subprocess
    import sys
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "%s\n"
        "maps = open('/proc/self/maps').read().splitlines()\n"
        "hip = {l.split()[-1] for l in maps if 'libamdhip64' in l}\n"
        "hsa = {l.split()[-1] for l in maps if 'libhsa-runtime64' in l}\n"
        "print(len(hip), len(hsa))\n"
    ) % (ROOT, "from sycldb_amd import _lib\nimport torch" if order == "library_first"
         else "import torch\nfrom sycldb_amd import _lib")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    assert out.stdout.split() == ["1", "1"], out.stdout


def test_codegen_compiles_every_instruction_form():
    """The generator's source for programs covering every instruction form,
    operand kind and chunk width compiles for gfx950 (hipRTC, no device)."""
    ok, why = E.codegen_available()
    if not ok:
        pytest.skip("runtime compilation unavailable: " + why)
    n = ctypes.c_int(0)
    L.check(L.lib.sdb_codegen_check(b"gfx950", ctypes.byref(n)))
    # seven single-program kernels, the two of them with a row tail also
    # without its queue, and one batched kernel
    assert n.value == 10


def test_queued_unroll_flag_matches_header():
    text = open(os.path.join(ROOT, "include", "sycldb.h")).read()
    assert re.search(r"#define SDB_UNROLL_QUEUED (\d+)", text).group(1) == str(E.UNROLL_QUEUED) == "16"
    # the flag goes with 1, 2 or 4 only (arguments are checked before the
    # engine, which is null here)
    for bad in (E.UNROLL_QUEUED, 3 | E.UNROLL_QUEUED, 8, 32 | 4):
        with pytest.raises(ValueError, match="unroll"):
            L.check(L.lib.sdb_engine_set_codegen(None, 0, bad))
    for good in (0, 1, 4 | E.UNROLL_QUEUED, 2 | E.UNROLL_QUEUED):
        with pytest.raises(ValueError, match="null engine"):
            L.check(L.lib.sdb_engine_set_codegen(None, 0, good))


def test_codegen_mode_validation():
    with pytest.raises(ValueError):
        E.Engine.set_codegen(E.Engine.__new__(E.Engine), "bogus")


API_CHECK = os.path.join(ROOT, "sycldb_amd", "sdb_api_check")


def test_cxx_layer_host_checks():
    """include/sycldb.hpp (the reference-named C++ layer) compiles, links and
    maps op strings and errors as the reference does (tests/cxx/api_check.cpp)."""
    import subprocess
    if not os.path.exists(API_CHECK):
        pytest.skip("api check program not built")
    out = subprocess.run([API_CHECK, "host"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "ok"


@pytest.mark.gpu
def test_cxx_layer_device_pipeline():
    """The reference-named C++ layer runs a selection / join / projection /
    aggregation / group-by pipeline on the GPU and matches values computed on
    the host."""
    import subprocess
    assert os.path.exists(API_CHECK), "api check program not built"
    print("cxx layer device pipeline", flush=True)
    out = subprocess.run([API_CHECK, "device"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "ok"


def test_sdb_op_layout_matches_ctypes(tmp_path):
    """The ctypes mirror of sdb_op has the C struct's size and offsets."""
    import subprocess
    fields = [f for f, _ in L.Op._fields_]
    src = tmp_path / "op_layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "sycldb.h"\nint main(void) {\n'
                   '    printf("%zu\\n", sizeof(sdb_op));\n' +
                   "".join('    printf("%%zu\\n", offsetof(sdb_op, %s));\n' % f for f in fields) + "    return 0;\n}\n")
    exe = tmp_path / "op_layout"
    subprocess.run(["gcc", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got[0] == ctypes.sizeof(L.Op)
    assert got[1:] == [getattr(L.Op, f).offset for f in fields]


def test_bundle_exec_rejects_bad_operators():
    """Checked on the host before any device work."""
    ops = (L.Op * 2)()
    ops[0].kind = 99
    with pytest.raises(ValueError, match="unknown operator kind"):
        L.check(L.lib.sdb_bundle_exec(ops, 1, 10, None))
    ops[0].kind = L.OP_SELECT_LIT  # null column
    with pytest.raises(ValueError, match="null"):
        L.check(L.lib.sdb_bundle_exec(ops, 1, 10, None))
    with pytest.raises(ValueError, match="operator count"):
        L.check(L.lib.sdb_bundle_exec(ops, L.BUNDLE_MAX_OPS + 1, 10, None))
    ops[0].kind = L.OP_GROUPBY
    ops[0].nkeys = 9
    with pytest.raises(ValueError, match="key count"):
        L.check(L.lib.sdb_bundle_exec(ops, 1, 10, None))
    # nothing to do: no device needed
    L.check(L.lib.sdb_bundle_exec(ops, 0, 10, None))
