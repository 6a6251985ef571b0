"""Plan-level CPU oracle: the reference's executor and TransientTable model.

TEST INFRASTRUCTURE ONLY (the checker; see oracle/oracle.py).

This restates, in Python over the C operator restatements of
oracle/sdb_oracle.c, what the reference does with a physical plan:

* ``parse_execution_info`` / ``dag_topological_sort``:
  operations/preprocessing.cpp:14-275
* ``execute_plan``: executor/ddor_executor.cpp:22-245
* ``_TransientTable``: models/transient_table.hpp:20-1928 (flag double
  buffer, ``apply_filter`` 750-981, ``apply_project`` 983-1224,
  ``apply_aggregate`` 1226-1647, ``apply_join`` 1649-1927)
* ``_Column``: models/models.hpp:1072-1649 (segmented storage, per-segment
  min/max, ``get_min_max`` 1257-1272, hash-table builders 1413-1548)
* ``format_result``: executor/result_writer.cpp:41-67

The reference defers kernels into per-segment bundles and flushes them at
joins/aggregates; every deferred kernel is row-local (or a commutative
integer reduction) and no buffer it touches is modified by host code before
the flush, so executing each operator eagerly, segment by segment, in program
order yields the same buffers.  The segment loop is kept (``segment_size``)
so the segmented code paths of the reference are exercised at small sizes.
"""
from collections import deque

import numpy as np

from . import oracle as O

SEGMENT_SIZE = 1 << 28  # common.hpp:12

# kernels/types.cpp:13-21
TABLE_COLUMN_NUMBERS = {"lineorder": 17, "part": 9, "supplier": 7, "customer": 8, "ddate": 17}


class OracleError(RuntimeError):
    pass


def _stoi(s):
    """std::stoi: leading whitespace, optional sign, decimal digits."""
    t = str(s).lstrip()
    i = 0
    if i < len(t) and t[i] in "+-":
        i += 1
    j = i
    while j < len(t) and t[j].isdigit():
        j += 1
    if j == i:
        raise OracleError("stoi: no conversion for %r" % (s,))
    v = int(t[:j])
    if v < -2**31 or v > 2**31 - 1:
        raise OracleError("stoi: out of range %r" % (s,))
    return v


def _to_int32(v):
    """(int) of an int64 literal (transient_table.hpp:882, 1005)."""
    return ((int(v) + 2**31) % 2**32) - 2**31


def _get_comp_op(op):
    """kernels/selection.hpp:23-39 (unknown, including "=", -> EQ)."""
    return {"==": O.EQ, ">=": O.GE, "<=": O.LE, "<": O.LT, ">": O.GT,
            "!=": O.NE, "<>": O.NE}.get(op, O.EQ)


def _get_logical_op(op):
    """kernels/selection.hpp:41-49 (unknown or "" -> NONE)."""
    return {"AND": O.AND, "OR": O.OR}.get(op, O.NONE)


def _is_filter_logical(op):
    """kernels/selection.hpp:53-59: every character alphabetic ("" too)."""
    return all(c.isalpha() for c in op)


def _get_binop(op):
    """kernels/aggregation.hpp:37-48 (unknown -> invalid_argument)."""
    try:
        return {"*": O.MUL, "/": O.DIV, "+": O.ADD, "-": O.SUB}[op]
    except KeyError:
        raise ValueError("Unknown operation: " + str(op))


class _Column:
    """models/models.hpp:1072-1649: a column cut into segments of
    ``segment_size`` rows, each with its own min/max."""

    def __init__(self, data, segment_size, is_aggregate=False, minmax=None):
        self.data = data
        self.is_aggregate = is_aggregate
        self.segment_size = segment_size
        n = len(data)
        self.bounds = [(s, min(s + segment_size, n)) for s in range(0, n, segment_size)]
        if minmax == "compute":  # Segment ctor, models.hpp:71-77
            self.seg_minmax = [O.minmax(data[a:b]) for a, b in self.bounds]
        elif minmax is None:     # unset (materialised, filled later)
            self.seg_minmax = [(0, 0) for _ in self.bounds]
        else:
            self.seg_minmax = [minmax for _ in self.bounds]

    def seg(self, i):
        a, b = self.bounds[i]
        return self.data[a:b]

    def get_min_max(self):
        """models/models.hpp:1257-1272"""
        if self.is_aggregate:
            raise OracleError("can't get min/max of aggregate result column")
        mn = min(m for m, _ in self.seg_minmax)
        mx = max(m for _, m in self.seg_minmax)
        return mn, mx


class _EmptyColumn:
    """A base column the table does not hold (models.hpp:1670-1673)."""
    is_aggregate = False

    def __init__(self, name):
        self.name = name

    def __getattr__(self, item):
        raise OracleError("column %s is not loaded" % self.name)


class _TransientTable:
    """models/transient_table.hpp:20-1928, one device, eager execution."""

    def __init__(self, base_columns, nrows, segment_size):
        self.nrows = nrows
        self.segment_size = segment_size
        # ctor: flags filled with true (transient_table.hpp:84-106)
        self.flags = np.ones(nrows, dtype=np.uint8)
        self.flags_secondary = None
        self.current_columns = list(base_columns)
        self.group_by_column = None
        self.group_by_column_index = 0

    def nseg(self):
        return (self.nrows + self.segment_size - 1) // self.segment_size

    def seg_range(self, i):
        a = i * self.segment_size
        return a, min(a + self.segment_size, self.nrows)

    def set_group_by_column(self, col):
        """transient_table.hpp:189-193"""
        self.group_by_column = self.current_columns[col]
        self.group_by_column_index = col

    def _ensure_secondary(self):
        if self.flags_secondary is None:
            self.flags_secondary = np.empty(self.nrows, dtype=np.uint8)

    def _swap(self):
        """transient_table.hpp:300-310"""
        self.flags, self.flags_secondary = self.flags_secondary, self.flags

    def _col(self, idx):
        c = self.current_columns[idx]
        if c is None:
            raise OracleError("column %d is null (dropped by a join)" % idx)
        return c

    # -------------------------------------------------------------- filter
    def apply_filter(self, expr, parent_op):
        """transient_table.hpp:750-981"""
        if expr.get("exprType") != "EXPR":
            return  # the reference prints an error and ignores it (757-761)
        op = expr.get("op", "")
        if op == "SEARCH":
            self._apply_search(expr, parent_op)
        elif _is_filter_logical(op):
            # left fold: parent op to the first operand, own op to the rest
            used = False
            for operand in expr.get("operands", []):
                self.apply_filter(operand, op if used else parent_op)
                used = True
        else:
            ops = expr.get("operands", [])
            if len(ops) != 2:
                return  # reference: error message, no kernel (863-867)
            cols = [None, None]
            literal = None
            for i in range(2):
                kind = ops[i]["exprType"]
                if kind == "COLUMN":
                    cols[i] = self._col(ops[i]["input"])
                elif kind == "LITERAL":
                    literal = _to_int32(ops[i]["literal"]["value"])
                else:
                    return
            if cols[0] is None:
                raise OracleError("comparison needs a column as first operand")
            cmp = _get_comp_op(op)
            logic = _get_logical_op(parent_op)
            self._ensure_secondary()
            for s in range(self.nseg()):
                a, b = self.seg_range(s)
                if literal is not None and cols[1] is None:
                    O.select_lit(cols[0].seg(s), self.flags[a:b], self.flags_secondary[a:b],
                                 cmp, logic, literal)
                else:
                    O.select_col(cols[0].seg(s), cols[1].seg(s), self.flags[a:b],
                                 self.flags_secondary[a:b], cmp, logic)
            self._swap()

    def _apply_search(self, expr, parent_op):
        """SEARCH: Segment::search_operator, models/models.hpp:604-718."""
        col = self._col(expr["operands"][0]["input"])
        rs = expr["operands"][1]["literal"]["rangeSet"]
        self._ensure_secondary()
        logic = _get_logical_op(parent_op)
        for s in range(self.nseg()):
            a, b = self.seg_range(s)
            data = col.seg(s)
            local = np.empty(b - a, dtype=np.uint8)
            if len(rs) == 1:
                lower, upper = _stoi(rs[0][1]), _stoi(rs[0][2])
                O.select_lit(data, local, local, O.GE, O.NONE, lower)
                O.select_lit(data, local, local, O.LE, O.AND, upper)
            else:
                first, second = _stoi(rs[0][1]), _stoi(rs[1][1])
                O.select_lit(data, local, local, O.EQ, O.NONE, first)
                O.select_lit(data, local, local, O.EQ, O.OR, second)
            O.logical(self.flags[a:b], local, self.flags_secondary[a:b], logic)
        self._swap()

    # ------------------------------------------------------------- project
    def apply_project(self, exprs):
        """transient_table.hpp:983-1224"""
        new_columns = []
        for i, expr in enumerate(exprs):
            kind = expr["exprType"]
            if kind == "COLUMN":
                new_columns.append(self.current_columns[expr["input"]])
                if expr["input"] == self.group_by_column_index:
                    self.group_by_column_index = i
            elif kind == "LITERAL":
                v = _to_int32(expr["literal"]["value"])
                data = np.empty(self.nrows, dtype=np.int32)
                O.fill_i32(data, v)
                new_columns.append(_Column(data, self.segment_size, minmax=(v, v)))
            else:
                ops = expr["operands"]
                if len(ops) != 2:
                    raise OracleError("Project operation: unsupported number of operands")
                binop = _get_binop(expr["op"])
                data = np.zeros(self.nrows, dtype=np.int32)
                newc = _Column(data, self.segment_size)
                k0, k1 = ops[0]["exprType"], ops[1]["exprType"]
                for s in range(self.nseg()):
                    a, b = self.seg_range(s)
                    f = self.flags[a:b]
                    if k0 == "COLUMN" and k1 == "COLUMN":
                        ca, cb = self._col(ops[0]["input"]), self._col(ops[1]["input"])
                        O.binop_cc(data[a:b], ca.seg(s), cb.seg(s), f, binop)
                        (mna, mxa), (mnb, mxb) = ca.seg_minmax[s], cb.seg_minmax[s]
                        newc.seg_minmax[s] = (min(mna, mnb), max(mxa, mxb))  # models.hpp:774-775
                    elif k0 == "LITERAL" and k1 == "COLUMN":
                        cb = self._col(ops[1]["input"])
                        O.binop_lc(data[a:b], _to_int32(ops[0]["literal"]["value"]), cb.seg(s), f, binop)
                        newc.seg_minmax[s] = cb.seg_minmax[s]  # models.hpp:832-833
                    elif k0 == "COLUMN" and k1 == "LITERAL":
                        ca = self._col(ops[0]["input"])
                        O.binop_cl(data[a:b], ca.seg(s), _to_int32(ops[1]["literal"]["value"]), f, binop)
                        newc.seg_minmax[s] = ca.seg_minmax[s]  # models.hpp:803-804
                    else:
                        raise OracleError("Project operation: unsupported operand kinds")
                new_columns.append(newc)
        self.current_columns = new_columns

    # ----------------------------------------------------------- aggregate
    def apply_aggregate(self, agg, group):
        """transient_table.hpp:1226-1647 (always a SUM of operands[0])"""
        if not agg.get("operands"):
            raise OracleError("aggregate without operand")
        agg_col = self._col(agg["operands"][0])
        if len(group) == 0:
            acc = np.zeros(1, dtype=np.uint64)
            for s in range(self.nseg()):
                a, b = self.seg_range(s)
                O.sum_flagged(agg_col.seg(s), self.flags[a:b], acc)
            self.nrows = 1
            self.flags = np.ones(1, dtype=np.uint8)
            self.flags_secondary = None
            self.current_columns = [_Column(acc, self.segment_size, is_aggregate=True)]
            return
        keys = [self._col(g) for g in group]
        mins, maxs = [], []
        prod = 1
        for c in keys:
            mn, mx = c.get_min_max()
            mins.append(mn)
            maxs.append(mx)
            # prod_ranges *= max - min + 1 (int expression into uint64)
            prod = (prod * (((mx - mn + 1) + 2**31) % 2**32 - 2**31)) % 2**64
        if prod > (1 << 34):
            raise OracleError("group-by range product too large: %d" % prod)
        sums = np.zeros(prod, dtype=np.uint64)
        seen = np.zeros(prod, dtype=np.uint32)
        out_keys = [np.zeros(prod, dtype=np.int32) for _ in keys]
        for s in range(self.nseg()):
            a, b = self.seg_range(s)
            O.groupby_dense([c.seg(s) for c in keys], mins, maxs, agg_col.seg(s),
                            self.flags[a:b], out_keys, sums, seen, prod)
        flags = np.empty(prod, dtype=np.uint8)
        O.seen_to_flags(seen, flags)
        self.nrows = prod
        self.flags = flags
        self.flags_secondary = None
        # result key columns carry min = max = 0 (models.hpp:182-183)
        self.current_columns = [_Column(k, self.segment_size, minmax=(0, 0)) for k in out_keys]
        self.current_columns.append(_Column(sums, self.segment_size, is_aggregate=True))

    # ---------------------------------------------------------------- join
    def apply_join(self, right, rel):
        """transient_table.hpp:1649-1927"""
        cond = rel["condition"]
        left_column = cond["operands"][0]["input"]
        right_column = cond["operands"][1]["input"] - len(self.current_columns)
        if not (0 <= left_column < len(self.current_columns)
                and 0 <= right_column < len(right.current_columns)):
            raise ValueError("Invalid column indices in join condition.")
        probe = self._col(left_column)
        build = right._col(right_column)
        bmin, bmax = build.get_min_max()
        ht_len = _to_int32(bmax - bmin + 1)
        if ht_len <= 0:
            raise OracleError("hash table length overflows int")
        self._ensure_secondary()
        if rel.get("joinType") == "semi":
            ht = np.zeros(ht_len, dtype=np.uint8)  # alloc_zero, models.hpp:1424
            for s in range(right.nseg()):
                a, b = right.seg_range(s)
                O.build_keys(build.seg(s), right.flags[a:b], ht, ht_len, bmin)
            for s in range(self.nseg()):
                a, b = self.seg_range(s)
                O.probe_semi(probe.seg(s), self.flags[a:b], self.flags_secondary[a:b], ht, bmin, bmax)
            self._swap()
            self.current_columns.extend([None] * len(right.current_columns))
            return
        if right.group_by_column is None:
            raise OracleError("full join needs a group-by column on the build side")
        vals = right.group_by_column
        gb_min, gb_max = vals.get_min_max()
        new_data = np.zeros(self.nrows, dtype=np.int32)
        new_col = _Column(new_data, self.segment_size, minmax=(gb_min, gb_max))
        gb_index = len(self.current_columns) + right.group_by_column_index
        self.current_columns.extend([None] * len(right.current_columns))
        self.current_columns[gb_index] = new_col
        ht = np.zeros(2 * ht_len, dtype=np.int32)  # alloc_zero, models.hpp:1525
        for s in range(right.nseg()):
            a, b = right.seg_range(s)
            O.build_keyvals(build.seg(s), vals.seg(s), right.flags[a:b], ht, ht_len, bmin)
        for s in range(self.nseg()):
            a, b = self.seg_range(s)
            O.probe_full(probe.seg(s), new_data[a:b], self.flags[a:b], self.flags_secondary[a:b],
                         ht, bmin, bmax)
        self._swap()


# ------------------------------------------------------------------ planning

def _expr_columns(expr, out):
    """operations/preprocessing.cpp:5-12"""
    kind = expr.get("exprType")
    if kind == "COLUMN":
        out.add(expr["input"])
    elif kind == "EXPR":
        for o in expr.get("operands", []):
            _expr_columns(o, out)


def dag_topological_sort(rels):
    """operations/preprocessing.cpp:14-102 (the lineorder scan is deferred
    while other sources remain)."""
    n = len(rels)
    dag = [[0] * n for _ in range(n)]
    S = deque()
    for rel in rels:
        op = rel["relOp"]
        if op == "TABLE_SCAN":
            S.appendleft(rel["id"])
        elif op in ("FILTER", "PROJECT", "AGGREGATE", "SORT"):
            dag[rel["id"] - 1][rel["id"]] = 1
        elif op == "JOIN":
            if len(rel["inputs"]) != 2:
                return []
            dag[rel["inputs"][0]][rel["id"]] = 1
            dag[rel["inputs"][1]][rel["id"]] = 1
        else:
            return []
    order = []
    while S:
        node = S.pop()
        if S and rels[node]["relOp"] == "TABLE_SCAN" and rels[node]["tables"][1] == "lineorder":
            S.appendleft(node)
            continue
        order.append(node)
        for i in range(n):
            if dag[node][i] == 1:
                dag[node][i] = 0
                if not any(dag[j][i] == 1 for j in range(n)):
                    S.appendleft(i)
    if len(order) != n:
        return []
    return order


def parse_execution_info(rels):
    """operations/preprocessing.cpp:104-275 (the parts that steer execution:
    group-by column per table and the DAG order)."""
    ops_info = []
    group_by_columns = {}
    for rel in rels:
        op = rel["relOp"]
        if op == "TABLE_SCAN":
            t = rel["tables"][1]
            ops_info.append([(t, i) for i in range(TABLE_COLUMN_NUMBERS[t])])
        elif op == "FILTER" or op == "SORT":
            ops_info.append(list(ops_info[-1]))
        elif op == "PROJECT":
            last = ops_info[-1]
            info = []
            for e in rel["exprs"]:
                cols = set()
                _expr_columns(e, cols)
                if cols:
                    info.append(last[min(cols)])
            ops_info.append(info)
        elif op == "AGGREGATE":
            last = ops_info[-1]
            info = []
            for g in rel["group"]:
                t, c = last[g]
                info.append((t, c))
                group_by_columns[t] = c
            for a in rel["aggs"]:
                info.append(last[a["operands"][0]])
            ops_info.append(info)
        elif op == "JOIN":
            ops_info.append(list(ops_info[rel["inputs"][0]]) + list(ops_info[rel["inputs"][1]]))
    return group_by_columns, dag_topological_sort(rels)


class OracleResult:
    def __init__(self, table):
        self.nrows = table.nrows
        self.flags = table.flags.copy()
        self.columns = []   # (is_aggregate, ndarray) per non-null column
        for c in table.current_columns:
            if c is None or isinstance(c, _EmptyColumn):
                continue
            self.columns.append((c.is_aggregate, c.data))

    def rows(self):
        """Selected rows as tuples of Python ints (result_writer.cpp:41-67)."""
        idx = np.nonzero(self.flags)[0]
        cols = [(agg, d[idx]) for agg, d in self.columns]
        out = []
        for r in range(len(idx)):
            out.append(tuple(int(d[r]) for _, d in cols))
        return out

    def text(self):
        """The .res text of executor/result_writer.cpp:41-67."""
        return "".join(" ".join(str(v) for v in row) + "\n" for row in self.rows())


def execute_plan(plan, tables, segment_size=SEGMENT_SIZE):
    """executor/ddor_executor.cpp:22-245.

    plan: dict with "rels" (the JSON mirror of calciteserver.thrift PlanResult)
    tables: {table name: {column index: np.int32 array}} (all the same length)
    """
    rels = plan["rels"]
    group_by_columns, order = parse_execution_info(rels)
    if not order:
        raise OracleError("DAG topological sort failed")
    out_table = [-1] * len(rels)
    tts = []
    for rel in rels:
        if rel["relOp"] != "TABLE_SCAN":
            continue
        name = rel["tables"][1]
        cols = tables[name]
        nrows = len(next(iter(cols.values())))
        base = []
        for i in range(TABLE_COLUMN_NUMBERS[name]):
            if i in cols:
                base.append(_Column(np.ascontiguousarray(cols[i], dtype=np.int32), segment_size,
                                    minmax="compute"))
            else:
                base.append(_EmptyColumn("%s%d" % (name, i)))
        tt = _TransientTable(base, nrows, segment_size)
        if name in group_by_columns:
            tt.set_group_by_column(group_by_columns[name])
        tts.append(tt)
        out_table[rel["id"]] = len(tts) - 1
    for rid in order:
        rel = rels[rid]
        op = rel["relOp"]
        if op == "FILTER":
            t = out_table[rid - 1]
            tts[t].apply_filter(rel["condition"], "")
            out_table[rid] = t
        elif op == "PROJECT":
            t = out_table[rid - 1]
            tts[t].apply_project(rel["exprs"])
            out_table[rid] = t
        elif op == "AGGREGATE":
            t = out_table[rid - 1]
            tts[t].apply_aggregate(rel["aggs"][0], rel["group"])
            out_table[rid] = t
        elif op == "JOIN":
            lt, rt = out_table[rel["inputs"][0]], out_table[rel["inputs"][1]]
            tts[lt].apply_join(tts[rt], rel)
            out_table[rid] = lt
        elif op == "SORT":
            out_table[rid] = out_table[rid - 1]
    return OracleResult(tts[out_table[len(rels) - 1]])
