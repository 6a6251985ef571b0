"""Pins the operator oracle (oracle/sdb_oracle.c) against definitions that do
not share its code: Python's comparison / logical operators keyed by the
reference's op strings, big-integer arithmetic reduced mod 2^32 / 2^64, set
and dictionary membership for the joins, a dict group-by, and the calendar for
the date generator.  CPU only."""
import datetime
import operator

import numpy as np
import pytest

from oracle import oracle as O
from oracle import plan_oracle as PO

BOUNDS = [-2**31, -2**31 + 1, -1, 0, 1, 2**31 - 2, 2**31 - 1]
PY_CMP = {"==": operator.eq, "!=": operator.ne, "<>": operator.ne, "<": operator.lt, "<=": operator.le,
          ">": operator.gt, ">=": operator.ge}


def _wrap32(v):
    return ((v + 2**31) % 2**32) - 2**31


def test_comparison_strings_map_to_reference_enum():
    # get_comp_op, kernels/selection.hpp:23-39
    assert [PO._get_comp_op(s) for s in ["==", "!=", "<", "<=", ">", ">="]] == [O.EQ, O.NE, O.LT, O.LE, O.GT, O.GE]
    assert PO._get_comp_op("<>") == O.NE
    assert PO._get_comp_op("=") == O.EQ          # Calcite's "=" is unknown -> EQ
    assert PO._get_comp_op("LIKE") == O.EQ
    assert PO._get_logical_op("AND") == O.AND and PO._get_logical_op("OR") == O.OR
    assert PO._get_logical_op("") == O.NONE and PO._get_logical_op("NOT") == O.NONE
    assert PO._is_filter_logical("AND") and PO._is_filter_logical("") and not PO._is_filter_logical("<=")
    with pytest.raises(ValueError):
        PO._get_binop("%")


@pytest.mark.parametrize("op", ["==", "!=", "<", "<=", ">", ">="])
@pytest.mark.parametrize("logic", ["", "AND", "OR"])
def test_select_lit_exhaustive(op, logic):
    cols, flags, lits = [], [], []
    for a in BOUNDS:
        for v in BOUNDS:
            for f in (0, 1):
                cols.append(a)
                flags.append(f)
                lits.append(v)
    col = np.array(cols, np.int32)
    fl = np.array(flags, np.uint8)
    for v in BOUNDS:
        out = np.full(len(col), 7, np.uint8)
        O.select_lit(col, fl, out, PO._get_comp_op(op), PO._get_logical_op(logic), v)
        for i in range(len(col)):
            c = PY_CMP[op](int(col[i]), v)
            want = {"": c, "AND": bool(fl[i]) and c, "OR": bool(fl[i]) or c}[logic]
            assert out[i] == int(want), (op, logic, col[i], v, fl[i])


@pytest.mark.parametrize("op", ["==", "!=", "<", "<=", ">", ">="])
def test_select_col_and_inplace(op):
    a = np.array([x for x in BOUNDS for _ in BOUNDS], np.int32)
    b = np.array([y for _ in BOUNDS for y in BOUNDS], np.int32)
    fl = np.ones(len(a), np.uint8)
    fl[::3] = 0
    want = np.array([int(bool(f) and PY_CMP[op](int(x), int(y))) for x, y, f in zip(a, b, fl)], np.uint8)
    O.select_col(a, b, fl, fl, PO._get_comp_op(op), O.AND)   # in == out
    assert (fl == want).all()


def test_logical_none_overwrites():
    # logical(NONE, a, b) = b (kernels/selection.hpp:91-92): the input flags
    # are ignored, so a previously rejected row can come back
    inf = np.array([0, 0, 1, 1], np.uint8)
    loc = np.array([0, 1, 0, 1], np.uint8)
    out = np.zeros(4, np.uint8)
    O.logical(inf, loc, out, O.NONE)
    assert out.tolist() == [0, 1, 0, 1]
    O.logical(inf, loc, out, O.AND)
    assert out.tolist() == [0, 0, 0, 1]
    O.logical(inf, loc, out, O.OR)
    assert out.tolist() == [0, 1, 1, 1]


def _cdiv(a, b):
    if b == 0:
        return 0
    q = abs(a) // abs(b)
    return _wrap32(q if (a >= 0) == (b >= 0) else -q)


@pytest.mark.parametrize("op,fn", [(O.MUL, lambda a, b: _wrap32(a * b)), (O.ADD, lambda a, b: _wrap32(a + b)),
                                   (O.SUB, lambda a, b: _wrap32(a - b)), (O.DIV, _cdiv)])
def test_binop_wrapping(op, fn):
    rng = np.random.default_rng(1)
    vals = BOUNDS + [int(x) for x in rng.integers(-2**31, 2**31, 40)] + [7, -7, 3, -3]
    a = np.array([x for x in vals for _ in vals], np.int32)
    b = np.array([y for _ in vals for y in vals], np.int32)
    fl = (rng.random(len(a)) < 0.7).astype(np.uint8)
    r = np.full(len(a), 12345, np.int32)
    O.binop_cc(r, a, b, fl, op)
    for i in range(len(a)):
        want = fn(int(a[i]), int(b[i])) if fl[i] else 12345   # unflagged rows untouched
        assert int(r[i]) == want, (op, a[i], b[i])
    # literal-first / literal-second forms
    r2 = np.full(len(a), 9, np.int32)
    O.binop_lc(r2, 5, b, fl, op)
    assert all(int(r2[i]) == (fn(5, int(b[i])) if fl[i] else 9) for i in range(len(b)))
    r3 = np.full(len(a), 9, np.int32)
    O.binop_cl(r3, a, -3, fl, op)
    assert all(int(r3[i]) == (fn(int(a[i]), -3) if fl[i] else 9) for i in range(len(a)))


def test_semi_join_is_set_membership():
    rng = np.random.default_rng(2)
    keys = rng.permutation(np.arange(100, 1100, dtype=np.int32))   # unique build keys
    bflags = (rng.random(len(keys)) < 0.3).astype(np.uint8)
    mn, mx = int(keys.min()), int(keys.max())
    ht = np.zeros(mx - mn + 1, np.uint8)
    O.build_keys(keys, bflags, ht, mx - mn + 1, mn)
    selected = {int(k) for k, f in zip(keys, bflags) if f}
    probe = rng.integers(0, 1300, 5000).astype(np.int32)            # includes out-of-range keys
    pin = (rng.random(len(probe)) < 0.8).astype(np.uint8)
    out = np.zeros(len(probe), np.uint8)
    O.probe_semi(probe, pin, out, ht, mn, mx)
    want = np.array([int(bool(f) and int(k) in selected) for k, f in zip(probe, pin)], np.uint8)
    assert (out == want).all()


def test_full_join_is_dictionary_lookup():
    rng = np.random.default_rng(3)
    keys = rng.permutation(np.arange(-500, 500, dtype=np.int32))
    vals = rng.integers(-10**6, 10**6, len(keys)).astype(np.int32)
    bflags = (rng.random(len(keys)) < 0.5).astype(np.uint8)
    mn, mx = int(keys.min()), int(keys.max())
    ht = np.zeros(2 * (mx - mn + 1), np.int32)
    O.build_keyvals(keys, vals, bflags, ht, mx - mn + 1, mn)
    d = {int(k): int(v) for k, v, f in zip(keys, vals, bflags) if f}
    probe = rng.integers(-700, 700, 4000).astype(np.int32)
    pin = (rng.random(len(probe)) < 0.9).astype(np.uint8)
    val_out = np.full(len(probe), -77, np.int32)
    out = np.zeros(len(probe), np.uint8)
    O.probe_full(probe, val_out, pin, out, ht, mn, mx)
    for i, k in enumerate(probe):
        hit = bool(pin[i]) and int(k) in d
        assert out[i] == int(hit)
        assert int(val_out[i]) == (d[int(k)] if hit else -77)


def test_build_keys_writes_every_row_flag():
    # BuildKeysHTKernel writes true AND false (join.hpp:31): a later false
    # build row clears a slot
    keys = np.array([5, 6, 7], np.int32)
    ht = np.ones(3, np.uint8)
    O.build_keys(keys, np.array([1, 0, 1], np.uint8), ht, 3, 5)
    assert ht.tolist() == [1, 0, 1]


def test_sum_flagged_mod_2_64_and_accumulates():
    rng = np.random.default_rng(4)
    col = rng.integers(-2**31, 2**31, 100_000).astype(np.int32)
    fl = (rng.random(len(col)) < 0.5).astype(np.uint8)
    acc = np.array([123], np.uint64)
    O.sum_flagged(col, fl, acc)
    want = (123 + sum(int(x) for x, f in zip(col, fl) if f)) % 2**64
    assert int(acc[0]) == want
    # negative total wraps to the two's-complement image
    acc2 = np.zeros(1, np.uint64)
    O.sum_flagged(np.array([-5, 2], np.int32), np.array([1, 1], np.uint8), acc2)
    assert int(acc2[0]) == 2**64 - 3


def test_groupby_dense_matches_dict():
    rng = np.random.default_rng(5)
    n = 200_000
    k0 = rng.integers(3, 8, n).astype(np.int32)       # range 3..7 -> 5
    k1 = rng.integers(-2, 2, n).astype(np.int32)      # range -2..1 -> 4
    agg = rng.integers(-1000, 10**6, n).astype(np.int32)
    fl = (rng.random(n) < 0.3).astype(np.uint8)
    mins, maxs = [3, -2], [7, 1]
    prod = 5 * 4
    sums = np.zeros(prod, np.uint64)
    seen = np.zeros(prod, np.uint32)
    ok0 = np.zeros(prod, np.int32)
    ok1 = np.zeros(prod, np.int32)
    O.groupby_dense([k0, k1], mins, maxs, agg, fl, [ok0, ok1], sums, seen, prod)
    d = {}
    for a, b, v, f in zip(k0.tolist(), k1.tolist(), agg.tolist(), fl.tolist()):
        if f:
            d[(a, b)] = d.get((a, b), 0) + v
    for (a, b), v in d.items():
        s = (a - 3) + (b + 2) * 5                     # first column fastest
        assert seen[s] == 1 and int(sums[s]) == v % 2**64 and ok0[s] == a and ok1[s] == b
    assert int(seen.sum()) == len(d)
    fl2 = np.zeros(prod, np.uint8)
    O.seen_to_flags(seen, fl2)
    assert (fl2 == (seen != 0)).all()


def test_count_and_compact():
    rng = np.random.default_rng(6)
    fl = (rng.random(10_001) < 0.1).astype(np.uint8)
    assert O.count_flags(fl) == int(fl.sum())
    assert (O.compact(fl) == np.nonzero(fl)[0]).all()
    assert O.count_flags(np.zeros(0, np.uint8)) == 0


def test_generators():
    u = O.gen_uniform(100_000, 42, 7, -3, 4)
    assert u.min() == -3 and u.max() == 4
    assert (O.gen_uniform(100_000, 42, 7, -3, 4) == u).all()            # deterministic
    assert not (O.gen_uniform(100_000, 43, 7, -3, 4) == u).all()        # seed matters
    tail = O.gen_uniform(1000, 42, 7, -3, 4, first_row=99_000)          # any row range
    assert (tail == u[99_000:]).all()
    base = O.gen_uniform(1000, 1, 1, 0, 4)
    ch = O.gen_child(base, 1, 2, 5, 0, 4)
    assert ((ch // 5) == base).all()
    dk, yr, ym = O.gen_dates(2556)
    d0 = datetime.date(1992, 1, 1)
    for i in range(2556):
        d = d0 + datetime.timedelta(days=i)
        assert dk[i] == d.year * 10000 + d.month * 100 + d.day and yr[i] == d.year and ym[i] == d.year * 100 + d.month
    assert dk[-1] == 19981230
    lk = O.gen_lookup(50_000, 3, 9, dk)
    assert set(lk.tolist()) <= set(dk.tolist()) and len(set(lk.tolist())) > 2000


def test_minmax():
    rng = np.random.default_rng(7)
    c = rng.integers(-10**9, 10**9, 12345).astype(np.int32)
    assert O.minmax(c) == (int(c.min()), int(c.max()))


def test_fill_i32_closed_form():
    # FillKernel (kernels/projection.hpp:18-21) overwrites every row
    for v in (0, -1, 7, 2**31 - 1, -2**31):
        dst = np.arange(1000, dtype=np.int32)
        O.fill_i32(dst, v)
        assert (dst == v).all() and dst.dtype == np.int32
    O.fill_i32(np.zeros(0, np.int32), 5)


def _slot_bigint(row_keys, mins, maxs, prod):
    """The group slot of kernels/aggregation.hpp:315-321 in Python integers:
    `int hash`, `int mult` wrap at 32 bits, then `hash %= prod_ranges`
    converts hash to uint64 (a negative hash becomes 2^64 + hash) and the
    remainder goes back into the int."""
    def i32(x):
        x &= 0xFFFFFFFF
        return x - (1 << 32) if x >= 1 << 31 else x
    h, m = 0, 1
    for c, lo, hi in zip(row_keys, mins, maxs):
        h = i32(h + i32(i32(c - lo) * m))
        m = i32(m * i32(hi - lo + 1))
    return i32((h % (1 << 64)) % prod)


def test_groupby_slot_outside_range_wraps_through_uint64():
    # keys below the minimum give a negative int hash: (2^64 + hash) % prod,
    # which differs from both C's and Python's signed remainder
    assert _slot_bigint([3], [10], [16], 7) == 2          # 2^64 - 7 = 2 (mod 7)
    assert (-7) % 7 == 0
    rng = np.random.default_rng(9)
    n = 50_000
    # one narrow key column with keys far outside [10, 16], and a wide pair
    # whose mixed-radix index overflows int32
    cases = [
        ([rng.integers(-2**31, 2**31, n).astype(np.int32)], [10], [16], 7),
        ([rng.integers(-100, 100, n).astype(np.int32), rng.integers(0, 2**31, n).astype(np.int32)],
         [-5, 3], [70_000, 70_003], 4375),
        ([rng.integers(-2**31, 2**31, n).astype(np.int32) for _ in range(3)], [-7, 0, 123], [9, 999, 200_000],
         97 * 89),
    ]
    for keys, mins, maxs, prod in cases:
        fl = (rng.random(n) < 0.6).astype(np.uint8)
        agg = rng.integers(-10**6, 10**6, n).astype(np.int32)
        sums = np.zeros(prod, np.uint64)
        seen = np.zeros(prod, np.uint32)
        outk = [np.zeros(prod, np.int32) for _ in keys]
        O.groupby_dense(keys, mins, maxs, agg, fl, outk, sums, seen, prod)
        want_sum, want_first = {}, {}
        cols = [k.tolist() for k in keys]
        for i in range(n):
            if not fl[i]:
                continue
            s = _slot_bigint([c[i] for c in cols], mins, maxs, prod)
            assert 0 <= s < prod
            want_sum[s] = want_sum.get(s, 0) + int(agg[i])
            want_first.setdefault(s, i)
        assert sorted(np.flatnonzero(seen).tolist()) == sorted(want_sum)
        for s, v in want_sum.items():
            assert int(sums[s]) == v % 2**64
            # the oracle stores the keys of the first flagged row of the slot
            assert [int(k[s]) for k in outk] == [c[want_first[s]] for c in cols]
