"""Every operator of the C ABI on the GPU against the CPU oracle, bit for bit,
on seeded inputs: empty, tiny and ragged lengths, aligned and misaligned
pointers (vector and row paths), out-of-range probe keys, in-place flags,
both group-by strategies."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402

SIZES = [0, 1, 3, 4, 5, 1023, 4099, 100_003]
OFFSETS = [0, 1, 4]  # 16-byte aligned, misaligned, 4-byte aligned flags (16-byte aligned int32)


@pytest.fixture(scope="module")
def ops():
    from sycldb_amd import ops as ops_mod
    return ops_mod


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def padded(a, off):
    """Device copy of `a` starting `off` elements into a larger buffer (so
    the pointer the kernel sees is misaligned when off > 0)."""
    buf = np.zeros(len(a) + off + 4, dtype=a.dtype)
    buf[off:off + len(a)] = a
    t = dev(buf)
    return t, t[off:off + len(a)]


def rand_cols(rng, n, lo=-50, hi=50):
    return rng.integers(lo, hi, n).astype(np.int32)


def rand_flags(rng, n, p=0.6):
    return (rng.random(n) < p).astype(np.uint8)


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("off", OFFSETS)
def test_select_lit_all_ops(ops, n, off):
    rng = np.random.default_rng(n * 7 + off)
    col, fl = rand_cols(rng, n), rand_flags(rng, n)
    _, c = padded(col, off)
    _, f = padded(fl, off)
    for cmp in range(6):
        for logic in range(3):
            _, out = padded(np.full(n, 9, np.uint8), off)
            ops.select_lit(c, f, out, cmp, logic, 3)
            want = np.full(n, 9, np.uint8)
            O.select_lit(col, fl, want, cmp, logic, 3)
            assert (host(out) == want).all(), (cmp, logic)


@pytest.mark.parametrize("n", [5, 4099, 100_003])
def test_select_inplace_and_columns(ops, n):
    rng = np.random.default_rng(n)
    a, b, fl = rand_cols(rng, n), rand_cols(rng, n), rand_flags(rng, n)
    ta, tb, tf = dev(a), dev(b), dev(fl)
    for cmp in range(6):
        for logic in range(3):
            t = tf.clone()
            ops.select_col(ta, tb, t, t, cmp, logic)          # in == out
            want = fl.copy()
            O.select_col(a, b, want, want, cmp, logic)
            assert (host(t) == want).all()
    # the reference free function: selection(flags, arr, op, value, parent_op)
    t = tf.clone()
    ops.selection(t, ta, "<=", 7, "AND", n)
    want = fl.copy()
    O.select_lit(a, want, want, O.LE, O.AND, 7)
    assert (host(t) == want).all()


@pytest.mark.parametrize("n", [0, 15, 16, 17, 100_003])
@pytest.mark.parametrize("off", OFFSETS)
def test_logical_any_nonzero_byte(ops, n, off):
    """Flags are bool bytes: any nonzero byte is true (0x80, 0xff, 2, ...),
    through the 16-flags-per-thread path (aligned) and the byte path."""
    rng = np.random.default_rng(12)
    a = np.where(rng.random(n) < 0.5, rng.integers(1, 256, n), 0).astype(np.uint8)
    b = np.where(rng.random(n) < 0.5, rng.integers(1, 256, n), 0).astype(np.uint8)
    a[: min(n, 4)] = [0x80, 0xFF, 0x01, 0x7F][: min(n, 4)]
    for logic in range(3):
        _, ta = padded(a, off)
        _, tb = padded(b, off)
        buf, out = padded(np.zeros(n, np.uint8), off)
        ops.logical(ta, tb, out, logic)
        want = np.zeros(n, np.uint8)
        O.logical(a, b, want, logic)
        got = host(buf)
        assert (got[off:off + n] == want).all(), logic
        assert (got[:off] == 0).all() and (got[off + n:] == 0).all()


@pytest.mark.parametrize("n", [0, 7, 100_003])
def test_logical_and_fill(ops, n):
    rng = np.random.default_rng(11)
    a, b = rand_flags(rng, n), rand_flags(rng, n, 0.3)
    for logic in range(3):
        out = torch.zeros(n, dtype=torch.uint8, device="cuda")
        ops.logical(dev(a), dev(b), out, logic)
        want = np.zeros(n, np.uint8)
        O.logical(a, b, want, logic)
        assert (host(out) == want).all()
    for off in OFFSETS:
        buf, t = padded(np.zeros(n, np.int32), off)
        ops.fill(t, -123)
        got = host(buf)
        assert (got[off:off + n] == -123).all() and (got[:off] == 0).all() and (got[off + n:] == 0).all()


@pytest.mark.parametrize("n", [0, 1, 5, 4099, 100_003])
@pytest.mark.parametrize("off", OFFSETS)
def test_binops(ops, n, off):
    rng = np.random.default_rng(n + 3 * off)
    a = rng.integers(-2**31, 2**31, n).astype(np.int32)
    b = rng.integers(-2**31, 2**31, n).astype(np.int32)
    b[::7] = 0           # division by zero is defined as 0
    b[::11] = -1
    a[::13] = -2**31     # INT_MIN / -1 wraps
    fl = rand_flags(rng, n)
    _, ta = padded(a, off)
    _, tb = padded(b, off)
    _, tf = padded(fl, off)
    for op in range(4):
        for kind in range(3):
            _, r = padded(np.full(n, 555, np.int32), off)
            want = np.full(n, 555, np.int32)
            if kind == 0:
                ops.binop(r, ta, tb, tf, op)
                O.binop_cc(want, a, b, fl, op)
            elif kind == 1:
                ops.binop(r, -7, tb, tf, op)
                O.binop_lc(want, -7, b, fl, op)
            else:
                ops.binop(r, ta, -3, tf, op)
                O.binop_cl(want, a, -3, fl, op)
            assert (host(r) == want).all(), (op, kind)
    with pytest.raises(ValueError):
        ops.perform_operation(r, ta, tb, tf, n, "%")


@pytest.mark.parametrize("hoff", [0, 1])
@pytest.mark.parametrize("length", [1, 31, 777, 2500])
def test_probe_semi_tables(ops, hoff, length):
    """Semi probes over key tables of odd lengths, a table pointer off
    16-byte alignment, any nonzero byte present, keys on both sides of the
    range, a row count off 4."""
    rng = np.random.default_rng(length + hoff)
    mn = -37
    table = np.where(rng.random(length) < 0.4, rng.integers(1, 256, length), 0).astype(np.uint8)
    _, ht = padded(table, hoff)
    n = 300_001
    probe = rng.integers(mn - 40, mn + length + 40, n).astype(np.int32)
    pfl = rand_flags(rng, n, 0.7)
    out = torch.full((n,), 3, dtype=torch.uint8, device="cuda")
    ops.probe_semi(dev(probe), dev(pfl), out, ht, mn, mn + length - 1)
    want = np.zeros(n, np.uint8)
    O.probe_semi(probe, pfl, want, table, mn, mn + length - 1)
    assert (host(out) == want).all()


@pytest.mark.parametrize("nb,np_", [(1, 10), (1000, 100_003), (20_000, 4099)])
def test_join_build_and_probe(ops, nb, np_):
    rng = np.random.default_rng(nb)
    keys = (rng.permutation(nb) + 17).astype(np.int32)
    vals = rng.integers(-10**6, 10**6, nb).astype(np.int32)
    bfl = rand_flags(rng, nb, 0.4)
    mn, mx = int(keys.min()), int(keys.max())
    length = mx - mn + 1
    # semi
    ht = torch.zeros(length, dtype=torch.uint8, device="cuda")
    ops.build_keys_ht(dev(keys), dev(bfl), nb, ht, length, mn)
    want_ht = np.zeros(length, np.uint8)
    O.build_keys(keys, bfl, want_ht, length, mn)
    assert (host(ht) == want_ht).all()
    probe = rng.integers(mn - 50, mx + 50, np_).astype(np.int32)
    pfl = rand_flags(rng, np_, 0.8)
    for off in OFFSETS:
        _, tp = padded(probe, off)
        _, tf = padded(pfl, off)
        _, out = padded(np.full(np_, 3, np.uint8), off)
        ops.probe_semi(tp, tf, out, ht, mn, mx)
        want = np.zeros(np_, np.uint8)
        O.probe_semi(probe, pfl, want, want_ht, mn, mx)
        assert (host(out) == want).all()
    # full
    ht2 = torch.zeros(2 * length, dtype=torch.int32, device="cuda")
    ops.build_key_vals_ht(dev(keys), dev(vals), dev(bfl), nb, ht2, length, mn)
    want_ht2 = np.zeros(2 * length, np.int32)
    O.build_keyvals(keys, vals, bfl, want_ht2, length, mn)
    assert (host(ht2) == want_ht2).all()
    for off in OFFSETS:
        _, tp = padded(probe, off)
        _, tf = padded(pfl, off)
        _, out = padded(np.full(np_, 3, np.uint8), off)
        _, vo = padded(np.full(np_, -99, np.int32), off)
        ops.probe_full(tp, vo, tf, out, ht2, mn, mx)
        want = np.zeros(np_, np.uint8)
        want_v = np.full(np_, -99, np.int32)
        O.probe_full(probe, want_v, pfl, want, want_ht2, mn, mx)
        assert (host(out) == want).all() and (host(vo) == want_v).all()
    # in-place free function (join.hpp:314-355)
    tf = dev(pfl)
    vo = torch.full((np_,), -99, dtype=torch.int32, device="cuda")
    ops.full_join(dev(probe), vo, tf, np_, ht2, mn, mx)
    assert (host(tf) == want).all()


@pytest.mark.parametrize("n", [0, 1, 5, 4099, 1_000_003])
@pytest.mark.parametrize("off", OFFSETS)
def test_sum_flagged(ops, n, off):
    rng = np.random.default_rng(n + off)
    col = rng.integers(-2**31, 2**31, n).astype(np.int32)
    fl = rand_flags(rng, n)
    _, tc = padded(col, off)
    _, tf = padded(fl, off)
    acc = torch.tensor([42], dtype=torch.int64, device="cuda")
    ops.sum_flagged(tc, tf, acc)
    ops.aggregate_operation(tc, tf, n, acc)         # accumulates, no reset
    want = np.array([42], np.uint64)
    O.sum_flagged(col, fl, want)
    O.sum_flagged(col, fl, want)
    assert int(host(acc).view(np.uint64)[0]) == int(want[0])


@pytest.mark.parametrize("ranges,n", [([(0, 24), (0, 24), (1992, 1998)], 300_007), ([(1992, 1998), (0, 999)], 200_003),
                                      ([(0, 299_999)], 100_000), ([(5, 5)], 10), ([(-3, 3), (10, 12)], 0)])
def test_groupby_dense(ops, ranges, n):
    rng = np.random.default_rng(n + len(ranges))
    keys = [rng.integers(lo, hi + 1, n).astype(np.int32) for lo, hi in ranges]
    agg = rng.integers(-10**6, 10**6, n).astype(np.int32)
    fl = rand_flags(rng, n, 0.2)
    mins = [lo for lo, _ in ranges]
    maxs = [hi for _, hi in ranges]
    prod = int(np.prod([hi - lo + 1 for lo, hi in ranges]))
    sums = torch.zeros(prod, dtype=torch.int64, device="cuda")
    seen = torch.zeros(prod, dtype=torch.int32, device="cuda")
    outk = [torch.full((prod,), -1, dtype=torch.int32, device="cuda") for _ in ranges]
    ops.group_by_aggregate([dev(k) for k in keys], dev(agg), maxs, mins, dev(fl), n, len(keys), outk, sums, seen,
                           prod)
    w_sums = np.zeros(prod, np.uint64)
    w_seen = np.zeros(prod, np.uint32)
    w_keys = [np.full(prod, -1, np.int32) for _ in ranges]
    O.groupby_dense(keys, mins, maxs, agg, fl, w_keys, w_sums, w_seen, prod)
    assert (host(sums).view(np.uint64) == w_sums).all()
    s = host(seen).view(np.uint32)
    assert (s == w_seen).all()
    for a, b in zip(outk, w_keys):
        ha = host(a)
        assert (ha[s != 0] == b[s != 0]).all()
    tflags = torch.zeros(prod, dtype=torch.uint8, device="cuda")
    ops.seen_to_flags(seen, tflags, prod)
    assert (host(tflags) == (w_seen != 0)).all()


def test_groupby_out_of_range_keys(ops):
    # keys outside [min, max] wrap through % prod exactly as the reference
    rng = np.random.default_rng(5)
    n = 50_000
    k0 = rng.integers(-5, 30, n).astype(np.int32)
    agg = rng.integers(0, 100, n).astype(np.int32)
    fl = np.ones(n, np.uint8)
    prod = 25
    sums = torch.zeros(prod, dtype=torch.int64, device="cuda")
    seen = torch.zeros(prod, dtype=torch.int32, device="cuda")
    outk = [torch.zeros(prod, dtype=torch.int32, device="cuda")]
    ops.groupby_dense([dev(k0)], [0], [24], dev(agg), dev(fl), outk, sums, seen, prod)
    w_sums = np.zeros(prod, np.uint64)
    w_seen = np.zeros(prod, np.uint32)
    O.groupby_dense([k0], [0], [24], agg, fl, [np.zeros(prod, np.int32)], w_sums, w_seen, prod)
    assert (host(sums).view(np.uint64) == w_sums).all() and (host(seen).view(np.uint32) == w_seen).all()


@pytest.mark.parametrize("n", [0, 1, 5, 4099, 1_000_003])
def test_count_compact_gather_minmax(ops, n):
    rng = np.random.default_rng(n)
    fl = rand_flags(rng, n, 0.1)
    assert ops.count_true_flags(dev(fl), n) == O.count_flags(fl)
    ids = torch.full((max(n, 1),), -1, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    ops.compact(dev(fl), ids, cnt, n=n)
    c = int(host(cnt)[0])
    want = O.compact(fl)
    assert c == len(want) and (host(ids)[:c] == want).all()
    if n:
        col = rand_cols(rng, n, -10**9, 10**9)
        dst = torch.zeros(max(c, 1), dtype=torch.int32, device="cuda")
        ops.gather(dev(col), ids, c, dst)
        assert (host(dst)[:c] == col[want]).all()
        assert ops.minmax(dev(col)) == O.minmax(col)


def test_generators_match_oracle():
    import ctypes
    from sycldb_amd import _lib as L
    n = 100_003
    out = torch.zeros(n, dtype=torch.int32, device="cuda")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = ctypes.c_void_p(out.data_ptr())
    L.check(L.lib.sdb_gen_uniform(p, n, 42, 77, 5, -9, 1_000_000, s))
    assert (host(out) == O.gen_uniform(n, 42, 77, -9, 1_000_000, first_row=5)).all()
    base = rng_base = np.random.default_rng(1).integers(0, 5, n).astype(np.int32)
    tb = dev(rng_base)
    L.check(L.lib.sdb_gen_child(p, ctypes.c_void_p(tb.data_ptr()), n, 3, 4, 0, 40, 0, 39, s))
    assert (host(out) == O.gen_child(base, 3, 4, 40, 0, 39)).all()
    dk, _, _ = O.gen_dates(2556)
    td = dev(dk)
    L.check(L.lib.sdb_gen_lookup(p, n, 9, 10, 0, ctypes.c_void_p(td.data_ptr()), 2556, s))
    assert (host(out) == O.gen_lookup(n, 9, 10, dk)).all()
    L.check(L.lib.sdb_gen_rowid(p, n, 10, s))
    assert (host(out) == np.arange(11, n + 11, dtype=np.int32)).all()


def test_error_paths(ops):
    t = torch.zeros(4, dtype=torch.int32, device="cuda")
    f = torch.zeros(4, dtype=torch.uint8, device="cuda")
    with pytest.raises(ValueError):
        ops.select_lit(t, f, f, 9, 0, 1)          # comparison out of range
    with pytest.raises(ValueError):
        ops.binop(t, t, t, f, 7)                  # arithmetic op out of range
    with pytest.raises(ValueError):
        ops.probe_semi(t, f, f, f, 0, 100)        # table smaller than its range


def test_compress_sync_scatters_into_host_columns(ops):
    """Segment::compress_sync (models/models.hpp:990-1068): the flagged rows
    of the device columns land in the same rows of the host columns, the
    other host rows stay as they were."""
    rng = np.random.default_rng(17)
    for n, p, k in ((1_000_003, 0.3, 3), (4097, 0.0, 2), (65_536, 1.0, 1), (10_000, 0.5, 0), (0, 0.5, 2)):
        print("compress_sync n=%d p=%.1f cols=%d" % (n, p, k), flush=True)
        cols = [rng.integers(-2**31, 2**31, n).astype(np.int32) for _ in range(k)]
        fl = (rng.random(n) < p).astype(np.uint8)
        host = [np.full(n, -12345, np.int32) for _ in range(k)]
        cnt = ops.compress_sync([dev(c) for c in cols], dev(fl), host, n=n)
        assert cnt == int(fl.sum())
        sel = fl.astype(bool)
        for c, h in zip(cols, host):
            assert (h[sel] == c[sel]).all() and (h[~sel] == -12345).all()


@pytest.mark.parametrize("case", ["u16", "wide_values", "odd_flags", "misaligned"])
def test_probe_full_large_probe_small_table(ops, case):
    """Probes over at least 4M rows and 16 rows per table slot gather from a
    16-bit copy of the pairs built on the device; values spanning 2^16 or
    more fall back to the pairs.  Flags other than 0/1 in the table never
    match (the reference tests == 1).  All against the oracle."""
    rng = np.random.default_rng(23)
    n, mn, mx = 4_500_007, 1000, 200_999
    length = mx - mn + 1
    keys = rng.integers(mn - 500, mx + 500, n).astype(np.int32)       # some out of range
    pfl = (rng.random(n) < 0.7).astype(np.uint8)
    ht2 = np.zeros(2 * length, np.int32)
    ht2[0::2] = (rng.random(length) < 0.6).astype(np.int32)
    lo_v = -40_000
    ht2[1::2] = rng.integers(lo_v, lo_v + 60_000, length).astype(np.int32)
    if case == "wide_values":
        ht2[1::2][::1000] = 10**9                                      # range >= 2^16: pairs path
    if case == "odd_flags":
        ht2[0::2][::7] = 2                                             # neither 0 nor 1: no match
    off = 1 if case == "misaligned" else 0
    print("probe_full large case", case, flush=True)
    tp_buf, tp = padded(keys, off)
    tf_buf, tf = padded(pfl, off)
    out = torch.zeros(n + off, dtype=torch.uint8, device="cuda")[off:]
    vo = torch.full((n,), -77, dtype=torch.int32, device="cuda")
    ops.probe_full(tp, vo, tf, out, dev(ht2), mn, mx)
    want_v = np.full(n, -77, np.int32)
    want = np.zeros(n, np.uint8)
    O.probe_full(keys, want_v, pfl, want, ht2, mn, mx)
    assert (host(out) == want).all()
    assert (host(vo) == want_v).all()
