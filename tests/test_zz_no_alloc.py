"""sdb_compact_noalloc, the compaction entry point that allocates nothing
(its tile counts live in a fixed device array of the library), against
sdb_compact through the C ABI: two calls in a row, the second smaller, plus
the size limit and the empty input."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 28  # SDB_COMPACT_NOALLOC_MAX_ROWS


def _compact(fn, flags, n):
    import torch
    from sycldb_amd import _lib as L
    ids = torch.full((max(n, 1),), -1, dtype=torch.int32, device="cuda:0")
    cnt = torch.full((1,), 0xFFFF, dtype=torch.int32, device="cuda:0")
    L.check(fn(ctypes.c_void_p(flags.data_ptr()), n, ctypes.c_void_p(ids.data_ptr()),
               ctypes.c_void_p(cnt.data_ptr()), None))
    torch.cuda.synchronize()
    c = int(cnt.item())
    return c, ids[:c].cpu().numpy()


def test_noalloc_matches_compact_twice_smaller_second(dev):
    import torch
    from sycldb_amd import _lib as L
    rng = np.random.default_rng(11)
    for n, p in ((10_000_019, 0.37), (777_777, 0.02)):
        print("noalloc compaction n=%d p=%.2f" % (n, p), flush=True)
        host = (rng.random(n) < p).astype(np.uint8)
        flags = torch.from_numpy(host).to(dev)
        want_c, want = _compact(L.lib.sdb_compact, flags, n)
        got_c, got = _compact(L.lib.sdb_compact_noalloc, flags, n)
        assert got_c == want_c == int(host.sum())
        assert (got == want).all()
        assert (got == np.flatnonzero(host)).all()


def test_noalloc_empty_and_limit(dev):
    import torch
    from sycldb_amd import _lib as L
    flags = torch.ones(4, dtype=torch.uint8, device=dev)
    c, _ = _compact(L.lib.sdb_compact_noalloc, flags, 0)
    assert c == 0
    ids = torch.zeros(4, dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int32, device=dev)
    # rejected on the host before any launch (the buffers are never touched)
    rc = L.lib.sdb_compact_noalloc(ctypes.c_void_p(flags.data_ptr()), MAX_ROWS + 1, ctypes.c_void_p(ids.data_ptr()),
                                   ctypes.c_void_p(cnt.data_ptr()), None)
    assert rc == 1  # SDB_ERR_INVALID_ARG
    assert b"2^28" in L.lib.sdb_last_error()


def test_noalloc_at_the_limit(dev):
    import torch
    from sycldb_amd import _lib as L
    n = MAX_ROWS
    print("noalloc compaction at the limit n=%d" % n, flush=True)
    flags = torch.zeros(n, dtype=torch.uint8, device=dev)
    flags[::1000] = 1
    flags[n - 1] = 1
    want_c, want = _compact(L.lib.sdb_compact, flags, n)
    got_c, got = _compact(L.lib.sdb_compact_noalloc, flags, n)
    assert got_c == want_c == (n + 999) // 1000 + 1
    assert (got == want).all()
