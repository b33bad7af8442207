"""pk_count_noalloc (the allocation-free count) against pk_count, through the
C ABI: bit-exact tables on two calls in a row, the second with a smaller
input (the shared scratch keeps the first call's contents), at k = 32 and
k = 64; inputs beyond the scratch and tables beyond the output capacity fail
with PK_ENOMEM; the context's own device memory does not grow."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from golden_util import read_fasta

pytestmark = pytest.mark.gpu


def _p(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctxs():
    from pakman_plus_amd import api
    a, b = api.Context(0), api.Context(0)
    yield a, b
    a.close()
    b.close()


def _dev(fa):
    import torch
    return torch.from_numpy(np.frombuffer(bytes(fa), dtype=np.uint8).copy()).to("cuda:0")


def _noalloc(ctx, dev, k, capacity):
    import torch
    words = 1 if k <= 32 else 2
    keys = torch.zeros(max(capacity, 1) * words, dtype=torch.int64, device="cuda:0")
    counts = torch.zeros(max(capacity, 1), dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()
    n = ctx.count_noalloc(dev.data_ptr(), dev.numel(), k, keys.data_ptr(), counts.data_ptr(), capacity)
    torch.cuda.synchronize()
    kh = keys.cpu().numpy().view(np.uint64)[: n * words].reshape(n, words)
    return kh, counts.cpu().numpy()[:n]


def _inputs():
    from pakman_plus_amd import gen
    big = bytes(gen.make_fasta(gen.config_small(scale=0.1)))  # 100 kbp, 30x: 3.3 MB
    mid = read_fasta("syn_mid.fa.gz")
    tiny = read_fasta("syn_tiny.fa.gz")
    return big, mid, tiny


@pytest.mark.parametrize("k", [32, 64])
def test_noalloc_matches_count_twice(ctxs, k):
    ctx, ref = ctxs
    big, mid, tiny = _inputs()
    first, second = (big, mid) if k == 32 else (mid, tiny)
    held = ctx.device_bytes
    for i, fa in enumerate((first, second)):
        _p(f"case: noalloc k={k} call {i + 1} ({len(fa)} bytes)")
        dev = _dev(fa)
        ref.count(fa, k)
        want_k, want_c = ref.table_host()
        got_k, got_c = _noalloc(ctx, dev, k, want_c.shape[0] + 16)
        assert got_k.shape == want_k.shape
        assert np.array_equal(got_k, want_k) and np.array_equal(got_c, want_c)
    assert ctx.device_bytes == held


def test_noalloc_limits(ctxs):
    from pakman_plus_amd import api, gen
    ctx, ref = ctxs
    held = ctx.device_bytes
    fa = bytes(gen.make_fasta(gen.config_small()))  # 33 MB: beyond the scratch
    _p(f"case: noalloc beyond the scratch ({len(fa)} bytes)")
    with pytest.raises(api.PakHipError) as e:
        _noalloc(ctx, _dev(fa), 32, 1 << 20)
    assert e.value.code == api.PK_ENOMEM and "scratch" in str(e.value)
    _p("case: noalloc with too small an output")
    small = read_fasta("syn_tiny.fa.gz")
    with pytest.raises(api.PakHipError) as e:
        _noalloc(ctx, _dev(small), 32, 10)
    assert e.value.code == api.PK_ENOMEM and "output" in str(e.value)
    # the context still counts after the failures
    ref.count(small, 32)
    want_k, want_c = ref.table_host()
    got_k, got_c = _noalloc(ctx, _dev(small), 32, want_c.shape[0])
    assert np.array_equal(got_k, want_k) and np.array_equal(got_c, want_c)
    assert ctx.device_bytes == held
