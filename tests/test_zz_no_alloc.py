"""The allocation-free G1 MSM entry point (zk_msm_g1_device_noalloc) through
the C ABI: two calls in a row, the second smaller, each identical to
zk_msm_g1_device on the same device buffers and to the discrete-log closed
form; a workspace beyond the fixed scratch is refused before any launch."""
import numpy as np
import pytest

import oracle as O
import zyklop_amd as zk
from tests import pyref

pytestmark = pytest.mark.gpu
RR = pyref.R


def case(rng, n):
    k0, d = (int.from_bytes(rng.bytes(32), "little") % RR for _ in range(2))
    pts = O.g1_progression(k0, d, n)
    sc = np.zeros((n, 4), np.uint64)
    sc[:, :3] = rng.integers(0, 2**64, size=(n, 3), dtype=np.uint64)
    sc[:, 3] = rng.integers(0, 0x30644E72E131A029, size=n, dtype=np.uint64)
    ones = np.zeros((n, 4), np.uint64)
    ones[:, 0] = 1
    idx = np.zeros((n, 4), np.uint64)
    idx[:, 0] = np.arange(n, dtype=np.uint64)
    tot = (k0 * O.fr_dot(sc, ones) + d * O.fr_dot(sc, idx)) % RR
    p, inf = O.g1_mul(O.g1_gen(), tot)
    exp = O.field_op(O.FP, O.TO_MONT, p.reshape(2, 4)).reshape(8)
    return pts, sc, exp


def test_noalloc_msm_matches_allocating_path():
    ctx = zk.Context(0)
    rng = np.random.default_rng(99)
    bufs = []
    try:
        for n, window in ((60000, 0), (3000, 0), (3000, 11)):
            pts, sc, exp = case(rng, n)
            dp, ds = ctx.dev_alloc(pts.nbytes), ctx.dev_alloc(sc.nbytes)
            bufs += [dp, ds]
            ctx.upload(dp, pts)
            ctx.upload(ds, sc)
            got, inf = ctx.msm_g1_device_noalloc(dp, ds, n, window)
            ref, rinf = ctx.msm_g1_device(dp, ds, n, window)
            assert inf == rinf and np.array_equal(got, ref), (n, window)
            assert not inf and np.array_equal(got, exp), (n, window)
        # n = 0 is the point at infinity
        got, inf = ctx.msm_g1_device_noalloc(bufs[0], bufs[1], 0, 0)
        assert inf and not got.any()
        # 2^18 points need more than the fixed 64 MiB: refused on the host
        big = ctx.dev_alloc((1 << 18) * 64)
        bigs = ctx.dev_alloc((1 << 18) * 32)
        bufs += [big, bigs]
        with pytest.raises(zk.ZkError) as e:
            ctx.msm_g1_device_noalloc(big, bigs, 1 << 18, 0)
        assert e.value.code == 1 and "fixed scratch" in str(e.value)
    finally:
        for p in bufs:
            ctx.dev_free(p)
        ctx.close()
