"""Prime shards: the multi-GPU split of the multiply, run on one GPU.

A shard runs RNS, the NTT passes and the linear part of the CRT for its
primes only; the combine sums the shards' slot partials and finishes the CRT.
The result must be bit-identical to the unsharded multiply for every shard
count (and so to the oracle).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,W", [(65536, 1747), (4096, 256), (2048, 868), (1024, 4000), (64, 32), (1, 8)])
def test_sharded_equals_plain(O, pm, N, W):
    q = O.spec_q(N, W)
    L = O.limbs_of(W)
    a, b = O.operands("mixed", N, W, q, 77)
    with pm.Multiplier(N, W, O.q_limbs(q, L)) as m:
        da, db = pm.DeviceBuffer.from_array(a), pm.DeviceBuffer.from_array(b)
        dc = pm.DeviceBuffer(N * L * 4)
        m.run_device(da, db, dc)
        m.sync()
        want = dc.download((N, L))
        if N <= 4096:
            assert np.array_equal(want, O.OraclePlan(N, W, q).multiply(a, b))
        groups = (m.K + 7) // 8
        for shards in sorted({1, 2, 3, groups}):
            if shards > groups:
                continue
            bounds = m.shard_bounds(shards)
            print("N=%d W=%d K=%d shards=%d bounds=%s" % (N, W, m.K, shards, bounds), flush=True)
            assert bounds[0] == 0 and bounds[-1] == m.K
            assert all(x % 8 == 0 for x in bounds[:-1]) and all(x < y for x, y in zip(bounds, bounds[1:]))
            dc.upload(np.zeros(N * L, np.uint32))
            m.run_device_sharded(da, db, dc, shards)
            m.sync()
            assert np.array_equal(dc.download((N, L)), want), "shards=%d" % shards
        with pytest.raises(pm.PolymultError, match="shards must be between 1 and"):
            m.shard_bounds(groups + 1)


def test_shard_timing_reports(O, pm):
    N, W = 65536, 1747
    q = O.spec_q(N, W)
    L = O.limbs_of(W)
    a, b = O.operands("uniform", N, W, q, 5)
    with pm.Multiplier(N, W, O.q_limbs(q, L)) as m:
        da, db = pm.DeviceBuffer.from_array(a), pm.DeviceBuffer.from_array(b)
        dc = pm.DeviceBuffer(N * L * 4)
        t = m.time_shard(da, db, dc, 8, 3)
        print("shard timing", t, flush=True)
        assert all(v > 0 for v in t.values())
        # the timed runs leave the sharded result in dc
        m.run_device(da, db, dc)
        m.sync()
        want = dc.download((N, L))
        m.run_device_sharded(da, db, dc, 8)
        m.sync()
        assert np.array_equal(dc.download((N, L)), want)
