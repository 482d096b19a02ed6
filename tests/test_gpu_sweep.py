"""Seeded random sweep over the points the point check accepts (solve.cu:873-913):
N a power of two, W from 2 to 15000 bits (both RNS epilogues, the split CRT
and its global-memory finish), and q = 2^W - delta for an odd delta anywhere
from 1 to its limit (below 2^63 and below 2^(W-1)), a third of them at the
limit. Every product goes through the host and the device path and must equal
the oracle's, bit for bit.

The points are drawn from a fixed seed, so a failure names a point that
reproduces on the CPU with the oracle alone.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KINDS = ["uniform", "max", "mixed", "sparse", "small16", "zero"]


def _points(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        # W spread over 2 .. 15000 on a log scale, so narrow points are not rare;
        # N up to 4096, and up to 256 above 4064 bits (the oracle's time)
        W = max(2, min(15000, int(round(2 ** rng.uniform(1, np.log2(15000))))))
        logN = int(rng.integers(0, 13 if W <= 4064 else 9))
        dmax = min(63, W - 1)  # delta below 2^dmax
        if rng.integers(0, 3) == 0:
            delta = (1 << dmax) - 1
        else:
            dbits = int(rng.integers(1, dmax + 1))
            delta = int(rng.integers(0, 1 << (dbits - 1))) | (1 << (dbits - 1)) | 1 if dbits > 1 else 1
        out.append((1 << logN, W, delta, KINDS[int(rng.integers(0, len(KINDS)))]))
    return out


SWEEP = _points(48, 20261204)


@pytest.mark.parametrize("N,W,delta,kind", SWEEP)
def test_random_point(O, pm, N, W, delta, kind):
    q = (1 << W) - delta
    L = O.limbs_of(W)
    assert q % 2 == 1 and q.bit_length() > 32 * (L - 1) and (1 << W) - q < (1 << 63)
    P = O.OraclePlan(N, W, q)
    a, b = O.operands(kind, N, W, q, N + W + delta % 997)
    want = P.multiply(a, b)
    with pm.Multiplier(N, W, O.q_limbs(q, L)) as m:
        print("sweep N=%d W=%d K=%d delta bits %d %s" % (N, W, m.K, delta.bit_length(), kind), flush=True)
        assert m.K == P.K
        got = m.run(a, b)
        assert np.array_equal(got, want), "host path: %d of %d coefficients differ" % (
            int(np.count_nonzero((got != want).any(axis=1))), N)
        da, db = pm.DeviceBuffer.from_array(a), pm.DeviceBuffer.from_array(b)
        dc = pm.DeviceBuffer(N * L * 4)
        m.run_device(da, db, dc)
        m.sync()
        assert np.array_equal(dc.download((N, L)), want), "device path"
