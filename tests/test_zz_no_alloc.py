"""The allocation-free device call (pm_run_device_scratch) against the
ordinary one (pm_run_device), through the C ABI.

The allocation-free call keeps its residues in one fixed array per GPU that
every call shares, so two calls in a row, the second on a smaller point
(fewer residues, the first call's contents still in the array), must each
give exactly what pm_run_device gives on the same operands.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(O, pm, m, N, W, kind, seed):
    q = O.spec_q(N, W)
    a, b = O.operands(kind, N, W, q, seed)
    L = O.limbs_of(W)
    da, db = pm.DeviceBuffer.from_array(a), pm.DeviceBuffer.from_array(b)
    want_buf, got_buf = pm.DeviceBuffer(N * L * 4), pm.DeviceBuffer(N * L * 4)
    m.run_device(da, db, want_buf)
    m.sync()
    return da, db, want_buf, got_buf


@pytest.mark.parametrize("on_stream", [False, True])
def test_scratch_call_matches_device_call(O, pm, on_stream):
    """65536/1747 then the smaller 4096/256, each allocation-free call against
    pm_run_device on the same operands (and the small one against the oracle)."""
    stream = pm.Stream() if on_stream else None
    handle = stream.handle if stream else None
    points = [(65536, 1747, "mixed", 31), (4096, 256, "uniform", 32)]
    for N, W, kind, seed in points:
        q = O.spec_q(N, W)
        L = O.limbs_of(W)
        print("scratch call N=%d W=%d stream=%s" % (N, W, on_stream), flush=True)
        with pm.Multiplier(N, W, O.q_limbs(q, L)) as m:
            da, db, want_buf, got_buf = _pair(O, pm, m, N, W, kind, seed)
            m.run_device_scratch(da, db, got_buf, handle)
            if stream:
                stream.sync()
            m.sync()
            want = want_buf.download((N, L))
            got = got_buf.download((N, L))
            assert np.array_equal(got, want), "N=%d W=%d" % (N, W)
            if N <= 4096:
                a, b = O.operands(kind, N, W, q, seed)
                assert np.array_equal(got, O.OraclePlan(N, W, q).multiply(a, b))
    if stream:
        stream.close()


def test_scratch_call_refuses_large_point(O, pm):
    """A point whose residues exceed the array is refused with a message, and
    the state's other calls keep working."""
    N, W = 65536, 4000
    q = O.spec_q(N, W)
    L = O.limbs_of(W)
    with pm.Multiplier(N, W, O.q_limbs(q, L)) as m:
        assert 2 * m.K * N > 24 << 20
        a, b = O.operands("uniform", N, W, q, 5)
        da, db = pm.DeviceBuffer.from_array(a), pm.DeviceBuffer.from_array(b)
        dc = pm.DeviceBuffer(N * L * 4)
        with pytest.raises(pm.PolymultError, match="too large for the allocation-free call"):
            m.run_device_scratch(da, db, dc)
        m.run_device(da, db, dc)
        m.sync()
        assert np.array_equal(dc.download((N, L)), m.run(a, b))
