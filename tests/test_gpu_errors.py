"""Error paths of the batched and served forms, through the C ABI.

Each refused call must return a nonzero status with its message, and must
leave the object it was made on working: a refused configure keeps the
previous configuration, a refused batch run keeps the pointers the batch
last ran with, and a server that refused a request still serves the next.
No refused call reaches a kernel: every check here fires on the host.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Offset:
    """A device pointer moved by `off` bytes inside a buffer that holds it."""

    def __init__(self, buf, off):
        self.ptr = buf.ptr.value + off


def _products(O, points, seed):
    prods, ops = [], []
    for i, (N, W) in enumerate(points):
        q = O.spec_q(N, W)
        prods.append((N, W, q))
        ops.append(O.operands("uniform", N, W, q, seed + i))
    return prods, ops


def _expect(O, prods, ops, dc):
    for i, (N, W, q) in enumerate(prods):
        got = dc[i].download((N, O.limbs_of(W)))
        assert np.array_equal(got, O.OraclePlan(N, W, q).multiply(*ops[i])), "product %d" % i


def test_batch_refused_configure_keeps_config(O, pm):
    points = [(1024, 256), (256, 1747)]
    prods, ops = _products(O, points, 900)
    B = pm.Batch(prods)
    da = [pm.DeviceBuffer.from_array(a) for a, _ in ops]
    db = [pm.DeviceBuffer.from_array(b) for _, b in ops]
    dc = [pm.DeviceBuffer(N * O.limbs_of(W) * 4) for N, W in points]
    B.configure(col_tile=16)
    bad = [dict(col_tile=7), dict(col_tile=64), dict(branches=0), dict(branches=99), dict(blocks_per_cu=9),
           dict(blocks_per_cu=-1), dict(ceil_split=2), dict(group_mb=-1)]
    for kw in bad:
        print("configure", kw, flush=True)
        with pytest.raises(pm.PolymultError, match="col_tile must be 8, 16 or 32|batch configuration out of range"):
            B.configure(**kw)
    B.run(da, db, dc)
    B.sync()
    _expect(O, prods, ops, dc)
    B.close()


def test_batch_bad_product_index(pm):
    B = pm.Batch([(64, 32, (1 << 32) - 5)])
    info = pm.PmInfo()
    rc = pm.load().pm_batch_product_info(B._h, 1, ctypes.byref(info))
    assert rc != 0
    assert "bad product index" in pm.load().pm_batch_last_error().decode()
    B.close()


def test_batch_empty(pm):
    h = ctypes.c_void_p()
    z = np.zeros(1, np.uint32)
    p = pm._ptr(z)
    rc = pm.load().pm_batch_create(0, p, p, p, ctypes.byref(h))
    assert rc != 0 and not h.value
    assert "empty batch" in pm.load().pm_batch_last_error().decode()


def test_batch_refused_run_keeps_pointers(O, pm):
    """A run refused for its second product's pointer must not keep the
    first product's new pointer either: the next run with that pointer has
    to read the new operand, not the one the device table still holds."""
    points = [(512, 300), (2048, 100)]
    prods, ops = _products(O, points, 910)
    B = pm.Batch(prods)
    da = [pm.DeviceBuffer.from_array(a) for a, _ in ops]
    db = [pm.DeviceBuffer.from_array(b) for _, b in ops]
    dc = [pm.DeviceBuffer(N * O.limbs_of(W) * 4) for N, W in points]
    B.run(da, db, dc)
    B.sync()
    _expect(O, prods, ops, dc)

    N0, W0, q0 = prods[0]
    a0x, _ = O.operands("uniform", N0, W0, q0, 977)
    da0x = pm.DeviceBuffer.from_array(a0x)
    # a buffer 256 bytes longer than the operand, so the moved pointer stays inside it
    N1, W1, _ = prods[1]
    roomy = pm.DeviceBuffer(N1 * O.limbs_of(W1) * 4 + 256)
    for moved in ([da0x, _Offset(roomy, 4)], [_Offset(roomy, 8), da[1]]):
        with pytest.raises(pm.PolymultError, match="16-byte aligned"):
            B.run(moved, db, dc)
    with pytest.raises(pm.PolymultError, match="null operand pointer"):
        B.run([da0x, da[1]], db, [dc[0], _Offset(roomy, -roomy.ptr.value)])
    for d in dc:
        d.upload(np.zeros(d.nbytes // 4, np.uint32))
    B.run([da0x, da[1]], db, dc)
    B.sync()
    _expect(O, prods, [(a0x, ops[0][1]), ops[1]], dc)
    B.close()


def test_server_refused_configs(O, pm):
    cat = [(256, 300, O.spec_q(256, 300))]
    for kw, msg in ((dict(max_inflight=9), "max_inflight above 8"), (dict(col_tile=7), "col_tile must be 8, 16 or 32"),
                    (dict(col_tile=12), "col_tile must be 8, 16 or 32")):
        print("server", kw, flush=True)
        with pytest.raises(pm.PolymultError, match=msg):
            pm.Server(cat, **kw)


def test_server_refused_requests_then_serves(O, pm):
    N, W = 256, 300
    q = O.spec_q(N, W)
    L = O.limbs_of(W)
    a, b = (np.ascontiguousarray(x, dtype=np.uint32) for x in O.operands("uniform", N, W, q, 920))
    with pm.Server([(N, W, q)]) as S:
        lib = pm.load()
        c = np.empty((N, L), np.uint32)
        t = ctypes.c_uint64()
        vp = pm._ptr
        rc = lib.pm_server_submit(S._h, 1, vp(a), vp(b), vp(c), ctypes.byref(t))
        assert rc < 0 and "point index out of range" in lib.pm_server_last_error().decode()
        rc = lib.pm_server_submit(S._h, 0, vp(a), None, vp(c), ctypes.byref(t))
        assert rc < 0 and "null operand pointer" in lib.pm_server_last_error().decode()
        with pytest.raises(pm.PolymultError, match="unknown ticket"):
            S.wait(123456789, 1.0)
        for args, msg in (((0.0, 0.1, [1.0]), "rate and seconds must be positive"),
                          ((100.0, 0.0, [1.0]), "rate and seconds must be positive"),
                          ((100.0, 0.1, [-1.0]), "mix weights must be non-negative"),
                          ((100.0, 0.1, [0.0]), "mix weights sum to zero")):
            with pytest.raises(pm.PolymultError, match=msg):
                S.drive(*args)
        for depth in (0, 4097):
            with pytest.raises(pm.PolymultError, match=r"depth must be in 1\.\.4096"):
                S.saturate(depth, 0.1, [1.0])
        with pytest.raises(pm.PolymultError, match="mix weights must be non-negative"):
            S.saturate(4, 0.1, [float("nan")])
        got = S.wait(S.submit(0, a, b))
        assert got is not None
        assert np.array_equal(got[0], O.OraclePlan(N, W, q).multiply(a, b))
        st = S.stats()
        assert st["completed"] == 1
