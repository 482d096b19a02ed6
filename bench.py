"""Benchmark of the negacyclic multiply c = a b mod (X^N + 1, q) on one MI355X.

Workload: the plain single product at N=65536, W=1747 (L=55 limbs, K=114 RNS
primes, q = 2^1747 - 9699327), operands mixed per coefficient (1/2 uniform in
[0, q), 1/4 q-1, 1/4 below 2^32), seeded. One step = one whole multiply with
a and b already resident in HBM (RNS, three NTT passes, CRT), launched
through the C ABI (pm_run_device). `value` is multiplies per second over the
timed steps. The PCIe-inclusive rate of the host-buffer call (pm_run, the
fherma_run replacement) is reported beside it, never as `value`.

    python bench.py [--gpus 1] [--steps 2000] [--warmup 200] [--no-cpu] [--serve-seconds 30] [--no-serve]
                    [--legs plain,points,shard,e2e,batch,serve,cpu] [--copy-probe BYTES] [--settle-ms 1000]

Before the timed steps the multiply runs for at least --settle-ms (1 s) and
--warmup steps, so the clocks have settled whatever --warmup is; the timed
steps are exactly --steps. Counter runs pass --settle-ms 0 to keep the
profiler's per-dispatch output small.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "polymult-challenge_amd"))
sys.path.insert(0, str(ROOT))

N, W = 65536, 1747
SEED = 20260
HBM_PEAK_GBS = 8000.0
PCIE_PEAK_GBS = 64.0  # PCIe Gen5 x16, per direction (vendor figure)
# 32-bit integer VALU: 1024 SIMDs x 16 lanes x 2.4 GHz lane-operations per second;
# a Shoup butterfly (modmul, add, sub, all canonical) is 12 of them. This is the
# nominal figure; the measured one (pm_probe_valu) prices the butterfly by the
# measured issue rates of its 9 add-class instructions, 1 v_mul_hi_u32 and 2
# v_mul_lo_u32.
VALU_INT32_OPS = 1024 * 16 * 2.4e9
BUTTERFLY_VALU = 12
PEAK_MODMUL_PER_S = VALU_INT32_OPS / BUTTERFLY_VALU
# Time-based warm-up before the timed steps, so the clocks have settled
# whatever --warmup is: after 0.3 s the per-kernel times still read 3-4%
# above those of a run that had done 0.5 s more work.
SETTLE_S = 1.0
KERNEL_TIMING_LAUNCHES = 100  # launches per kernel in the HIP-event kernel times
METRIC = "mults/s at N=65536/W=1747 (plain vs batched), %HBM/PCIe peak, serve p99 ms"
KERNELS = ["rns", "col_fwd", "row_fused", "col_inv", "crt"]


def kernel_bytes(K: int, L: int) -> dict:
    """Algorithmic HBM bytes of each kernel per multiply: the operands,
    residues and result it must read and write once (no twiddles, no
    re-reads)."""
    r = K * N * 4  # one residue array
    c = N * L * 4  # one coefficient array
    return {"rns": 2 * c + 2 * r, "col_fwd": 4 * r, "row_fused": 3 * r, "col_inv": 2 * r, "crt": r + c}


def modmuls(K: int) -> dict:
    """Modular multiplications per kernel per multiply (butterflies count one)."""
    logn = N.bit_length() - 1
    half = N // 2
    w1 = logn // 2
    w2 = logn - w1
    return {"rns": 0, "col_fwd": 2 * K * half * w1, "row_fused": K * (3 * half * w2 + N),
            "col_inv": K * half * w1, "crt": K * N}


def cpu_baseline(q: int, sample_cases: int) -> dict | None:
    """The reference's own solve.cu on its CPU emulator (oracle/_ref), pinned
    to one core, on `sample_cases` products of this workload."""
    from oracle import oracle as O

    exe = O.ref_binary()
    if exe is None:
        return None
    with tempfile.TemporaryDirectory() as d:
        cases = [O.operands("mixed", N, W, q, SEED + 1 + i) for i in range(sample_cases)]
        O.write_point(d, N, W, q, cases)
        core = sorted(os.sched_getaffinity(0))[0]
        t0 = time.perf_counter()
        subprocess.run([str(exe), d], check=True, capture_output=True,
                       preexec_fn=lambda: os.sched_setaffinity(0, {core}))
        wall = time.perf_counter() - t0
        res = O.read_results(d)
        secs = [c["seconds"] for c in res["cases"] if c["status"] == "ok"]
        # spot-check one output against the oracle
        L = O.limbs_of(W)
        ok = bool(np.array_equal(O.read_output(d, 0, N, L), O.OraclePlan(N, W, q).multiply(*cases[0])))
    total = sum(secs)
    return {"value": len(secs) / total if total else None, "unit": "mults/s", "cores": 1, "kind": "reference",
            "sample": "%d products at N=65536/W=1747 (mixed operands), fherma_run time from the reference's "
                      "results.json; solve.cu on pm_emu.h built by oracle/Makefile, pinned to one core "
                      "(wall %.1f s incl. init)" % (len(secs), wall),
            "per_call_s": secs, "output_matches_oracle": ok}


def batched(pmlib, O, count: int, points, seed: int, check: int, iters: int, tune: bool = True) -> dict:
    """One hipGraph launch over `points`; mults/s = products / launch time."""
    prods, ops = [], []
    for i, (n, w) in enumerate(points):
        q = O.spec_q(n, w)
        prods.append((n, w, q))
        ops.append(O.operands("uniform", n, w, q, seed + i))
    B = pmlib.Batch(prods)
    da = [pmlib.DeviceBuffer.from_array(a) for a, _ in ops]
    db = [pmlib.DeviceBuffer.from_array(b) for _, b in ops]
    dc = [pmlib.DeviceBuffer(n * O.limbs_of(w) * 4) for n, w in points]
    tuned = B.autotune(da, db, dc, 3) if tune else "default configuration"
    B.run(da, db, dc)
    B.sync()
    from concurrent.futures import ThreadPoolExecutor

    def exact_one(i):
        n, w, q = prods[i]
        return bool(np.array_equal(dc[i].download((n, O.limbs_of(w))), O.OraclePlan(n, w, q).multiply(*ops[i])))

    with ThreadPoolExecutor(max_workers=8) as ex:  # the oracle's ctypes calls release the GIL
        exact = all(ex.map(exact_one, range(min(check, count))))
    B.time(da, db, dc, 10)  # warm-up: the clocks settle after the oracle check
    ms = B.time(da, db, dc, iters)
    B.close()
    # the plain per-call path on the same products: one device-resident
    # multiply per product, back to back (HIP events on the state's stream)
    plain_ms = {}
    for (n, w, q), a_, b_, c_ in zip(prods, da, db, dc):
        if (n, w) not in plain_ms:
            with pmlib.Multiplier(n, w, O.q_limbs(q, O.limbs_of(w))) as m:
                # by time, whatever the size: 30 ms of warm-up (clocks), then
                # at least 100 ms and 20 calls timed
                def calls(k):
                    for _ in range(k):
                        m.run_device(a_, b_, c_)
                    m.sync()
                t = time.perf_counter()
                while time.perf_counter() - t < 0.03:
                    calls(10)
                k, t = 0, time.perf_counter()
                while k < 20 or time.perf_counter() - t < 0.1:
                    calls(10)
                    k += 10
                plain_ms[(n, w)] = 1e3 * (time.perf_counter() - t) / k
    plain_total = sum(plain_ms[(n, w)] for n, w, _ in prods)
    io = sum(12 * n * O.limbs_of(w) for n, w in points)
    return {"products": count, "ms_per_launch": round(ms, 4), "mults_per_s": round(count / (ms * 1e-3), 1),
            "plain_per_call_ms": round(plain_total, 4),
            "plain_per_call_mults_per_s": round(count / (plain_total * 1e-3), 1),
            "speedup_vs_plain": round(plain_total / ms, 3),
            "io_bytes": io, "bit_exact_checked": min(check, count), "bit_exact": exact, "autotune": tuned}


def point_line(pmlib, O, n: int, w: int, kinds, seed: int, e2e_calls: int = 20) -> dict:
    """One plain point: for each operand distribution, the device-resident
    result and the host-buffer (pm_run) result checked against the oracle,
    then the device-resident rate (at least 0.2 s of back-to-back calls after
    SETTLE_S of warm-up) and the median host-buffer call on the first."""
    q = O.spec_q(n, w)
    L = O.limbs_of(w)
    P = O.OraclePlan(n, w, q)
    out = {"N": n, "W": w, "L": L, "operands": list(kinds), "bit_exact": {}}
    with pmlib.Multiplier(n, w, O.q_limbs(q, L)) as m:
        out["K"] = m.K
        dev = None
        for i, kind in enumerate(kinds):
            a, b = O.operands(kind, n, w, q, seed + i)
            da, db = pmlib.DeviceBuffer.from_array(a), pmlib.DeviceBuffer.from_array(b)
            dc = pmlib.DeviceBuffer(n * L * 4)
            m.run_device(da, db, dc)
            m.sync()
            want = P.multiply(a, b)
            out["bit_exact"][kind] = bool(np.array_equal(dc.download((n, L)), want)) and \
                bool(np.array_equal(m.run(a, b), want))
            if dev is None:
                dev = (a, b, da, db, dc)
        a, b, da, db, dc = dev

        def calls(k):
            for _ in range(k):
                m.run_device(da, db, dc)
            m.sync()
        t = time.perf_counter()
        while time.perf_counter() - t < SETTLE_S:
            calls(20)
        k, t = 0, time.perf_counter()
        while k < 100 or time.perf_counter() - t < 0.2:
            calls(20)
            k += 20
        dt = time.perf_counter() - t
        out["device_mults_per_s"] = round(k / dt, 1)
        c_host = np.empty((n, L), np.uint32)
        m.run(a, b, c_host)
        ts = []
        for _ in range(e2e_calls):
            t = time.perf_counter()
            m.run(a, b, c_host)
            ts.append(time.perf_counter() - t)
        med = float(np.median(ts))
        out["end_to_end_median_ms"] = round(1e3 * med, 4)
        out["end_to_end_mults_per_s"] = round(1.0 / med, 1)
        out["pcie_bytes"] = 12 * n * L
    return out


def cpu_point(O, n: int, w: int, cases: int, seed: int) -> dict | None:
    """The reference's solve.cu on its CPU emulator, pinned to one core, at one
    point: mults/s over `cases` uniform products (fherma_run times from its
    results.json), one output checked against the oracle."""
    exe = O.ref_binary()
    if exe is None:
        return None
    q = O.spec_q(n, w)
    with tempfile.TemporaryDirectory() as d:
        ops = [O.operands("uniform", n, w, q, seed + i) for i in range(cases)]
        O.write_point(d, n, w, q, ops)
        core = sorted(os.sched_getaffinity(0))[0]
        subprocess.run([str(exe), d], check=True, capture_output=True,
                       preexec_fn=lambda: os.sched_setaffinity(0, {core}))
        res = O.read_results(d)
        secs = [c["seconds"] for c in res["cases"] if c["status"] == "ok"]
        ok = bool(np.array_equal(O.read_output(d, 0, n, O.limbs_of(w)), O.OraclePlan(n, w, q).multiply(*ops[0])))
    total = sum(secs)
    return {"value": round(len(secs) / total, 2) if total else None, "unit": "mults/s", "cores": 1,
            "kind": "reference", "sample": "%d uniform products at N=%d/W=%d, solve.cu on pm_emu.h, one core"
                                          % (len(secs), n, w),
            "output_matches_oracle": ok}


SERVE_POINTS = [(8192, 438), (32768, 868), (65536, 1747)]
SERVE_MIX = [0.70, 0.25, 0.05]


def plain_mix(pmlib, O) -> dict:
    """The plain per-call path on the serving mix: pm_run (the fherma_run
    replacement) from pageable host buffers, one request at a time; median
    seconds per point. main() runs it early: the first state made right after
    gigabytes of device memory were freed ran its host path up to 40% slower,
    which would flatter the comparison with the server."""
    per_point = {}
    for (n, w) in SERVE_POINTS:
        q = O.spec_q(n, w)
        L = O.limbs_of(w)
        a, b = O.operands("uniform", n, w, q, SEED + 7)
        c = np.empty((n, L), np.uint32)
        with pmlib.Multiplier(n, w, O.q_limbs(q, L)) as m:
            m.run(a, b, c)
            ts = []
            for _ in range(10):
                t = time.perf_counter()
                m.run(a, b, c)
                ts.append(time.perf_counter() - t)
        per_point["%d/%d" % (n, w)] = float(np.median(ts))
    return per_point


def serving(pmlib, O, rates, seconds: float, per_point: dict) -> dict:
    """Open-loop Poisson load on the serving mix through the native driver
    (pm_server_drive); latency from each scheduled arrival to its result in
    the caller's buffer. One server, one run per rate."""
    cat = [(n, w, O.spec_q(n, w)) for n, w in SERVE_POINTS]
    out = {"mix": {"%d/%d" % p: f for p, f in zip(SERVE_POINTS, SERVE_MIX)}, "seconds_per_rate": seconds,
           "bytes_per_request_mean": round(sum(f * 12 * n * O.limbs_of(w) for (n, w), f in zip(SERVE_POINTS, SERVE_MIX))),
           "rates": []}
    def row_of(rep):
        p50, p99 = rep.pop("point_p50_ms"), rep.pop("point_p99_ms")
        row = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in rep.items()}
        row["per_point"] = {"%d/%d" % pt: {"p50_ms": round(p50[i], 3), "p99_ms": round(p99[i], 3)}
                            for i, pt in enumerate(SERVE_POINTS)}
        return row

    with pmlib.Server(cat) as S:
        S.drive(200.0, 0.5, SERVE_MIX, seed=SEED)  # warm-up
        for r in rates:
            out["rates"].append(row_of(S.drive(float(r), seconds, SERVE_MIX, seed=SEED + int(r))))
            print("bench: serve at %d req/s done" % r, file=sys.stderr, flush=True)
        # capacity: closed loop with 64 requests in flight; the achieved rate
        # is the server's throughput on this mix
        cap = row_of(S.saturate(64, min(seconds, 3.0), SERVE_MIX, seed=SEED + 1))
        out["capacity"] = {k: cap[k] for k in ("requests", "achieved_rate", "mean_batch", "p50_ms", "p99_ms",
                                               "failed")}
        out["capacity"]["in_flight"] = 64
    mean_s = sum(f * per_point["%d/%d" % p] for p, f in zip(SERVE_POINTS, SERVE_MIX))
    out["plain_per_call"] = {"median_ms": {k: round(1e3 * v, 3) for k, v in per_point.items()},
                             "mix_req_per_s": round(1.0 / mean_s, 1)}
    out["capacity_vs_plain"] = round(out["capacity"]["achieved_rate"] / (1.0 / mean_s), 3)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 200 warm-up steps (26 ms) let the clocks settle after the oracle check;
    # with 20 the timed steps ran about 10% slower at N=65536/W=1747
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--e2e-calls", type=int, default=20)
    ap.add_argument("--cpu-cases", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=30.0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--legs", default="plain,points,shard,e2e,batch,serve,cpu",
                    help="comma list of parts to run (plain is always run: it gives value)")
    ap.add_argument("--no-autotune", action="store_true", help="batched legs on the default configuration")
    ap.add_argument("--settle-ms", type=float, default=1e3 * SETTLE_S,
                    help="plain leg: run at least this long before the clock (default 1000)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run the plain device copy kernel over BYTES (the counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the multiply runs on a single MI355X (the path would shard by RNS "
              "prime across GPUs, see the design notes); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    legs = set(args.legs.split(","))
    t_start = time.perf_counter()

    def note(what):
        # one progress line per finished part, on stderr (stdout carries the JSON line)
        print("bench: %s done at %.1f s" % (what, time.perf_counter() - t_start), file=sys.stderr, flush=True)

    if args.no_serve:
        legs.discard("serve")
    if args.no_cpu:
        legs.discard("cpu")

    import pmlib
    from oracle import oracle as O

    if args.copy_probe:
        print(json.dumps({"copy_bytes": args.copy_probe, "GBps": round(pmlib.probe_copy(args.copy_probe, 10), 1)}))
        return

    q = O.spec_q(N, W)
    L = O.limbs_of(W)
    a, b = O.operands("mixed", N, W, q, SEED)
    m = pmlib.Multiplier(N, W, O.q_limbs(q, L))
    K = m.K
    da, db = pmlib.DeviceBuffer.from_array(a), pmlib.DeviceBuffer.from_array(b)
    dc = pmlib.DeviceBuffer(N * L * 4)

    # correctness of the measured configuration, once, before timing
    m.run_device(da, db, dc)
    m.sync()
    c_dev = dc.download((N, L))
    correct = bool(np.array_equal(c_dev, O.OraclePlan(N, W, q).multiply(a, b)))

    # settle: at least --settle-ms and --warmup steps before the clock
    t_settle = time.perf_counter()
    done = 0
    while done < args.warmup or time.perf_counter() - t_settle < 1e-3 * args.settle_ms:
        for _ in range(50):
            m.run_device(da, db, dc)
        m.sync()
        done += 50
    settle_ms = 1e3 * (time.perf_counter() - t_settle)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        m.run_device(da, db, dc)
    m.sync()
    dt = time.perf_counter() - t0
    ms_step = 1e3 * dt / args.steps
    value = args.steps / dt

    # per-kernel times from HIP events on the stream the kernels run on
    ms = m.time_device(da, db, dc, KERNEL_TIMING_LAUNCHES)
    kms = dict(zip(KERNELS, ms[1:]))
    kb = kernel_bytes(K, L)
    dom = max(KERNELS, key=lambda k: kms[k])
    ach = kb[dom] / (kms[dom] * 1e-3) / 1e9
    traffic = None
    tf = ROOT / "profiles" / "traffic.json"
    if tf.exists():
        traffic = json.loads(tf.read_text()).get(dom)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": "profiles/traffic.json: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of this kernel, "
                                  "scaled by the copy-kernel calibration (not measured in this run)",
                "algorithmic_bytes": kb[dom], "kernel_ms": round(kms[dom], 4), "timed_launches": KERNEL_TIMING_LAUNCHES}
    mm = modmuls(K)
    # the same kernel against the VALU: butterflies per second over the
    # butterfly rate the measured issue rates allow
    valu = pmlib.probe_valu(256)
    t_bfly = 9 / valu["add_u32"] + 1 / valu["mul_hi_u32"] + 2 / valu["mul_lo_u32"]  # 1e-12 s per lane
    peak_bfly = 1e12 / t_bfly
    roofline_valu = {"bound": "valu", "kernel": dom, "achieved": round(mm[dom] / (kms[dom] * 1e-3) / 1e9, 1),
                     "peak": round(peak_bfly / 1e9, 1), "unit": "G butterflies/s",
                     "frac": round(mm[dom] / (kms[dom] * 1e-3) / peak_bfly, 4),
                     "measured_issue_Tops": valu,
                     "butterfly": "9 add-class + 1 v_mul_hi_u32 + 2 v_mul_lo_u32 lane instructions (Shoup CT)",
                     "nominal_peak_G": round(PEAK_MODMUL_PER_S / 1e9, 1)}
    per_kernel = {k: {"ms": round(kms[k], 4), "GBps": round(kb[k] / (kms[k] * 1e-3) / 1e9, 1),
                      "Gmodmul_per_s": round(mm[k] / (kms[k] * 1e-3) / 1e9, 1),
                      "modmul_frac_of_valu_peak": round(mm[k] / (kms[k] * 1e-3) / PEAK_MODMUL_PER_S, 3),
                      "modmul_frac_of_measured_valu_peak": round(mm[k] / (kms[k] * 1e-3) / peak_bfly, 3)}
                  for k in KERNELS}
    note("plain (%.1f mults/s, bit-exact %s)" % (value, correct))

    # the other plain points: C1 (the CPU-reference comparison point) and C2
    # (four operand sets)
    points = None
    if "points" in legs:
        points = {"C1": point_line(pmlib, O, 4096, 256, ["uniform"], SEED + 11),
                  "C2": point_line(pmlib, O, 32768, 868, ["uniform", "max", "sparse", "small16"], SEED + 21)}
        if "cpu" in legs:
            points["C1"]["cpu_baseline"] = cpu_point(O, 4096, 256, 100, SEED + 31)
        note("points")

    # prime shards: what one GPU of an 8-GPU prime-sharded multiply runs
    shard = None
    if "shard" in legs:
        G = 8
        t = m.time_shard(da, db, dc, G, 20)
        m.run_device_sharded(da, db, dc, G)
        m.sync()
        part = (L + 3) * N * 8
        chain = sum(t[k] for k in ("rns", "col_fwd", "row_fused", "col_inv", "crt_partial"))
        shard = {"shards": G, "bounds": m.shard_bounds(G),
                 "shard0_ms": {k: round(v, 4) for k, v in t.items() if k != "sharded_multiply"},
                 "shard0_chain_ms": round(chain, 4),
                 "all_shards_here_ms": round(t["sharded_multiply"], 4),
                 "partial_bytes_per_shard": part,
                 "bit_exact_vs_plain": bool(np.array_equal(dc.download((N, L)), c_dev)),
                 "note": "shards run one after another on this GPU; on G GPUs each runs its chain at once, then a "
                         "reduce-scatter of the partials and the combine on N/G coefficients each"}
        note("shard")

    # end to end from pageable host buffers (the fherma_run replacement)
    pcie_bytes = 12 * N * L
    end_to_end = None
    if "e2e" in legs:
        c_host = np.empty((N, L), np.uint32)
        m.run(a, b, c_host)
        e2e = []
        for _ in range(args.e2e_calls):
            t = time.perf_counter()
            m.run(a, b, c_host)
            e2e.append(time.perf_counter() - t)
        e2e_med = float(np.median(e2e))
        # the same against the pinned copy rates measured here: one product's
        # floor is its inputs in (8 N L bytes) then its result out (4 N L)
        tr = pmlib.probe_transfers(64 << 20, 16, 3)
        floor_s = 8 * N * L / (tr["h2d_pinned"] * 1e9) + 4 * N * L / (tr["d2h_pinned"] * 1e9)
        end_to_end = {"mults_per_s": round(1.0 / e2e_med, 2), "median_ms": round(1e3 * e2e_med, 3),
                      "pct_pcie_peak": round(100 * pcie_bytes / e2e_med / (PCIE_PEAK_GBS * 1e9), 1),
                      "measured_pinned_GBps": {"h2d": tr["h2d_pinned"], "d2h": tr["d2h_pinned"]},
                      "floor_ms_at_measured": round(1e3 * floor_s, 3),
                      "pct_of_measured_floor": round(100 * floor_s / e2e_med, 1),
                      "bit_exact": bool(np.array_equal(c_host, c_dev))}
        note("e2e")

    # batched: the same point 16 times, and the mixed batch of 64 uneven products
    del m, da, db, dc
    # the serving leg's plain baseline, before the batch legs free gigabytes
    plain_serve = plain_mix(pmlib, O) if "serve" in legs else None
    batch_same = batch_mix = None
    if "batch" in legs:
        tune = not args.no_autotune
        batch_same = batched(pmlib, O, 16, [(N, W)] * 16, SEED + 100, 16, 50, tune)
        note("batched same point")
        rng = np.random.default_rng(SEED)
        mix = [(int(2 ** rng.integers(11, 17)), int(rng.integers(128, 2049))) for _ in range(64)]
        batch_mix = batched(pmlib, O, 64, mix, SEED + 200, 64, 50, tune)
        batch_mix["workload"] = "64 products, N log-uniform in 2^11..2^16, W uniform in [128, 2048], seed %d" % SEED
        note("batched mix")

    serve = serving(pmlib, O, (200, 1000, 5000), args.serve_seconds, plain_serve) if "serve" in legs else None
    if serve is not None:
        note("serve")

    cpu = cpu_baseline(q, args.cpu_cases) if "cpu" in legs else None
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "mults/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "settle_ms": round(settle_ms, 1), "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": "plain single product N=65536 W=1747 L=55 K=%d, mixed operands, "
                               "device-resident (pm_run_device)" % K, "N": N, "W": W, "L": L, "K": K,
                   "operands": "mixed per coefficient: 1/2 uniform [0,q), 1/4 q-1, 1/4 < 2^32", "seed": SEED},
        "bit_exact": correct,
        "pct_hbm_peak": round(100 * pcie_bytes / (ms_step * 1e-3) / (HBM_PEAK_GBS * 1e9), 3),
        # M = K (3 (N/2) log2 N + 2N) modmuls per product, over the throughput, against the VALU peak
        "modmuls_per_product": sum(mm.values()),
        "pct_valu_modmul_peak": round(100 * sum(mm.values()) / (ms_step * 1e-3) / PEAK_MODMUL_PER_S, 2),
        "pct_measured_valu_modmul_peak": round(100 * sum(mm.values()) / (ms_step * 1e-3) / peak_bfly, 2),
        "serial_multiply_ms": round(ms[0], 4),
        "roofline": roofline, "roofline_valu": roofline_valu, "kernels": per_kernel, "end_to_end": end_to_end,
        "points": points, "prime_shard": shard,
        "batched_same_point": batch_same, "batched_mixed": batch_mix, "serving": serve, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
