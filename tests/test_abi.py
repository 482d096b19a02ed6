"""CPU checks of the drop-in boundary: libpolymult.so loads, exports every
symbol include/polymult.h declares plus the reference's three C++ entry
points, and its host-side point validation agrees with the oracle and with
the reference's messages (solve.cu:873-913). No compute call touches a GPU.
"""
import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared_c_symbols():
    text = (ROOT / "include" / "polymult.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(pm_\w+)\s*\(", text, re.M)))


def test_header_and_binding_agree(pm):
    assert sorted(pm.C_SYMBOLS) == _declared_c_symbols()


def test_library_exports(pm):
    lib = pm.LIB_PATH
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib)], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for name in _declared_c_symbols():
        assert name in exported, name
    for name in pm.CXX_SYMBOLS:
        assert name in exported, name


@pytest.mark.parametrize("N,W", [(1, 8), (64, 32), (4096, 256), (8192, 438), (32768, 868), (65536, 1747),
                                 (512, 4000)])
def test_point_info_matches_oracle(O, pm, N, W):
    q = O.spec_q(N, W)
    L = O.limbs_of(W)
    K, delta, primes = pm.point_info(N, W, L, O.q_limbs(q, L))
    P = O.OraclePlan(N, W, q)
    assert K == P.K and delta == P.delta
    assert np.array_equal(primes, P.primes)


@pytest.mark.parametrize("args,msg", [
    ((48, 32, 1, "q"), "N must be a power of two"),
    ((64, 32, 2, "q2"), "inconsistent W / L"),
    ((64, 32, 1, "even"), "q out of range"),
    ((64, 32, 1, "tiny"), "too far below"),
    ((64, 64, 2, "toplimb"), "top limb is zero"),
    ((1 << 26, 4, 1, "q13"), "not enough NTT primes"),
    ((1, 15900, 497, "wide"), "too many RNS primes"),
])
def test_point_errors(O, pm, args, msg):
    """The reference's refusals (solve.cu:873-913), and only those: N = 2^26
    has a single prime 1 mod 2^27 below 2^31, and W = 15900 needs more than
    1023 primes."""
    N, W, L, which = args
    q = O.spec_q(64, 32)
    qv = {"q": np.array([q], np.uint32), "q2": np.array([q, 0], np.uint32),
          "even": np.array([q - 1], np.uint32), "tiny": np.array([3], np.uint32),
          "toplimb": np.array([q, 0], np.uint32), "q13": np.array([13], np.uint32),
          "wide": O.q_limbs((1 << 15900) - 3, 497)}[which]
    with pytest.raises(pm.PolymultError, match=msg):
        pm.point_info(N, W, L, qv)


@pytest.mark.parametrize("N,W,delta", [(131072, 64, 0), (262144, 32, 0), (1 << 20, 16, 15), (1 << 25, 8, 5),
                                       (512, 4065, 1000001), (64, 5000, 1000001), (16, 8192, 1000001),
                                       (1, 15840, 3)])
def test_points_beyond_two_passes_and_4064_bits(O, pm, N, W, delta):
    """Points the reference accepts past N = 65536 (three and four NTT passes)
    and past W = 4064 (the 64-bit RNS epilogue, the split CRT): pm_point_info
    takes them, with the oracle's K and primes (delta 0: the spec's prime q;
    else q = 2^W - delta, any odd modulus the reference accepts)."""
    q = O.spec_q(N, W) if delta == 0 else (1 << W) - delta
    L = O.limbs_of(W)
    K, d, primes = pm.point_info(N, W, L, O.q_limbs(q, L))
    assert d == (1 << W) - q
    if N * K <= 1 << 22:
        P = O.OraclePlan(N, W, q)
        assert K == P.K and np.array_equal(primes, P.primes)
    else:
        # 2^25: the two primes 1 mod 2^26 below 2^31, largest first
        assert list(primes) == [2013265921, 1811939329]


def test_q_limb_count_checked(O, pm):
    q = O.spec_q(64, 64)
    with pytest.raises(pm.PolymultError, match="q must have L limbs"):
        pm.point_info(64, 64, 2, O.q_limbs(q, 3))


def test_struct_layouts_match_header(pm, tmp_path):
    """The ctypes mirrors of the header's structs have the C compiler's sizes and offsets."""
    import subprocess
    structs = {"pm_info": pm.PmInfo, "pm_server_config": pm.ServerConfig, "pm_server_stats": pm.ServerStats,
               "pm_serve_report": pm.ServeReport}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "polymult.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in py._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    root = Path(__file__).resolve().parents[1]
    subprocess.run(["gcc", "-std=c99", "-I", str(root / "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n") if l)
    for cname, py in structs.items():
        assert int(got[cname]) == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got["%s.%s" % (cname, f)]) == getattr(py, f).offset, (cname, f)


def test_twiddle_companion_recovers_twiddle(O):
    """The NTT tables hold only the Shoup companion w' = floor(w 2^32 / p) of
    each twiddle; the kernels recover w = floor(w' p / 2^32) + 1. That holds
    for every w in [1, p) (no twiddle is zero, all are powers of psi) and
    fails for w = 0. Checked exhaustively for small primes, and for the 114
    primes of the N=65536/W=1747 plan on w = 1, 2, p - 1 and 20,000 random w
    each, plus the plan's own twiddles psi^k for k < 4096."""
    def recover(w, p):
        c = (w << 32) // p
        return ((c * p) >> 32) + 1

    for p in (3, 5, 17, 97, 257, 7681, 12289):
        assert all(recover(w, p) == w for w in range(1, p)), p
        assert recover(0, p) != 0
    rng = np.random.default_rng(2026)
    P = O.OraclePlan(65536, 1747, O.spec_q(65536, 1747))
    for p, psi in zip(P.primes.tolist(), P.psi.tolist()):
        ws = [1, 2, p - 1] + rng.integers(1, p, 20000).tolist()
        assert all(recover(w, p) == w for w in ws), p
        w = 1
        for _ in range(4096):
            assert recover(w, p) == w, (p, w)
            w = w * psi % p
