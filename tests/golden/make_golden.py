"""Writes the golden fixtures in this directory.

Inputs come from the seeded generator in oracle/oracle.py; the expected
products come from the Kronecker-substitution bigint product (no NTT, no RNS,
no CRT), so the fixtures pin the oracle rather than echo it. When the
reference's own emulator build (oracle/_ref/solution_emu) is present, every
fixture is also checked against it before it is written.

    python tests/golden/make_golden.py
"""
import json
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
from oracle import oracle as O  # noqa: E402

CASES = [  # (N, W, operand kind, seed)
    (1, 8, "uniform", 1),
    (2, 40, "uniform", 2),
    (16, 33, "max", 3),
    (64, 32, "uniform", 4),
    (64, 4000, "small16", 5),
    (256, 1747, "mixed", 6),
    (1024, 512, "sparse", 7),
]


def main():
    ref = O.ref_binary()
    index = {"method": "kronecker substitution (Python bigint)", "checked_against_reference_binary": bool(ref),
             "cases": []}
    for N, W, kind, seed in CASES:
        q = O.spec_q(N, W)
        a, b = O.operands(kind, N, W, q, seed)
        L = O.limbs_of(W)
        c = O.pack(O.kronecker(O.unpack(a), O.unpack(b), q), L)
        if ref is not None:
            with tempfile.TemporaryDirectory() as d:
                O.write_point(d, N, W, q, [(a, b)])
                subprocess.run([str(ref), d], check=True, capture_output=True)
                assert np.array_equal(O.read_output(d, 0, N, L), c), (N, W)
        name = "negacyclic_N%d_W%d_%s.npz" % (N, W, kind)
        np.savez_compressed(HERE / name, a=a, b=b, c=c)
        index["cases"].append({"file": name, "N": N, "W": W, "kind": kind, "seed": seed, "q": hex(q)})
        print("wrote", name)
    (HERE / "index.json").write_text(json.dumps(index, indent=1) + "\n")


if __name__ == "__main__":
    main()
