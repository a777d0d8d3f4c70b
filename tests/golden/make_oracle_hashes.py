"""SHA-256s of the CPU oracle's results on the large parity cases, for GPU tests
that cannot afford to run the serial oracle at full size.

Cases: the benchmark's 800,000-point six-ellipsoid union at depth 9 in mixed
precision; the 250,000-point noisy torus at depth 8 with point weights 0 and
4; the reference's 1,000,000-point torus fixture at depth 9
(docs/results/torus-1m-d9.json); the 20,000-point sphere at depth 6.  Per case:
the per-level basis key hashes, the hashes of the field's keys, level
corrections and cumulative coefficients, the isovalue's bits, and the hashes
of values, gradients and validity at a fixed query set.  Writes
oracle_hashes.json beside this script; run once in the build container (a few
minutes of serial CPU), the output is committed.
"""
import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
for p in (ROOT, ROOT / "warp-psr_amd", ROOT / "tests"):
    sys.path.insert(0, str(p))

import oracle  # noqa: E402
from golden_cases import LARGE_CASES as CASES, large_case_inputs as case_inputs  # noqa: E402


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def summary(name):
    p, n, q, params = case_inputs(name)
    t = time.perf_counter()
    r = oracle.reconstruct(p, n, **params)
    seconds = time.perf_counter() - t
    depth = params["depth"]
    values, grads, valid = r.evaluate(q)
    return {
        "params": params, "points": int(len(p)), "queries": int(len(q)), "oracle_seconds": round(seconds, 1),
        "input_sha256": digest(np.column_stack([p, n])),
        "basis_counts": [int(len(r.array("basis_keys", d))) for d in range(depth + 1)],
        "basis_keys": [digest(r.array("basis_keys", d)) for d in range(depth + 1)],
        "keys": digest(r.array("keys")), "coefficients": digest(r.array("coefficients")),
        "cumulative": digest(r.array("cumulative")), "iso_value": float(r.iso_value),
        "iso_hex": np.asarray(r.iso_value, dtype=r.array("coefficients").dtype).tobytes().hex(),
        "values": digest(values), "gradients": digest(grads), "valid": digest(valid),
    }


def main():
    out_path = HERE / "oracle_hashes.json"
    out = json.loads(out_path.read_text()) if out_path.exists() else {}
    for name in (sys.argv[1:] or CASES):
        out[name] = summary(name)
        print(name, out[name]["oracle_seconds"], "s", flush=True)
        out_path.write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
