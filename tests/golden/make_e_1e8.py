"""Write tests/golden/e_digits_gmp.json: SHA-256 of the first D significant
digits of e ("2.ddd...", the reference's output format) for large D, computed
with GMP by the reference's formula (n_total = D + 20 guard digits, N =
split_terms_needed(n_total), floor((B+P) 10^n_total / B), truncated).  The
reference itself cannot run these sizes (its transform length stops at 2^24).

    python tests/golden/make_e_1e8.py
"""
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import _oracle as O  # noqa: E402


def main():
    out = {}
    for d in (10 ** 6, 10 ** 7, 10 ** 8):
        t0 = time.time()
        s = O.gmp_e_digits(d)
        s = s[0] + "." + s[1:]
        out[str(d)] = {"sha256": hashlib.sha256(s.encode()).hexdigest(), "len": len(s),
                       "head": s[:64], "tail": s[-64:], "gmp_seconds": round(time.time() - t0, 1)}
        print(d, out[str(d)]["sha256"], out[str(d)]["gmp_seconds"], "s", flush=True)
    with open(os.path.join(HERE, "e_digits_gmp.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
