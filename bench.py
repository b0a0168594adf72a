"""Benchmark: one 2^30-bit x 2^30-bit product per step on one MI355X.

Workload: two seeded uniform-random 2^30-bit integers (top bit set), 16-bit
coefficients, 3-prime CRT, NTT length 2^27, operands already resident in HBM.
A step is one full device product (split, 3 x (2 forward + 1 inverse) length-
2^27 transforms, pointwise, Garner, carry); `value` is ms per product.

Roofline: the NTT stages (forward passes, pointwise, inverse passes) are
priced with the algorithmic bytes B_mul = 104 * k * L (k = 3 primes, L = 2^27:
2 HBM passes of 16 B/coefficient per transform, 3 transforms, plus one extra
8 B/coefficient read of the second spectrum) against 8 TB/s.  Their time is
measured with HIP events recorded on the engine's own stream around every
launch.  The dominant kernel is the forward column pass (k_fast_pass<LOGS, 0>);
its per-launch algorithmic bytes are 2 operands x 3 primes x 16 B x L.

CPU baseline: the oracle's 3-prime NTT multiply (one OpenMP thread per prime)
on a bounded sample (a 2^26-bit product, L = 2^23), scaled to 2^30 bits by
L log L; GMP mpz_mul on the same sample is reported beside it.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
K_PRIMES = 3
FWD = "k_fast_pass<fwd>"     # the dominant kernel: a forward column pass over both operands


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--bits", type=int, default=30, help="log2 of operand bits")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--profile-only", action="store_true",
                    help="run warmup + steps of the product only (for rocprofv3), print nothing else")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 1 GiB device-to-device copy (torch) to calibrate rocprofv3 byte counters")
    return ap.parse_args()


def kernel_times(L):
    names = ctypes.create_string_buffer(32 * 64)
    ms = (ctypes.c_double * 64)()
    n = L.ntt_engine_kernel_times(names, ms, 64)
    out = []
    for i in range(n):
        nm = names.raw[32 * i: 32 * (i + 1)].split(b"\0", 1)[0].decode()
        out.append((nm, ms[i]))
    return out


def cpu_baseline(sample_bits=26, target_bits=30):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    a = O.random_operand(101, 1 << sample_bits)
    b = O.random_operand(102, 1 << sample_bits)
    t0 = time.perf_counter()
    O.mul_ntt3(a, b, threads=3)
    t_port = time.perf_counter() - t0
    t_gmp = O.gmp().gmpref_time_mul(O._ptr(a), len(a), O._ptr(b), len(b), 1)
    ls, lt = sample_bits - 3, target_bits - 3        # log2 transform lengths
    scale = (2 ** lt * lt) / (2 ** ls * ls)
    return {"value": t_port * scale * 1e3, "unit": "ms/mul (2^30-bit, scaled from sample)", "cores": 3,
            "kind": "port",
            "sample": f"one 2^{sample_bits}-bit x 2^{sample_bits}-bit product (L=2^{ls}) with the oracle's "
                      f"3-prime NTT multiply, 3 OpenMP threads: {t_port * 1e3:.1f} ms; scaled by L log L x{scale:.2f}",
            "gmp_sample_ms": t_gmp * 1e3, "gmp_scaled_ms": t_gmp * 1e3 * (2 ** (target_bits - sample_bits)) ** 1.1,
            "nproc": os.cpu_count()}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        # a plain copy kernel of known size: 1 GiB read + 1 GiB written
        import torch
        src = torch.ones(1 << 27, dtype=torch.int64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 2 * (1 << 30)}))
        return
    import cryptography_amd as E
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O

    L = E.lib()
    E.init()
    nbits = 1 << args.bits
    a = O.random_operand(1, nbits)
    b = O.random_operand(2, nbits)
    na, nb = len(a), len(b)
    logL = E.log_len(na, nb)
    Llen = 1 << logL
    da, db = E.DeviceBuffer.from_array(a), E.DeviceBuffer.from_array(b)
    dc = E.DeviceBuffer(na + nb)

    for _ in range(args.warmup):
        E.mul_device(dc, da, na, db, nb)
    # correctness of what is timed: full product vs GMP (bit-exact) on the first step
    got = dc.download()
    want = O.gmp_mul(a, b)
    if not np.array_equal(got, want):
        print("bench.py: GPU product differs from GMP", file=sys.stderr)
        sys.exit(1)

    if args.profile_only:
        for _ in range(args.steps):
            E.mul_device(dc, da, na, db, nb, sync=False)
        E.engine.check(L.ntt_engine_sync())
        print(json.dumps({"profile_only": True, "steps": args.steps, "transform_len": Llen}))
        return

    # timed region: K device products back to back, synchronized on both sides
    E.engine.check(L.ntt_engine_sync())
    t0 = time.perf_counter()
    for _ in range(args.steps):
        E.mul_device(dc, da, na, db, nb, sync=False)
    E.engine.check(L.ntt_engine_sync())
    t_total = time.perf_counter() - t0
    ms_per_step = t_total * 1e3 / args.steps

    # per-launch event timings on the engine stream (separate, instrumented steps)
    L.ntt_engine_set_timing(1)
    ntt_ms, tot_ms, per = [], [], {}
    for _ in range(max(3, args.steps // 2)):
        E.mul_device(dc, da, na, db, nb)
        ntt_ms.append(L.ntt_engine_last_ntt_ms())
        tot_ms.append(L.ntt_engine_last_total_ms())
        for nm, ms in kernel_times(L):
            per.setdefault(nm, []).append(ms)
    L.ntt_engine_set_timing(0)
    sq_ms = []
    for _ in range(3):
        E.sqr_device(dc, da, na)
    E.engine.check(L.ntt_engine_sync())
    t0 = time.perf_counter()
    for _ in range(args.steps):
        E.sqr_device(dc, da, na, sync=False)
    E.engine.check(L.ntt_engine_sync())
    sq_ms = (time.perf_counter() - t0) * 1e3 / args.steps

    t_ntt = float(np.median(ntt_ms))
    b_mul = 104 * K_PRIMES * Llen
    n_fwd = sum(1 for nm, _ in kernel_times(L) if nm == FWD) or 1
    fwd_avg = float(np.mean(per.get(FWD, [0.0])))
    fwd_bytes = 2 * K_PRIMES * 16 * Llen
    roof = {"bound": "hbm", "kernel": FWD, "achieved": fwd_bytes / (fwd_avg * 1e-3) / 1e9 if fwd_avg else None,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "traffic": None,
            "launch_ms": fwd_avg, "launches_per_mul": n_fwd,
            "ntt_stages_ms": t_ntt, "ntt_stages_achieved": b_mul / (t_ntt * 1e-3) / 1e9}
    roof["frac"] = roof["achieved"] / HBM_PEAK_GBS if roof["achieved"] else None
    roof["ntt_stages_frac"] = roof["ntt_stages_achieved"] / HBM_PEAK_GBS

    line = {
        "metric": "ms per 2^30-bit mul (%HBM peak)",
        "value": ms_per_step, "unit": "ms/mul", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic",
        "config": {"workload": f"2^{args.bits}-bit x 2^{args.bits}-bit product, 16-bit coefficients, "
                               f"3-prime CRT, NTT length 2^{logL}, operands resident in HBM",
                   "operand_bits": nbits, "transform_len": Llen, "primes": K_PRIMES, "seed": [1, 2]},
        "square_ms": sq_ms,
        "kernels_ms": {k: float(np.mean(v)) for k, v in per.items()},
        "total_event_ms": float(np.median(tot_ms)),
        "roofline": roof,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    for x in (da, db, dc):
        x.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
