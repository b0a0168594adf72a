"""Benchmark of the MI355X CRT-NTT multiply engine, one GPU.

Headline (the JSON line's `value`): one 2^30-bit x 2^30-bit product per step,
16-bit coefficients, 3-prime CRT, NTT length 2^27, operands resident in HBM;
`value` is ms per product (lower is better).  A step is the whole device
product: split, 3 x (2 forward + 1 inverse) length-2^27 transforms, pointwise
product, Garner CRT, carry.  The product is checked bit-exact against GMP
mpz_mul before timing.

Roofline: the NTT stages (forward passes, fused pointwise, inverse passes) are
priced with the algorithmic bytes B_mul = 104 * k * L (k = 3 primes,
L = 2^27: two HBM passes of 16 B per coefficient for each of the 3 transforms
plus one 8 B read of the second spectrum) against 8 TB/s.  Kernel times come
from HIP events recorded after every launch on the engine's own stream.  The
dominant kernel is the launch with the longest median time; its algorithmic
bytes per launch come from launch_table(), its HBM traffic per launch from the
committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (profiles/).

Other workloads (reported under "workloads"; skip with --headline-only):
  mul22   one 2^22 x 2^22-bit product (L = 2^19): latency, and 256 of them as
          one graph-replayed batch
  batch   4096 products, operand bits 2^(12 + 10u) (u uniform, independent),
          one hipGraph replay per step, vs the same products launched one by one
  e       e to 10^8 digits with the level-batched product tree (digits checked
          by SHA-256 against GMP), and the tree alone at 10^6 and 10^7 digits,
          level-batched vs one launch sequence per product
  serve   Poisson arrivals, operand bits 2^(10 + 18u), operands resident,
          64 GB working-memory budget, swept arrival rates: products/s, p50/p99
  convert bigint_to_decimal of a 10^6-digit integer and newton_reciprocal of a
          2^22-bit integer (the reference's base_convert.h / newton.h entry points)
CPU baseline: the oracle's 3-prime NTT multiply of the headline's 2^30-bit
operands on up to 16 host threads, measured at full size, with GMP mpz_mul of
the same operands (one core) beside it.
"""
import argparse
import ctypes
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
K_PRIMES = 3


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--bits", type=int, default=30, help="log2 of operand bits (headline)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--headline-only", action="store_true", help="skip the other workloads")
    ap.add_argument("--workloads", default="mul22,batch,e,serve,convert")
    ap.add_argument("--e-digits", type=int, default=10 ** 8)
    ap.add_argument("--profile-only", action="store_true",
                    help="run warmup + steps of the product only (for rocprofv3), print nothing else")
    ap.add_argument("--profile-op", choices=("mul", "sqr"), default="mul",
                    help="with --profile-only: the product (mul) or the square of the first operand (sqr)")
    ap.add_argument("--tuning", default=os.path.join(ROOT, "cryptography_amd", "tuning_mi355x.json"),
                    help="tuning file to load (ntt_engine_autotune_save format); 'none' keeps the defaults")
    ap.add_argument("--tree-compare", type=int, default=0, metavar="DIGITS",
                    help="only compare the batched and per-product product trees of e at DIGITS digits")
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


def b_alg(logL):
    return 104 * K_PRIMES * (1 << logL)


def pass_plan(L, logL):
    """The engine's pass split of a length-2^logL transform (ntt_engine_get_plan)."""
    L.ntt_engine_get_plan.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    out = (ctypes.c_int * 8)()
    n = L.ntt_engine_get_plan(logL, out, 8)
    return [out[i] for i in range(n)]


def has_xtab(logS, logT):
    """Whether a pass of 2^logS x 2^logT takes its cross twiddles from a table (xtab_for in
    ntt_engine.hip: S of 8..1024 points and a block of 2^8..2^20 residues)."""
    return 3 <= logS <= 10 and 8 <= logS + logT <= 20


def launch_table(L, logL, na, nb, square=False):
    """Per launch of one product (in launch order): rocprof kernel name, blocks, algorithmic bytes.
    Algorithmic bytes: residues are 8 B, k = 3 primes; forward passes cover both operands (one for a
    square), the first one reads the operand limbs instead of residues; the fused middle pass reads
    both spectra (one for a square) and writes one; inverse passes read and write one row set;
    Garner reads 3 residue rows and writes the product limbs."""
    nops = 1 if square else 2
    Lw = 1 << logL
    ntiles = max(1, Lw // 4096)
    plan = pass_plan(L, logL)
    L.ntt_engine_get_tile_batch.restype = ctypes.c_int
    batch = L.ntt_engine_get_tile_batch()

    def blocks(kind):
        """Blocks per (operand, prime) row: tiles per block as enqueue_ntt_class sets them."""
        cap = {6: max(batch, 8), 2: max(batch, 8), 7: max(batch, 8), 5: max(batch, 4), 9: max(batch, 4)}.get(kind, batch)
        tpb = min(max(ntiles * K_PRIMES // 4096, 1), cap)
        return -(-ntiles // tpb)

    out, done = [], 0
    for i, s in enumerate(plan[:-1]):
        # the first pass takes its chain ratios from a table at S = 2^6 and 2^9 (kind 8)
        kind = (8 if s in (6, 9) else 3) if i == 0 else (6 if has_xtab(s, logL - done - s) else 0)
        byt = 8 * (na + (0 if square else nb)) + nops * K_PRIMES * 8 * Lw if i == 0 else nops * K_PRIMES * 16 * Lw
        out.append((f"k_fast_pass<{s}, {kind}>", blocks(kind) * K_PRIMES * nops, byt))
        done += s
    # the wave-local fused pass (S = 2^6, 2^9) takes several tiles per block
    mid_blocks = blocks(2) if len(plan) > 1 and plan[-1] in (6, 9) else ntiles
    out.append((f"k_fast_pass<{plan[-1]}, {2 if len(plan) > 1 else 4}>", mid_blocks * K_PRIMES,
                K_PRIMES * (16 if square else 24) * Lw))
    for i in range(len(plan) - 2, -1, -1):
        done -= plan[i]
        logT = logL - done - plan[i]
        logC = min(logT, 12 - plan[i])
        kind = 5 if plan[i] >= 3 and logC + plan[i] == 12 else 1
        if kind == 5:   # cross twiddles from a table (7), or chain ratios from one (9)
            kind = 7 if has_xtab(plan[i], logT) else 9
        out.append((f"k_fast_pass<{plan[i]}, {kind}>", blocks(kind) * K_PRIMES, K_PRIMES * 16 * Lw))
    out.append(("k_garner_pack", None, K_PRIMES * 8 * Lw + 8 * (na + nb)))
    return out


def profile_entry(name, blocks, square=False):
    """A launch shape's entry in the committed rocprofv3 summary (profiles/mul_2p30_kernels.json)."""
    path = os.path.join(ROOT, "profiles", "mul_2p30_kernels.json")
    if not os.path.exists(path):
        return {}
    return json.load(open(path))["kernels"].get(f"{name}|{blocks}" + ("|sqr" if square else ""), {})


def profile_traffic(name, blocks, square=False):
    """HBM bytes per launch of a kernel from the committed rocprofv3 counter summary
    (FETCH_SIZE and WRITE_SIZE, scaled by a copy calibration)."""
    k = profile_entry(name, blocks, square)
    if k.get("fetch_bytes") is None or k.get("write_bytes") is None:
        return None
    return k["fetch_bytes"] + k["write_bytes"]


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ---- headline: one 2^30-bit product ----------------------------------------

def headline(args, E, O):
    L = E.lib()
    nbits = 1 << args.bits
    a = O.random_operand(1, nbits)
    b = O.random_operand(2, nbits)
    na, nb = len(a), len(b)
    logL = E.log_len(na, nb)
    Llen = 1 << logL
    da, db = E.DeviceBuffer.from_array(a), E.DeviceBuffer.from_array(b)
    dc = E.DeviceBuffer(na + nb)
    for _ in range(max(1, args.warmup)):
        E.mul_device(dc, da, na, db, nb)
    # GMP's product is the parity check; its time (one core, limb import and
    # export included) is reported beside the CPU baseline
    t0 = time.perf_counter()
    want = O.gmp_mul(a, b)
    t_gmp_mul = time.perf_counter() - t0
    if not np.array_equal(dc.download(), want):
        raise SystemExit("bench.py: GPU product differs from GMP")
    if args.profile_only:
        for _ in range(args.steps):
            if args.profile_op == "sqr":
                E.sqr_device(dc, da, na, sync=False)
            else:
                E.mul_device(dc, da, na, db, nb, sync=False)
        E.engine.check(L.ntt_engine_sync())
        return {"profile_only": True, "op": args.profile_op, "steps": args.steps, "transform_len": Llen}
    E.engine.check(L.ntt_engine_sync())
    t0 = time.perf_counter()
    for _ in range(args.steps):
        E.mul_device(dc, da, na, db, nb, sync=False)
    E.engine.check(L.ntt_engine_sync())
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    # per-launch event timings on the engine stream, in steady state: each
    # sample is the last of four products enqueued back to back (the first
    # pass of the first one or two products after the device idles runs up to
    # 0.6 ms slower: with identical L2 hit counts and HBM bytes, it is the
    # idle-to-busy ramp, not the kernel)
    L.ntt_engine_set_timing(1)
    ntt_ms, tot_ms, per = [], [], {}
    for _ in range(max(3, args.steps // 2)):
        for _ in range(4):
            E.mul_device(dc, da, na, db, nb, sync=False)
        E.engine.check(L.ntt_engine_sync())
        ntt_ms.append(L.ntt_engine_last_ntt_ms())
        tot_ms.append(L.ntt_engine_last_total_ms())
        for i, (nm, ms) in enumerate(kernel_times(L)):
            per.setdefault(f"{i}:{nm}", []).append(ms)
    L.ntt_engine_set_timing(0)
    for _ in range(2):
        E.sqr_device(dc, da, na)
    t0 = time.perf_counter()
    want_sq = O.gmp_mul(a, a)
    t_gmp_sqr = time.perf_counter() - t0
    if not np.array_equal(dc.download(), want_sq):
        raise SystemExit("bench.py: GPU square differs from GMP")
    del want_sq
    t0 = time.perf_counter()
    for _ in range(args.steps):
        E.sqr_device(dc, da, na, sync=False)
    E.engine.check(L.ntt_engine_sync())
    sq_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    # per-launch event timings of the square
    L.ntt_engine_set_timing(1)
    sq_ntt, sq_per = [], {}
    for _ in range(max(3, args.steps // 2)):
        for _ in range(4):
            E.sqr_device(dc, da, na, sync=False)
        E.engine.check(L.ntt_engine_sync())
        sq_ntt.append(L.ntt_engine_last_ntt_ms())
        for i, (nm, ms) in enumerate(kernel_times(L)):
            sq_per.setdefault(f"{i}:{nm}", []).append(ms)
    L.ntt_engine_set_timing(0)
    for x in (da, db, dc):
        x.close()
    # the same product through the host-buffer ABI (ntt_mul_limbs: operands in
    # and product out over PCIe), for reference only: value is device-resident
    E.mul(a[:1024], b[:1024])
    host_ms = []
    for _ in range(3):
        t0 = time.perf_counter()
        E.mul(a, b)
        host_ms.append((time.perf_counter() - t0) * 1e3)
    t_ntt = float(np.median(ntt_ms))
    kern = {k: float(np.median(v)) for k, v in per.items()}
    # the dominant kernel: the launch with the longest median time
    table = launch_table(L, logL, na, nb)
    times = list(kern.values())
    i_dom = int(np.argmax(times[:len(table)]))
    name, blocks, alg = table[i_dom]
    t_dom = times[i_dom]
    roof = {"bound": "hbm", "kernel": f"{name} (launch {i_dom} of the product)",
            "achieved": alg / (t_dom * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "traffic": profile_traffic(name, blocks) if (nbits == 1 << 30) else None,
            "alg_bytes_per_launch": alg, "launch_ms": t_dom,
            "ntt_stages_ms": t_ntt, "ntt_stages_achieved": b_alg(logL) / (t_ntt * 1e-3) / 1e9,
            "b_alg_bytes": b_alg(logL)}
    roof["frac"] = roof["achieved"] / HBM_PEAK_GBS
    roof["ntt_stages_frac"] = roof["ntt_stages_achieved"] / HBM_PEAK_GBS
    # every launch of the product at its algorithmic bytes, for comparison with the dominant one
    roof["per_launch"] = [{"kernel": nm, "launch_ms": t, "alg_bytes": ab,
                           "achieved": ab / (t * 1e-3) / 1e9, "frac": ab / (t * 1e-3) / 1e9 / HBM_PEAK_GBS}
                          for (nm, _, ab), t in zip(table, times)]
    # VALU instructions per butterfly of the dominant kernel (committed SQ counters)
    if nbits == 1 << 30:
        roof["valu_instr_per_butterfly"] = profile_entry(name, blocks).get("valu_per_butterfly")
        roof["valu_instr_per_butterfly_by_kernel"] = {
            nm: round(profile_entry(nm, bl).get("valu_per_butterfly") or 0, 1) for nm, bl, _ in table
            if profile_entry(nm, bl).get("valu_per_butterfly")}
    # the square: B_sqr = 64 k L against its NTT stages, and its dominant launch
    b_sqr = 64 * K_PRIMES * Llen
    sq_table = launch_table(L, logL, na, na, square=True)
    sq_times = [float(np.median(v)) for v in sq_per.values()]
    j_dom = int(np.argmax(sq_times[:len(sq_table)]))
    sq_name, sq_blocks, sq_alg = sq_table[j_dom]
    t_sq_ntt = float(np.median(sq_ntt))
    sq_roof = {"bound": "hbm", "kernel": f"{sq_name} (launch {j_dom} of the square)",
               "achieved": sq_alg / (sq_times[j_dom] * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "traffic": profile_traffic(sq_name, sq_blocks, square=True) if nbits == 1 << 30 else None,
               "alg_bytes_per_launch": sq_alg, "launch_ms": sq_times[j_dom],
               "ntt_stages_ms": t_sq_ntt, "b_alg_bytes": b_sqr,
               "ntt_stages_achieved": b_sqr / (t_sq_ntt * 1e-3) / 1e9}
    sq_roof["frac"] = sq_roof["achieved"] / HBM_PEAK_GBS
    sq_roof["ntt_stages_frac"] = sq_roof["ntt_stages_achieved"] / HBM_PEAK_GBS
    sq_roof["per_launch"] = [{"kernel": nm, "launch_ms": t, "alg_bytes": ab,
                              "frac": ab / (t * 1e-3) / 1e9 / HBM_PEAK_GBS}
                             for (nm, _, ab), t in zip(sq_table, sq_times)]
    return {"operands": (a, b), "product": want, "gmp_mul_s": t_gmp_mul, "gmp_sqr_s": t_gmp_sqr,
            "ms_per_step": ms_per_step, "square_ms": sq_ms, "square_roofline": sq_roof, "host_api_ms": float(np.median(host_ms)), "logL": logL, "Llen": Llen, "nbits": nbits,
            "kernels_ms": kern, "total_event_ms": float(np.median(tot_ms)), "roofline": roof}


# ---- config 1: 2^22-bit products -------------------------------------------

def workload_mul22(E, O):
    from cryptography_amd import engine as EN
    a = O.random_operand(1, 1 << 22)
    b = O.random_operand(2, 1 << 22)
    na, nb = len(a), len(b)
    da, db = E.DeviceBuffer.from_array(a), E.DeviceBuffer.from_array(b)
    dc = E.DeviceBuffer(na + nb)
    E.mul_device(dc, da, na, db, nb)
    want = O.gmp_mul(a, b)
    ok = bool(np.array_equal(dc.download(), want))
    reps = 50
    t0 = time.perf_counter()
    for _ in range(reps):
        E.mul_device(dc, da, na, db, nb)          # synchronous: latency per product
    lat_ms = (time.perf_counter() - t0) * 1e3 / reps
    n = 256                                        # 256 copies as one graph-replayed batch
    pool = EN.DevicePool(n * (na + nb) + 16)
    items, outs = [], []
    for i in range(n):
        oc = pool.take(na + nb)
        outs.append(oc)
        items.append((da.ptr, na, db.ptr, nb, pool.ptr(oc), na + nb, None, 0))
    g = EN.BatchGraph(EN.make_descs(items))
    g.launch()
    EN.sync()
    ok = ok and bool(np.array_equal(pool.download(outs[-1], na + nb), want))
    t0 = time.perf_counter()
    for _ in range(5):
        g.launch()
    EN.sync()
    t = (time.perf_counter() - t0) / 5
    g.close()
    pool.close()
    for x in (da, db, dc):
        x.close()
    logL = E.log_len(na, nb)
    # CPU beside it: GMP mpz_mul on this product (one core), and the reference's schoolbook
    # (its path for this size, multiply.c:35-67, restated by the oracle) on a 2^18-bit sample
    # scaled by (2^22 / 2^18)^2
    t_gmp = O.gmp().gmpref_time_mul(O._ptr(a), na, O._ptr(b), nb, 5)
    sa, sb = a[: na // 16].copy(), b[: nb // 16].copy()
    t0 = time.perf_counter()
    O.mul_schoolbook(sa, sb)
    t_sb = (time.perf_counter() - t0) * 256
    return {"bit_exact_vs_gmp": ok, "latency_ms": lat_ms, "transform_len": 1 << logL,
            "batch256_products_per_s": n / t, "batch256_ms": t * 1e3,
            "batch256_ntt_alg_gbs": n * b_alg(logL) / t / 1e9,
            "cpu_gmp_ms": t_gmp * 1e3, "cpu_schoolbook_est_ms": t_sb * 1e3,
            "cpu_schoolbook_sample": "oracle schoolbook on 2^18 x 2^18 bits, x256"}


# ---- config 4: 4096 mixed products in one graph ----------------------------

def workload_batch(E, O):
    from cryptography_amd import engine as EN
    import _workloads as W
    n = 4096
    bits = W.config4_sizes(n)
    la, lb = (bits[:, 0] + 63) // 64, (bits[:, 1] + 63) // 64
    pool, items, host = W.config4_batch(EN, n)
    descs = EN.make_descs(items)
    t0 = time.perf_counter()
    g = EN.BatchGraph(descs)
    t_capture = time.perf_counter() - t0
    g.launch()
    EN.sync()
    got = [pool.download(oc, len(a) + len(b)) for (a, b, oc) in host]
    t0 = time.perf_counter()
    want = [O.gmp_mul(a, b) for (a, b, _) in host]      # GMP on one core: the CPU beside it
    t_gmp = time.perf_counter() - t0
    bad = sum(1 for g_, w_ in zip(got, want) if not np.array_equal(g_, w_))
    del got, want
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        g.launch()
    EN.sync()
    t = (time.perf_counter() - t0) / reps
    # the same products, one launch sequence each (device-resident, no batching)
    L = E.lib()
    t0 = time.perf_counter()
    for (pa, na, pb, nb, pc, _, _, _) in items:
        E.engine.check(L.ntt_mul_device_async(pc, pa, na, pb, nb))
    EN.sync()
    t_each = time.perf_counter() - t0
    thr = L.ntt_engine_get_threshold()
    ntt_bytes = sum(b_alg(E.log_len(int(x), int(y))) for x, y in zip(la, lb) if max(x, y) > thr)
    g.close()
    # the same graph at the reference's fixed crossover of 64 limbs (multiply.h:23-25)
    t_default = None
    if thr != 64:
        L.ntt_engine_set_threshold(64)
        g2 = EN.BatchGraph(descs)
        g2.launch()
        EN.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            g2.launch()
        EN.sync()
        t_default = (time.perf_counter() - t0) / reps
        g2.close()
        L.ntt_engine_set_threshold(thr)
    pool.close()
    return {"products": n, "all_bit_exact_vs_gmp": bad == 0, "mismatches": bad, "graph_ms": t * 1e3,
            "products_per_s": n / t, "ntt_alg_gbs": ntt_bytes / t / 1e9, "capture_ms": t_capture * 1e3,
            "unbatched_ms": t_each * 1e3, "speedup_vs_unbatched": t_each / t, "schoolbook_threshold_limbs": thr,
            "graph_ms_at_threshold_64": None if t_default is None else t_default * 1e3,
            "cpu_gmp_ms": t_gmp * 1e3, "cpu_gmp_note": "GMP mpz_mul of the same 4096 products, one core"}


# ---- config 3: e ------------------------------------------------------------

def workload_e(E, O, digits):
    from cryptography_amd import engine as EN
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "e_digits_gmp.json")))
    out = {"digits": digits}
    t0 = time.perf_counter()
    s, st = EN.compute_e(digits)
    out["seconds"] = time.perf_counter() - t0
    out["phases_s"] = {k: st[k] for k in ("t_tree", "t_recip", "t_div", "t_radix", "t_total")}
    out["n_terms"] = st["n_terms"]
    out["device_products"] = st["products"]
    ent = gold.get(str(digits))
    out["sha256_matches_gmp"] = bool(ent and hashlib.sha256(s.encode()).hexdigest() == ent["sha256"])
    comp = {}
    for d in (10 ** 6, 10 ** 7):
        comp[str(d)] = tree_compare(EN, O, d)
    out["tree_batched_vs_per_product"] = comp
    # CPU beside it: GMP binary splitting + division + conversion, one core, at 10^6 digits
    t0 = time.perf_counter()
    O.gmp_e_digits(10 ** 6)
    out["cpu_gmp_e_1e6_s"] = time.perf_counter() - t0
    return out


def tree_compare(EN, O, digits, per_product=True):
    """The product tree of e at `digits` digits, level-batched in one hipGraph
    against one launch sequence per product (the reference's plain path on the
    same device arena): end to end (planning, and graph capture for the
    batched tree, included) and evaluation alone, each on its own tree."""
    import ctypes
    L = EN.lib()
    vp = ctypes.c_void_p
    L.ntt_split_tree_create.argtypes = [ctypes.c_long, ctypes.c_long, ctypes.c_int, ctypes.POINTER(vp)]
    L.ntt_split_tree_run.argtypes = [vp]
    L.ntt_split_tree_destroy.argtypes = [vp]
    n = O.orc().orc_split_terms_needed(digits + 20)

    def leg(mode):
        h = vp()
        t0 = time.perf_counter()
        EN.check(L.ntt_split_tree_create(1, n + 1, mode, ctypes.byref(h)))
        t1 = time.perf_counter()
        EN.check(L.ntt_split_tree_run(h))
        t2 = time.perf_counter()
        runs = 3 if mode == 1 else 1
        for _ in range(runs if mode == 1 else 0):
            EN.check(L.ntt_split_tree_run(h))
        t3 = time.perf_counter()
        L.ntt_split_tree_destroy(h)
        return {"plan_s": t1 - t0, "first_run_s": t2 - t1, "end_to_end_s": t2 - t0,
                "replay_s": (t3 - t2) / runs if mode == 1 else t2 - t1}

    out = {"terms": n, "batched": leg(1)}
    if per_product:
        out["per_product"] = leg(2)
        out["speedup_end_to_end"] = out["per_product"]["end_to_end_s"] / out["batched"]["end_to_end_s"]
        out["speedup_evaluation"] = out["per_product"]["first_run_s"] / out["batched"]["replay_s"]
    return out


# ---- next to the path: decimal conversion (base_convert.h) ----------------

def workload_convert(E, O):
    """bigint_to_decimal of a random 10^6-digit integer (checked against the digit
    string it was made from), newton_reciprocal of a 2^22-bit integer, and
    ntt_to_decimal of a 10^7-digit integer against GMP (timed on one host core)."""
    import random
    import test_arith_ext as T
    if hasattr(sys, "set_int_max_str_digits"):
        sys.set_int_max_str_digits(0)
    L = T._setup(E.lib())
    t0 = time.perf_counter()
    L.base_convert_init(20)
    t_init = time.perf_counter() - t0
    rng = random.Random(6)
    s = "".join(rng.choice("0123456789") for _ in range(10 ** 6)).lstrip("0")
    v = int(s)
    t0 = time.perf_counter()
    out = T.to_dec(L, v)
    t_conv = time.perf_counter() - t0
    q = rng.getrandbits(1 << 22) | (1 << ((1 << 22) - 1))
    t0 = time.perf_counter()
    x = T.recip(L, q)
    t_rec = time.perf_counter() - t0
    ok_rec = x * q <= 1 << (2 << 22) < (x + 1) * q
    # 10^7 digits straight from limbs (ntt_to_decimal), checked against GMP's mpz_get_str
    import ctypes
    u64p = ctypes.POINTER(ctypes.c_uint64)
    L.ntt_to_decimal.argtypes = [u64p, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t,
                                 ctypes.POINTER(ctypes.c_size_t)]
    limbs = O.splitmix(7, 520000)
    cap = int(64 * len(limbs) * 0.30103) + 2
    buf, nd = ctypes.create_string_buffer(cap), ctypes.c_size_t()
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        rc = L.ntt_to_decimal(limbs.ctypes.data_as(u64p), len(limbs), buf, cap, ctypes.byref(nd))
        ts.append(time.perf_counter() - t0)
        if rc:
            raise RuntimeError(f"ntt_to_decimal failed ({rc})")
    t0 = time.perf_counter()
    ref = O.gmp_to_decimal(limbs)
    t_gmp = time.perf_counter() - t0
    return {"digits": len(s), "to_decimal_s": t_conv, "exact": out == s, "cache_init_level20_s": t_init,
            "reciprocal_2p22_bits_s": t_rec, "reciprocal_exact": bool(ok_rec),
            "to_decimal_1e7": {"digits": nd.value, "s": float(np.median(ts)), "exact": buf.raw[:nd.value].decode() == ref,
                               "cpu_gmp_s": t_gmp}}


# ---- config 5: serving -----------------------------------------------------

def workload_serve(E, O):
    """Poisson arrivals of requests with operands log-uniform in [2^10, 2^28] bits,
    resident in HBM, through a server with a 64 GB budget and three slots, at
    swept arrival rates; then the same requests all at once (the service's
    capacity, which bounds what any arrival rate can be delivered).  The last
    30 requests' products are checked against the same products alone (and
    GMP where both operands are at most 2^24 bits)."""
    from cryptography_amd import engine as EN
    import _workloads as W
    rng = np.random.default_rng(55)
    src_words = 1 << 24                    # operands: windows of 128 MiB of resident random limbs
    out_words = 1 << 28                    # 2 GiB of result space, reused round-robin
    pool = EN.DevicePool(src_words + out_words + 64)
    src = pool.take(src_words)
    src_host = O.splitmix(77, src_words)
    pool.upload(src, src_host)
    out_base = pool.take(out_words)

    def requests(n, seed):
        bits = W.config5_bits(n, seed)
        items, used = [], 0
        for i in range(n):
            la, lb = int((bits[i, 0] + 63) // 64), int((bits[i, 1] + 63) // 64)
            if used + la + lb > out_words:
                used = 0
            oa = int(rng.integers(0, src_words - la))
            ob = int(rng.integers(0, src_words - lb))
            items.append((pool.ptr(src + oa), la, pool.ptr(src + ob), lb, pool.ptr(out_base + used), la + lb,
                          None, 0, oa, ob, used))
            used += la + lb
        return items

    res = {}
    last = None
    for qps in (100, 300, 1000, 3000):
        n = int(min(1200, max(300, qps)))
        items = requests(n, 1000 + qps)
        arrival = np.cumsum(rng.exponential(1.0 / qps, size=n))
        t_done, st = EN.serve_replay(EN.make_descs([it[:8] for it in items]), arrival, 64 << 30, slots=3)
        lat = (t_done - arrival) * 1e3
        res[str(qps)] = {"requests": n, "offered_per_s": qps, "products_per_s": n / st["makespan_s"],
                         "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                         "batches": st["batches"], "max_inflight_batches": st["max_inflight"]}
        last = items
    # checks: the last requests of the last sweep point (before the capacity run
    # below reuses their result space)
    ok_alone, ok_gmp, n_gmp = True, True, 0
    for (pa, la, pb, lb, pc, n_out, _, _, oa, ob, oc) in last[-30:]:
        got = pool.download(out_base + oc, n_out)
        a, b = src_host[oa:oa + la], src_host[ob:ob + lb]
        with E.DeviceBuffer(n_out) as dc:
            E.engine.check(E.lib().ntt_mul_device(dc.ptr, pa, la, pb, lb))
            ok_alone &= bool(np.array_equal(got, dc.download()))
        if max(la, lb) <= (1 << 18):
            ok_gmp &= bool(np.array_equal(got, O.gmp_mul(a, b)))
            n_gmp += 1
    # capacity: every request queued at once
    items = requests(1200, 7)
    _, st = EN.serve_replay(EN.make_descs([it[:8] for it in items]), np.zeros(len(items)), 64 << 30, slots=3)
    capacity = len(items) / st["makespan_s"]
    pool.close()
    return {"budget_gb": 64, "slots": 3, "operand_bits": "2^(10 + 18u)", "by_qps": res,
            "capacity_products_per_s": capacity, "capacity_batches": st["batches"],
            "checked_last": 30, "equal_alone": ok_alone, "equal_gmp": ok_gmp, "gmp_checked": n_gmp}


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(O, h):
    """The oracle's 3-prime CT-DIT NTT multiply (the reference's algorithm, restated)
    on the headline's own 2^30-bit operands, every transform spread over the host
    cores this GPU's share of the box gives (at most 16), measured once at full
    size; its product must equal GMP's.  GMP mpz_mul (one core) beside it."""
    a, b = h["operands"]
    threads = max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    c = O.mul_ntt3_par(a, b, threads)
    t_port = time.perf_counter() - t0
    exact = bool(np.array_equal(c, h["product"]))
    return {"value": t_port * 1e3, "unit": "ms/mul (2^30-bit)", "cores": threads, "kind": "port",
            "sample": f"one full 2^30-bit x 2^30-bit product (the headline operands, L = 2^27, 3 primes) with "
                      f"the oracle's CT-DIT NTT multiply, every transform on {threads} OpenMP threads; "
                      f"product equals GMP's: {exact}",
            "exact": exact,
            "gmp_mul_ms": h["gmp_mul_s"] * 1e3, "gmp_sqr_ms": h["gmp_sqr_s"] * 1e3,
            "gmp_note": "GMP 6.2.1 mpz_mul on one core, the same operands (limb import and export included)",
            "cpu_model": cpu_model(), "nproc": os.cpu_count()}


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
    import _oracle as O
    E.init()
    tuning = None
    if args.tuning != "none" and os.path.exists(args.tuning):
        E.lib().ntt_engine_autotune_load.argtypes = [ctypes.c_char_p]
        E.engine.check(E.lib().ntt_engine_autotune_load(args.tuning.encode()))
        tuning = os.path.relpath(args.tuning, ROOT)
    if args.tree_compare:
        from cryptography_amd import engine as EN
        progress(f"product tree of e at {args.tree_compare} digits: batched, then per product")
        print(json.dumps({"tree_compare": tree_compare(EN, O, args.tree_compare), "digits": args.tree_compare}))
        return
    progress("headline: 2^30-bit product")
    h = headline(args, E, O)
    if args.profile_only:
        print(json.dumps(h))
        return
    line = {
        "metric": "ms per 2^30-bit mul (%HBM peak)",
        "value": h["ms_per_step"], "unit": "ms/mul", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": h["ms_per_step"], "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic",
        "config": {"workload": f"2^{args.bits}-bit x 2^{args.bits}-bit product, 16-bit coefficients, "
                               f"3-prime CRT, NTT length 2^{h['logL']}, operands resident in HBM",
                   "operand_bits": h["nbits"], "transform_len": h["Llen"], "primes": K_PRIMES, "seed": [1, 2]},
        "square_ms": h["square_ms"],
        "square_roofline": h["square_roofline"],
        "host_api_ms": h.get("host_api_ms"),
        "kernels_ms": h["kernels_ms"],
        "total_event_ms": h["total_event_ms"],
        "roofline": h["roofline"],
        "tuning": {"file": tuning, "threshold_limbs": E.lib().ntt_engine_get_threshold(),
                   "tile_batch": E.lib().ntt_engine_get_tile_batch(),
                   "pass_split_2p27": pass_plan(E.lib(), 27)},
    }
    if not args.headline_only:
        wl = {}
        todo = args.workloads.split(",")
        for name, fn in (("mul22", lambda: workload_mul22(E, O)), ("batch", lambda: workload_batch(E, O)),
                         ("e", lambda: workload_e(E, O, args.e_digits)), ("serve", lambda: workload_serve(E, O)),
                         ("convert", lambda: workload_convert(E, O))):
            if name in todo:
                progress(f"workload {name}")
                wl[name] = fn()
        line["workloads"] = wl
    if not args.no_cpu:
        progress("cpu baseline")
        line["cpu_baseline"] = cpu_baseline(O, h)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
