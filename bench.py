"""Benchmark of the ZODA column encode on one MI355X.

Workload: plain encode of a 2048 x 2048 square of 512-byte shares, held in
HBM: 2048 rows x 262144 columns of uniform BabyBear u32 (2 GiB) -> 4096 x
262144 (4 GiB).  A step is one encode launch over the whole square; the input
is device-resident before the timed region starts.

Prints one JSON line with the throughput (decimal Gb/s of payload,
side^2 * 512 bytes), the roofline of the encode kernel (algorithmic bytes
12 * s * C per launch over its HIP-event-timed duration, against 8 TB/s), and
the CPU oracle timed on a bounded sample of the same workload.

    python bench.py [--steps K] [--warmup W] [--side S] [--no-e2e] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "zoda-cu_amd"))

METRIC = "encode Gb/s + %HBM roofline (2048² square); batched vs plain-loop p50/p99"
PEAK_HBM_GBS = 8000.0
SHARE_BYTES = 512
P = 2013265921  # BabyBear


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=20,
                    help="untimed steps right before the timed ones (the clock settles within about 12)")
    ap.add_argument("--side", type=int, default=2048)
    ap.add_argument("--seed", type=int, default=2048)
    ap.add_argument("--input", choices=("uniform", "ramp"), default="uniform",
                    help="encode input: seeded uniform BabyBear, or the reference benchmark's ramp "
                         "(row*cols+col) mod p (src/benchmarks/utils.rs:81-84)")
    ap.add_argument("--no-e2e", action="store_true", help="skip the pinned-host end-to-end timing")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-cols", type=int, default=65536,
                    help="columns in the CPU sample (about 20 s of CPU-thread work on 16 threads)")
    ap.add_argument("--no-batched", action="store_true", help="skip the 96-square batched comparison")
    ap.add_argument("--no-served", action="store_true", help="skip the served-trace comparison")
    ap.add_argument("--served-requests", type=int, default=3000,
                    help="requests per served trace (p99 rests on 1 %% of them)")
    ap.add_argument("--no-rlc", action="store_true", help="skip the RLC leg")
    ap.add_argument("--no-decode", action="store_true", help="skip the erasure-decode leg")
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip the reference's own window: zoda_hip_intt + zoda_hip_ntt on the pinned square")
    ap.add_argument("--no-col", action="store_true", help="skip the column-major device encode leg")
    ap.add_argument("--no-4096", action="store_true", help="skip the 4096^2 square leg")
    ap.add_argument("--only-encode", action="store_true",
                    help="time the encode only (no end-to-end, batched, served or CPU legs); for profiling")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="instead of benchmarking, run 3 device copies of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.only_encode:
        args.no_e2e = args.no_cpu = args.no_batched = args.no_served = args.no_rlc = args.no_decode = True
        args.no_dropin = args.no_col = args.no_4096 = True
    return args


def traffic_from_profile(side, layout: str = ""):
    """HBM bytes per launch of the encode kernel measured by rocprofv3 --pmc
    (FETCH_SIZE and WRITE_SIZE, each scaled by what the counter reads for a
    copy kernel of known size), from the committed summary's
    `hbm_bytes_per_launch:` line (per encode call for the column-major layout,
    `layout="col_"`); null when no measurement exists."""
    import re
    path = os.path.join(ROOT, "profiles", f"encode_{layout}{side}_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        m = re.search(r"^hbm_bytes_per_launch:\s*(\d+)", f.read(), re.M)
    return int(m.group(1)) if m else None


def cpu_baseline(side: int, ncols: int, seed: int):
    """The oracle (a C restatement of src/ntt/cpu.rs and the encode of
    src/benchmarks/utils.rs) on `ncols` columns of the same square over the
    process's CPU share (at most 16 threads: one GPU's share of the host),
    plus a sixteenth of the sample on one core, the form of the reference's own
    single-threaded CPU encode (src/zoda/babybear.rs:41-44)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle
    import zoda_hip
    data = zoda_hip.uniform_numpy(side * ncols, seed).reshape(side, ncols)
    host_cores = os.cpu_count() or 1
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = host_cores
    threads = max(1, min(allowed, 16))
    t0 = time.perf_counter()
    oracle.encode(data, threads=threads)
    dt = time.perf_counter() - t0
    one = max(ncols // 16, 64)
    t0 = time.perf_counter()
    oracle.encode(np.ascontiguousarray(data[:, :one]), threads=1)
    dt1 = time.perf_counter() - t0
    gbps = side * ncols * 4 * 8 / 1e9 / dt
    gbps1 = side * one * 4 * 8 / 1e9 / dt1
    return {
        "value": round(gbps, 4),
        "unit": "Gb/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{ncols} of {128 * side} columns of the {side}^2 square on {threads} threads "
                  f"({dt:.2f} s); Gb/s of input data per second, the same unit as the GPU line",
        "host_cores": host_cores,
        "cores_allowed": allowed,
        "threads_note": "threads = the cores this process may run on, capped at 16 (one GPU's share of the host)",
        "single_core_value": round(gbps1, 4),
        "single_core_sample": f"{one} columns on 1 thread ({dt1:.2f} s)",
        "cpu_model": _cpu_model(),
    }


def batched_vs_plain(zoda_hip, torch, seed: int, reps: int = 7):
    """96 squares of sides drawn over {32..512} (30/30/20/15/5 %), side x 128 side
    each: one persistent launch replayed from a hipGraph against a plain loop of
    one encode launch per square on one stream.  Both timed with HIP events."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sides = rng.choice([32, 64, 128, 256, 512], size=96, p=[0.30, 0.30, 0.20, 0.15, 0.05]).tolist()
    ins, outs, specs = [], [], []
    for i, sd in enumerate(sides):
        t_in = torch.empty((sd, 128 * sd), dtype=torch.int32, device="cuda")
        zoda_hip.fill_uniform(t_in, seed=seed + i)
        t_out = torch.empty((2 * sd, 128 * sd), dtype=torch.int32, device="cuda")
        ins.append(t_in)
        outs.append(t_out)
        specs.append((t_in.data_ptr(), t_out.data_ptr(), sd, 128 * sd, sd))
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream
    torch.cuda.synchronize()
    tuned = zoda_hip.autotune_batch(specs, iters=3)

    def plain():
        for (di, do, k, c, n) in specs:
            zoda_hip.encode_device(di, k, c, n, do, 0, sptr)

    plan = zoda_hip.BatchPlan(specs)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ms = []
        for _ in range(reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            torch.cuda.synchronize()
            ms.append(a.elapsed_time(b))
        ms.sort()
        return ms[len(ms) // 2]

    t_plain = timed(plain)
    t_batch = timed(lambda: plan.launch(sptr))
    plan.destroy()
    # the two forms agree on the last run
    ok = True
    for t_out, (di, do, k, c, n) in zip(outs, specs):
        ref = torch.empty_like(t_out)
        zoda_hip.encode_device(di, k, c, n, ref.data_ptr(), 0, sptr)
        torch.cuda.synchronize()
        ok = ok and bool(torch.equal(ref, t_out))
    payload = sum(sd * sd * SHARE_BYTES for sd in sides)
    return {
        "workload": "96 squares, sides {32,64,128,256,512} at 30/30/20/15/5 %, side x 128*side u32, device-resident",
        "payload_bytes": payload,
        "plain_loop_ms": round(t_plain, 4),
        "batched_graph_ms": round(t_batch, 4),
        "plain_loop_gbps": round(payload * 8 / 1e9 / (t_plain / 1e3), 2),
        "batched_gbps": round(payload * 8 / 1e9 / (t_batch / 1e3), 2),
        "speedup": round(t_plain / t_batch, 3),
        "tuning_table": {
            str(k): {
                "choice": _candidate_name(v["candidates"][v["choice"]], v["choice"]),
                "times_us": {_candidate_name(c, i): t for i, (c, t) in enumerate(zip(v["candidates"], v["times_us"]))},
            } for k, v in tuned.items()},
        "tuning_note": "per side of the batch (largest first): the four work-queue plans and the side's own "
                       "fused launches, each timed as the whole batch (best of 3 after a warm-up, other sides at "
                       "their current entries, starting from the untuned defaults); the fastest is kept in the "
                       "device's table (zoda_hip_tune_save / _load persist it)",
        "bit_exact_vs_plain": ok,
    }


def _candidate_name(c: dict, i: int) -> str:
    d = "-".join(str(x) for x in c["digits"])
    if c["kind"] == "direct":
        return f"direct(W{c['W']},{c['threads']}t,{d})"
    return f"q{i}(W{c['W']},{d},{c['twiddles']})"


def served_vs_plain(zoda_hip, torch, seed: int, count: int, rates=(2000, 8000), slots: int = 3,
                    checked_per_trace: int = 24):
    """Poisson arrivals of squares with log-uniform sides {64..1024} (512 B
    shares, device-resident inputs drawn from 12 read-only buffers per side),
    replayed in real time against the continuous-batching server (HBM budget
    200 GB, 3 launch slots) and against a plain one-square-at-a-time loop.
    The server takes each request's output from its device pool at admission
    and returns it at completion (the plain loop writes one block per size, one
    request at a time); a seeded sample of requests per trace writes to
    buffers of its own instead, checked after the trace against the same
    square encoded alone."""
    import numpy as np
    sides_all = [64, 128, 256, 512, 1024]
    budget = 200 * 10**9
    pool = {}
    for sd in sides_all:
        pool[sd] = []
        for j in range(12):
            t_in = torch.empty((sd, 128 * sd), dtype=torch.int32, device="cuda")
            zoda_hip.fill_uniform(t_in, seed=seed + 100 * sd + j)
            pool[sd].append(t_in)
    torch.cuda.synchronize()
    # Untimed warm-up of both paths with one request of every side, so that
    # neither pays first-use costs (code object loading, twiddle tables) in
    # its timed trace: a running server keeps those resident.
    warm = [(pool[sd][0].data_ptr(), 0, sd, 128 * sd, sd) for sd in sides_all]
    for mode in (0, 1):
        zoda_hip.serve_trace(warm, np.zeros(len(warm)), hbm_budget=budget, slots=slots, mode=mode)
    out = {}
    all_ok = True
    for rate in rates:
        rng = np.random.default_rng(seed + rate)
        sides = rng.choice(sides_all, size=count).tolist()
        gaps = rng.exponential(1e6 / rate, size=count)
        arrivals = np.cumsum(gaps) - gaps[0]
        checked = sorted(rng.choice(count, size=min(checked_per_trace, count), replace=False).tolist())
        payload = sum(sd * sd * SHARE_BYTES for sd in sides)
        res = {}
        for mode, name in ((0, "served"), (1, "plain_loop")):
            own = {i: torch.zeros((2 * sides[i], 128 * sides[i]), dtype=torch.int32, device="cuda")
                   for i in checked}
            specs = []
            for i, sd in enumerate(sides):
                t_in = pool[sd][i % len(pool[sd])]
                specs.append((t_in.data_ptr(), own[i].data_ptr() if i in own else 0, sd, 128 * sd, sd))
            torch.cuda.synchronize()
            lat, st = zoda_hip.serve_trace(specs, arrivals, hbm_budget=budget, slots=slots, mode=mode)
            ok = True
            for i in checked:
                ref = zoda_hip.encode_tensor(pool[sides[i]][i % 12])
                torch.cuda.synchronize()
                ok = ok and bool(torch.equal(ref, own[i]))
                del ref
            all_ok = all_ok and ok
            del own
            torch.cuda.empty_cache()
            res[name] = {
                "p50_us": round(float(np.percentile(lat, 50)), 1),
                "p99_us": round(float(np.percentile(lat, 99)), 1),
                "mean_us": round(float(lat.mean()), 1),
                "gbps": round(payload * 8 / 1e9 / (st["makespan_us"] / 1e6), 2),
                "launches": st["batches"],
                "max_batch": st["max_batch"],
                "output_pool_peak_bytes": st["pool_peak_bytes"],
                "outputs_checked": len(checked),
                "outputs_match_alone": ok,
            }
        out[f"{rate}_req_per_s"] = res
    return {
        "workload": f"{count} requests, Poisson arrivals, sides log-uniform {{64..1024}}, 512 B shares, "
                    f"HBM budget 200 GB, {slots} slots; outputs from the server's pool (plain loop: one block "
                    f"per size); Gb/s = payload over first arrival to last completion",
        "outputs_checked": all_ok,
        **out,
    }


def measured_copy_gbps(torch, like, reps: int = 5):
    """Device-to-device copy bandwidth on a buffer the size of the input (read +
    write bytes over the HIP-event time, p50): the practical ceiling the
    roofline is also reported against."""
    src = torch.empty_like(like)
    dst = torch.empty_like(like)
    stream = torch.cuda.current_stream()
    dst.copy_(src)
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        dst.copy_(src)
        b.record(stream)
        torch.cuda.synchronize()
        ms.append(a.elapsed_time(b))
    ms.sort()
    nbytes = 2 * src.numel() * src.element_size()
    del src, dst
    return round(nbytes / (ms[len(ms) // 2] / 1e3) / 1e9, 1)


def rlc_leg(zoda_hip, torch, t_out, reps: int = 10):
    """RLC (src/zoda/rlc.rs:32-40) of every row of the encoded square with the
    column coefficients of a data root (derived on the device, rlc.rs:48-66):
    one streaming pass over the 2s x C output, HIP events, p50 of `reps`."""
    rows, cols = t_out.shape
    stream = torch.cuda.current_stream()
    coeffs = zoda_hip.rlc_coefficients("0" * 64, cols)  # warm-up (first launch loads the kernel)
    torch.cuda.synchronize()
    cms = []
    for _ in range(5):  # the call as a caller sees it (allocation and synchronisation included)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        coeffs = zoda_hip.rlc_coefficients("0" * 64, cols)
        e1.record(stream)
        torch.cuda.synchronize()
        cms.append(e0.elapsed_time(e1))
    coeff_ms = sorted(cms)[len(cms) // 2]
    zoda_hip.rlc_rows(t_out, coeffs)
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        zoda_hip.rlc_rows(t_out, coeffs)
        b.record(stream)
        torch.cuda.synchronize()
        ms.append(a.elapsed_time(b))
    ms.sort()
    t = ms[len(ms) // 2]
    alg = rows * cols * 4 + cols * 16
    return {
        "workload": f"RLC of all {rows} rows of the {rows}x{cols} encoded square, 4 limbs",
        "ms_p50": round(t, 4),
        "roofline": {"bound": "hbm", "achieved": round(alg / (t / 1e3) / 1e9, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(alg / (t / 1e3) / 1e9 / PEAK_HBM_GBS, 4)},
        "coefficients_ms": round(coeff_ms, 3),
    }


def decode_leg(zoda_hip, torch, t_in, t_out, seed: int, reps: int = 3):
    """Erasure decode of the square from a random half of its 2s rows (the
    data back, bit-exact), wall clock around the synchronous call (it includes
    the set-up of the erasure pattern), median of `reps`."""
    import numpy as np
    side, cols = t_in.shape
    known = np.random.default_rng(seed).choice(2 * side, size=side, replace=False)
    out = torch.empty_like(t_in)
    zoda_hip.decode_tensor(t_out, known, out=out)
    torch.cuda.synchronize()
    ok = bool(torch.equal(out, t_in))
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        zoda_hip.decode_tensor(t_out, known, out=out)
        times.append(time.perf_counter() - t0)
    times.sort()
    t = times[len(times) // 2]
    return {
        "workload": f"decode the {side}x{cols} data from {side} random rows of {2 * side}",
        "ms_p50": round(t * 1e3, 2),
        "payload_gbps": round(side * side * SHARE_BYTES * 8 / 1e9 / t, 1),
        "bit_exact": ok,
    }


def time_events(torch, fn, steps: int, warmup: int = 10):
    """HIP-event time of each of `steps` calls of fn on the current stream
    (after `warmup` untimed calls): (avg ms, p50 ms)."""
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for a, b in ev:
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    ms = sorted(a.elapsed_time(b) for a, b in ev)
    return sum(ms) / len(ms), ms[len(ms) // 2]


def roofline_of(alg_bytes: int, ms: float, traffic=None):
    gbs = alg_bytes / (ms / 1e3) / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(gbs / PEAK_HBM_GBS, 4), "traffic": traffic, "algorithmic_bytes": alg_bytes,
            "kernel_ms_avg": round(ms, 4)}


def spot_check(torch, t_in, t_out_rows, pick):
    """Columns `pick` of the encoded square (rows x cols tensor, or a callable
    returning the 2s x len(pick) block) against the oracle."""
    import numpy as np
    import oracle
    idx = torch.tensor(pick, device=t_in.device)
    d_s = t_in[:, idx].contiguous().cpu().numpy().view(np.uint32)
    o_s = t_out_rows(idx)
    return bool(np.array_equal(o_s, oracle.encode(d_s)))


def dropin_leg(zoda_hip, t_in, t_out, reps: int = 5):
    """The reference's own "GPU xform" window (src/benchmarks/utils.rs:137-145):
    intt_cuda(square, n_k, stride n_kn) then ntt_cuda(square, n_kn, stride
    n_kn) on the pinned column-major square (src/ntt/cuda.rs:31-67), here
    zoda_hip_intt + zoda_hip_ntt.  Host memory in and out (PCIe inside the
    window, as in the reference), wall clock, median of `reps` after one
    warm-up; the warm-up's result is checked against the device encode."""
    import numpy as np
    side, cols = t_in.shape
    nkn = 2 * side
    sq = zoda_hip.PinnedSquare(nkn, cols)
    try:
        a = sq.array.reshape(cols, nkn)
        a[:, :side] = t_in.cpu().numpy().view(np.uint32).T
        rk, rkn = zoda_hip.roots_of_unity(side), zoda_hip.roots_of_unity(nkn)
        ptr = sq.array.ctypes.data

        def xform():
            zoda_hip._check(zoda_hip.lib.zoda_hip_intt(ptr, rk.ctypes.data, side, cols, nkn), "zoda_hip_intt")
            zoda_hip._check(zoda_hip.lib.zoda_hip_ntt(ptr, rkn.ctypes.data, nkn, cols, nkn), "zoda_hip_ntt")

        xform()
        pick = [0, 1, cols // 3, cols - 1]
        ok = all(np.array_equal(a[c], t_out[:, c].cpu().numpy().view(np.uint32)) for c in pick)
        times = []
        for _ in range(reps):  # in place: later runs transform the previous result (same work)
            t0 = time.perf_counter()
            xform()
            times.append(time.perf_counter() - t0)
        times.sort()
        t = times[len(times) // 2]
        return {
            "window": "zoda_hip_intt(n_k, stride n_kn) + zoda_hip_ntt(n_kn, stride n_kn) on the pinned "
                      f"column-major {cols} x {nkn} square; host memory in and out; wall clock, median of {reps}",
            "ms_p50": round(t * 1e3, 2),
            "gbps": round(side * side * SHARE_BYTES * 8 / 1e9 / t, 2),
            "matches_device_encode": ok,
        }
    finally:
        sq.free()


def dropin_device_leg(zoda_hip, torch, t_in, t_out, steps: int = 5):
    """The same window on device memory: zoda_hip_ntt_device inverse (n_k,
    stride n_kn) then forward (n_kn, stride n_kn) over the column-major square
    held in HBM -- the drop-in transform kernel itself, without PCIe.  HIP
    events on the launch stream, average of `steps` after one run whose
    result is checked against the device encode and five untimed ones.  Algorithmic bytes: 8 per
    transformed value, 8 (n_k + n_kn) cols per window."""
    import numpy as np
    side, cols = t_in.shape
    nkn = 2 * side
    sq = torch.zeros((cols, nkn), dtype=torch.int32, device=t_in.device)
    sq[:, :side] = t_in.t()
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream

    def xform():
        zoda_hip.ntt_device(sq.data_ptr(), side, cols, nkn, True, sptr)
        zoda_hip.ntt_device(sq.data_ptr(), nkn, cols, nkn, False, sptr)

    xform()
    torch.cuda.synchronize()
    pick = [0, 1, cols // 3, cols - 1]
    ok = all(np.array_equal(sq[c].cpu().numpy(), t_out[:, c].cpu().numpy()) for c in pick)
    for _ in range(5):  # untimed, after the check's idle gap (the clock settles)
        xform()
    st = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    en = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    for i in range(steps):  # in place: later runs transform the previous result (same work)
        st[i].record(stream)
        xform()
        en[i].record(stream)
    torch.cuda.synchronize()
    ms = sorted(a.elapsed_time(b) for a, b in zip(st, en))
    avg = sum(ms) / len(ms)
    del sq
    torch.cuda.empty_cache()
    return {
        "window": f"zoda_hip_ntt_device inverse (n {side}, stride {nkn}) + forward (n {nkn}, stride {nkn}) on the "
                  f"column-major {cols} x {nkn} square in HBM; HIP events, average of {steps}",
        "ms_avg": round(avg, 4),
        "ms_p50": round(ms[len(ms) // 2], 4),
        "gbps": round(side * side * SHARE_BYTES * 8 / 1e9 / (avg / 1e3), 2),
        "roofline": roofline_of(8 * (side + nkn) * cols, avg),
        "matches_device_encode": ok,
    }


def column_major_leg(zoda_hip, torch, t_in, steps: int):
    """The device-resident encode in the reference's output layout
    (ZODA_LAYOUT_COL: out[c * n_kn + r], src/ntt/cuda.rs:31-67)."""
    side, cols = t_in.shape
    nkn = 2 * side
    t_col = torch.empty((cols, nkn), dtype=torch.int32, device=t_in.device)
    sptr = torch.cuda.current_stream().cuda_stream
    fn = lambda: zoda_hip.encode_device(t_in.data_ptr(), side, cols, side, t_col.data_ptr(), 1, sptr)
    avg, p50 = time_events(torch, fn, steps)
    import numpy as np
    ok = spot_check(torch, t_in, lambda idx: t_col[idx, :].t().contiguous().cpu().numpy().view(np.uint32),
                    [0, 1, cols // 3, cols - 1])
    del t_col
    return {
        "workload": f"{side}x{side} square ({side} x {cols} u32) encoded into the column-major {cols} x {nkn} "
                    "square, device-resident",
        "ms_avg": round(avg, 4), "ms_p50": round(p50, 4),
        "gbps": round(side * side * SHARE_BYTES * 8 / 1e9 / (avg / 1e3), 2),
        "roofline": roofline_of(12 * side * cols, avg, traffic_from_profile(side, "col_")),
        "spot_check": ok,
    }


def side_leg(zoda_hip, torch, side: int, seed: int, steps: int, e2e: bool):
    """A second square size end to end: device-resident encode (HIP events,
    roofline) and the pinned host-to-host encode."""
    import numpy as np
    cols = 128 * side
    dev = torch.device("cuda:0")
    t_in = torch.empty((side, cols), dtype=torch.int32, device=dev)
    t_out = torch.empty((2 * side, cols), dtype=torch.int32, device=dev)
    zoda_hip.fill_uniform(t_in, seed=seed)
    sptr = torch.cuda.current_stream().cuda_stream
    fn = lambda: zoda_hip.encode_device(t_in.data_ptr(), side, cols, side, t_out.data_ptr(), 0, sptr)
    avg, p50 = time_events(torch, fn, steps)
    pick = [0, 1, cols // 3, cols - 1]
    ok = spot_check(torch, t_in, lambda idx: t_out[:, idx].contiguous().cpu().numpy().view(np.uint32), pick)
    res = {
        "workload": f"plain ZODA encode of a {side}x{side} square of 512 B shares ({side} x {cols} u32 -> "
                    f"{2 * side} x {cols}), device-resident",
        "ms_avg": round(avg, 4), "ms_p50": round(p50, 4), "steps": steps,
        "gbps": round(side * side * SHARE_BYTES * 8 / 1e9 / (avg / 1e3), 2),
        "roofline": roofline_of(12 * side * cols, avg, traffic_from_profile(side)),
        "spot_check": ok,
    }
    if e2e:
        pin_in = pin_out = None
        try:
            pin_in = zoda_hip.PinnedSquare(side * cols, 1)
            pin_out = zoda_hip.PinnedSquare(2 * side * cols, 1)
            a_in = pin_in.array.reshape(side, cols)
            a_out = pin_out.array.reshape(2 * side, cols)
            a_in[:] = t_in.cpu().numpy().view(np.uint32)
            zoda_hip.encode_host(a_in, out=a_out)
            times = []
            for _ in range(3):
                t0 = time.perf_counter()
                zoda_hip.encode_host(a_in, out=a_out)
                times.append(time.perf_counter() - t0)
            times.sort()
            ok_e2e = bool(np.array_equal(a_out[:, pick], t_out[:, pick].cpu().numpy().view(np.uint32)))
            res["end_to_end_pinned"] = {
                "gbps": round(side * side * SHARE_BYTES * 8 / 1e9 / times[1], 2),
                "ms_p50": round(times[1] * 1e3, 2),
                "window": "pinned row-major host in -> pinned host out, wall clock, median of 3",
                "matches_device_encode": ok_e2e}
        except Exception as ex:  # report, do not fail the line
            res["end_to_end_pinned"] = {"error": str(ex)}
        finally:
            for pbuf in (pin_in, pin_out):
                if pbuf is not None:
                    pbuf.free()
    del t_in, t_out
    torch.cuda.empty_cache()
    return res


def _cpu_model():
    """Model name, family/model/stepping, sockets and logical CPUs from /proc/cpuinfo."""
    info = {}
    sockets, logical = set(), 0
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    continue
                key, val = (x.strip() for x in line.split(":", 1))
                if key == "processor":
                    logical += 1
                elif key == "physical id":
                    sockets.add(val)
                elif key in ("model name", "cpu family", "model", "stepping", "cpu MHz", "cache size",
                             "cpu cores") and key not in info:
                    info[key] = val
    except OSError:
        return "unknown"
    return (f"{info.get('model name', 'unknown')} (family {info.get('cpu family', '?')}, model "
            f"{info.get('model', '?')}, stepping {info.get('stepping', '?')}, {info.get('cache size', '?')} "
            f"cache, {info.get('cpu cores', '?')} cores per socket, {len(sockets) or 1} socket(s), "
            f"{logical} logical CPUs)")


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the encode shards by column slabs with no exchange, "
              f"and multi-GPU runs are not part of this build (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import zoda_hip

    if args.copy_calibration:
        n = args.copy_calibration << 18   # int32 elements in MIB MiB
        src = torch.ones(n, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_each_way": n * 4, "copies": 3}))
        return

    side = args.side
    cols = 128 * side
    rows_out = 2 * side
    dev = torch.device("cuda:0")
    t_in = torch.empty((side, cols), dtype=torch.int32, device=dev)
    t_out = torch.empty((rows_out, cols), dtype=torch.int32, device=dev)
    if args.input == "ramp":
        for r0 in range(0, side, 256):  # (row * cols + col) mod p, built on the device in slabs
            r1 = min(side, r0 + 256)
            idx = torch.arange(r0 * cols, r1 * cols, device=dev, dtype=torch.int64)
            t_in[r0:r1].copy_((idx % P).to(torch.int32).view(r1 - r0, cols))
            del idx
    else:
        zoda_hip.fill_uniform(t_in, seed=args.seed)
    stream = torch.cuda.current_stream(dev)
    sptr = stream.cuda_stream

    def step():
        zoda_hip.encode_device(t_in.data_ptr(), side, cols, side, t_out.data_ptr(), 0, sptr)

    # spot check of one encode's output (full parity lives in tests/), then
    # the warm-up steps right before the timed ones: after an idle gap the
    # GPU's clock takes about a dozen 2048^2 encodes to settle (2.27 ms for
    # the first, 2.02 ms from the twelfth on)
    step()
    torch.cuda.synchronize()
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    pick = [0, 1, cols // 3, cols - 1]
    idx = torch.tensor(pick, device=dev)
    d_s = t_in[:, idx].contiguous().cpu().numpy().view(np.uint32)
    o_s = t_out[:, idx].contiguous().cpu().numpy().view(np.uint32)
    if not np.array_equal(o_s, oracle.encode(d_s)):
        print("bench: output mismatch against the oracle", file=sys.stderr)
        sys.exit(1)
    for _ in range(args.warmup):
        step()

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        step()
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = sorted(s.elapsed_time(e) for s, e in zip(starts, ends))
    avg_ms = sum(kern_ms) / len(kern_ms)
    p50_ms = kern_ms[len(kern_ms) // 2]
    ms_per_step = wall * 1e3 / args.steps

    payload = side * side * SHARE_BYTES
    alg_bytes = 12 * side * cols          # 4 B read + 8 B written per input element
    value = payload * 8 / 1e9 / (ms_per_step / 1e3)
    achieved = alg_bytes / (avg_ms / 1e3) / 1e9
    traffic = traffic_from_profile(side)
    copy_gbps = measured_copy_gbps(torch, t_in)

    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Gb/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": ("synthetic: the reference benchmark's ramp (row*cols+col) mod p, device-resident"
                 if args.input == "ramp" else
                 "synthetic: seeded uniform BabyBear u32 (zoda_hip_fill_uniform), device-resident"),
        "config": {
            "workload": f"plain ZODA encode of a {side}x{side} square of 512 B shares, device-resident",
            "side": side, "rows_in": side, "cols": cols, "rows_out": rows_out,
            "field": "BabyBear p=2013265921 (exact; FP64-FMA lazy arithmetic)",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes": alg_bytes,
            "kernel_ms_avg": round(avg_ms, 4),
            "kernel_ms_p50": round(p50_ms, 4),
            "copy_gbps": copy_gbps,
            "frac_of_copy": round(achieved / copy_gbps, 4) if copy_gbps else None,
        },
    }

    if not args.no_e2e:
        e2e = {}
        try:
            pin_in = zoda_hip.PinnedSquare(side * cols, 1)
            pin_out = zoda_hip.PinnedSquare(rows_out * cols, 1)
            a_in = pin_in.array.reshape(side, cols)
            a_out = pin_out.array.reshape(rows_out, cols)
            a_in[:] = t_in.cpu().numpy().view(np.uint32)
            zoda_hip.encode_host(a_in, out=a_out)
            times = []
            for _ in range(3):
                t0 = time.perf_counter()
                zoda_hip.encode_host(a_in, out=a_out)
                times.append(time.perf_counter() - t0)
            times.sort()
            ok = bool(np.array_equal(a_out[:, pick], o_s))
            e2e = {"gbps": round(payload * 8 / 1e9 / times[1], 2), "ms_p50": round(times[1] * 1e3, 2),
                   "window": "pinned row-major host in -> pinned host out, wall clock, median of 3",
                   "spot_check": ok}
            pin_in.free()
            pin_out.free()
        except Exception as ex:  # report, do not fail the device-resident line
            e2e = {"error": str(ex)}
        result["end_to_end_pinned"] = e2e

    if not args.no_col:
        result["encode_col_device"] = column_major_leg(zoda_hip, torch, t_in, max(3, args.steps // 2))
        torch.cuda.empty_cache()
    if not args.no_dropin:
        try:
            result["gpu_xform_dropin"] = dropin_leg(zoda_hip, t_in, t_out)
        except Exception as ex:  # report, do not fail the device-resident line
            result["gpu_xform_dropin"] = {"error": str(ex)}
        result["gpu_xform_device"] = dropin_device_leg(zoda_hip, torch, t_in, t_out)
    if not args.no_rlc:
        result["rlc"] = rlc_leg(zoda_hip, torch, t_out)
    if not args.no_decode:
        result["decode"] = decode_leg(zoda_hip, torch, t_in, t_out, args.seed)

    del t_in, t_out
    torch.cuda.empty_cache()
    if not args.no_4096 and side != 4096:
        result["side_4096"] = side_leg(zoda_hip, torch, 4096, args.seed, 5, not args.no_e2e)
    if not args.no_batched:
        result["batched"] = batched_vs_plain(zoda_hip, torch, args.seed)
        torch.cuda.empty_cache()
    if not args.no_served:
        result["served"] = served_vs_plain(zoda_hip, torch, args.seed, args.served_requests)
        torch.cuda.empty_cache()

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(side, args.cpu_cols, args.seed)
    else:
        result["cpu_baseline"] = None

    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
