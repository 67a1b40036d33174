"""Benchmark of the MI355X search / correlate hot path (one GPU).

Default workload: GPS L1 C/A acquisition as pocket_acq runs it
(app/pocket_acq/pocket_acq.c:152-192): cpx8 IQ at 12 Msps, 1 ms coherent
(N = 12000, FFT length M = 24000), 10 non-coherent blocks, 21 Doppler bins
(+/-5 kHz), 32 PRNs, seeded synthetic IF with 8 visible satellites.  A "step"
is the whole search (all 672 cells) with the IF and code spectra resident in
HBM.  value = cells / s.

Also reported on the same line: the plain FP32 per-PRN GPU path (reference
dataflow, one search_code per PRN and block, unfused), the tracking
correlators (64 channels at 24 Msps, 44 taps, Msps*ch), the roofline of the
dominant kernel from HIP events, and the CPU baseline (the oracle restatement
of sdr_search_code with the reference PocketFFT, single thread).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "acq search cells/s (batched+MFMA vs plain FP32, 1 GPU) + correlator Msps*ch"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_plan(plan, steps, warmup):
    for _ in range(warmup):
        plan.run()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.run()
    return (time.perf_counter() - t0) / steps


def kernel_stats(plan, names):
    out = {}
    for n in names:
        r = plan.kernel_ms(n)
        if r:
            out[n] = dict(avg_ms=r[0], launches=r[1], elems=r[2])
    return out


def search_kernel_bytes(name, elems, M, N, n_bin, n_prn):
    """Algorithmic HBM bytes of one launch of a pass kernel processing `elems`
    complex points (elems / M transforms): compulsory reads + writes."""
    items = elems / M
    if name == "inv_row":      # read X (shared by PRNs) + Cf, write Z
        return elems * 8 + (items / n_prn) * M * 8 + min(items, n_prn) * M * 8
    if name == "inv_col":      # read Z, read-modify-write N lags of P
        return elems * 8 + items * N * 8
    if name == "fwd_col":      # read IF bytes (shared by bins), write tmp
        return M * 1 + elems * 8
    if name == "fwd_row":
        return elems * 16
    return elems * 16


def bench_acq(args, sdr, workloads, orc):
    from pocketsdr_amd import sdr_func
    w = workloads.acq_l1ca(seed=args.seed)
    s = w["sig"]
    n_prn, n_bin = s["code_fft"].shape[0], len(s["fds"])
    cells = n_prn * n_bin
    res = {}
    flags = sdr_func.SEARCH_BATCHED | (sdr_func.SEARCH_GRAPH if args.graph else 0)
    plan = sdr.SearchPlan([w["buff"]], [s], flags)
    log("batched plan built; warmup", args.warmup, "steps", args.steps)
    t_b = time_plan(plan, args.steps, args.warmup)
    cn0, ix, dop = plan.result(0)
    found = sorted(s["prns"][q] for q in range(n_prn) if cn0[q] >= 38.0)
    truth = sorted(t["prn"] for t in w["truth"])
    log("batched: %.3f ms/step, detections %s (truth %s)" % (t_b * 1e3, found, truth))
    # per-kernel HIP-event timing over timed steps on the plan's stream
    plan.timing(True)
    for _ in range(args.steps):
        plan.run()
    ks = kernel_stats(plan, ["fwd_col", "fwd_row", "inv_row", "inv_col", "peak", "fill"])
    plan.close()
    dom = max(ks, key=lambda k: ks[k]["avg_ms"] * ks[k]["launches"])
    kb = search_kernel_bytes(dom, ks[dom]["elems"], s["M"], s["N"], n_bin, n_prn)
    gbs = kb / (ks[dom]["avg_ms"] * 1e-3) / 1e9
    res["roofline"] = dict(bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                           frac=round(gbs / HBM_PEAK_GBS, 4), traffic=None, kernel=dom,
                           alg_bytes_per_launch=kb, avg_launch_ms=ks[dom]["avg_ms"])
    res["kernels_ms"] = {k: round(v["avg_ms"], 4) for k, v in ks.items()}
    # plain FP32 per-PRN path (the 5x target denominator)
    if not args.skip_plain:
        plan = sdr.SearchPlan([w["buff"]], [s], sdr_func.SEARCH_PLAIN)
        t_p = time_plan(plan, max(1, args.steps // 2), 1)
        cn0p, ixp, _ = plan.result(0)
        plan.close()
        res["plain_fp32_cells_per_s"] = cells / t_p
        res["batched_over_plain"] = t_p / t_b
        res["plain_same_indices"] = bool(np.array_equal(ix, ixp))
        log("plain: %.3f ms/step (batched is %.2fx)" % (t_p * 1e3, t_p / t_b))
    res["detections"] = dict(found=found, truth=truth)
    # CPU baseline: oracle sdr_search_code (reference PocketFFT), 1 thread
    if orc is not None and not args.skip_cpu:
        n_cpu_prn = min(n_prn, args.cpu_prns)
        t0 = time.perf_counter()
        for q in range(n_cpu_prn):
            P = np.zeros((n_bin, s["M"]), np.float32)
            for k in range(s["n_blk"]):
                orc.search_code(s["code_fft"][q], s["T"], w["buff"], k * s["blk_step"], s["fs"],
                                s["fi"], s["fds"], P)
        t_cpu = time.perf_counter() - t0
        res["cpu_baseline"] = dict(
            value=round(n_cpu_prn * n_bin / t_cpu, 2), unit="cells/s", cores=1,
            kind="port",
            sample="%d PRNs x %d bins x %d blocks of the same workload, oracle C restatement "
                   "of sdr_search_code with the reference PocketFFT (%s), no 1 ms sleeps"
                   % (n_cpu_prn, n_bin, s["n_blk"], "loaded" if orc.REF_LOADED else "absent"))
        log("cpu: %.1f cells/s" % (n_cpu_prn * n_bin / t_cpu))
    return cells / t_b, t_b, res, dict(workload="acq_l1ca_12msps_1ms_x10_21bins_32prn",
                                       fs=s["fs"], N=s["N"], M=s["M"], n_prn=n_prn,
                                       n_bins=n_bin, n_blk=s["n_blk"], cells=cells,
                                       path="batched" + ("+hipgraph" if args.graph else ""))


def bench_trk(args, sdr, workloads):
    w = workloads.trk_l1ca(seed=args.seed + 1, n_epoch=args.trk_epochs)
    plan = sdr.TrkPlan(w["buff"], w["chans"], w["pos"], w["n_epoch"])
    t = time_plan(plan, max(1, args.steps // 2), 1)
    plan.timing(True)
    plan.run()
    ks = plan.kernel_ms("trk_corr")
    plan.close()
    n_ch, N = len(w["chans"]), w["chans"][0]["N"]
    msps_ch = n_ch * N * w["n_epoch"] / t / 1e6
    return dict(correlator_msps_ch=round(msps_ch, 1), correlator_ms_per_step=round(t * 1e3, 3),
                correlator_config=dict(n_ch=n_ch, fs=24e6, n_epoch=w["n_epoch"], n_pos=len(w["pos"])),
                correlator_kernel_ms=round(ks[0], 4) if ks else None)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-trk", action="store_true")
    ap.add_argument("--cpu-prns", type=int, default=32)
    ap.add_argument("--trk-epochs", type=int, default=1000)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs on a single MI355X "
              "(--gpus %d is not supported)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import pocketsdr_amd  # noqa: F401  (fails loudly if the HIP library is missing)
    from pocketsdr_amd import sdr_func as sdr, workloads
    sdr.func_init()
    orc = None
    if not args.skip_cpu:
        from oracle import oracle as orc
    value, t_step, res, cfg = bench_acq(args, sdr, workloads, orc)
    if not args.skip_trk:
        res.update(bench_trk(args, sdr, workloads))
    line = dict(metric=METRIC, value=round(value, 1), unit="cells/s", n_gpus=1, steps=args.steps,
                warmup=args.warmup, ms_per_step=round(t_step * 1e3, 4), higher_is_better=True,
                scaling="weak", vs_baseline=None, dtype="f32",
                data="synthetic (seeded cpx8 IF, 2-bit, 8 visible PRNs, C/N0 38-50 dB-Hz)",
                config=cfg)
    line.update(res)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
