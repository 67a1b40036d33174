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


SPEC_G = {150: 16, 160: 16, 200: 16, 240: 16, 300: 16, 320: 12, 400: 16, 480: 8, 1000: 8, 960: 4}
MODES = {"fwd_col": (2, 0), "fwd_row": (0, 0), "inv_row": (3, 0), "inv_col": (0, 1)}


def kernel_instance(cls, M):
    """Name of the pass kernel a (class, FFT length) launches, as rocprof
    lists it: the four-step split M = M1 * M2 (rows M2 a multiple of 16 when
    possible, then the most balanced), column passes of length M1, row passes
    of length M2."""
    if cls not in MODES:
        return cls
    M1, M2 = M, 1
    if M > 4096:
        cands = [(0 if (M // d) % 16 == 0 else 1, abs(d - M // d), d) for d in range(1, 4097)
                 if M % d == 0 and M // d <= 4096]
        M1 = min(cands)[2]
        M2 = M // M1
    L = M1 if cls in ("fwd_col", "inv_col") else M2
    lm, sm = MODES[cls]
    if L in SPEC_G:
        return "pass_kernel_ct<%d, %d, %d, %d>" % (lm, sm, L, SPEC_G[L])
    return "pass_kernel_rt<%d, %d>" % (lm, sm)


def build_workload(name, seed):
    from pocketsdr_amd import workloads
    if name == "l1ca":
        w = workloads.acq_l1ca(seed=seed)
        w["sig"]["name"] = "L1CA"
        return [w["buff"]], [w["sig"]], {"L1CA": w["truth"]}, \
            "acq_l1ca_12msps_1ms_x10_21bins_32prn"
    if name == "multi":
        w = workloads.acq_multi(seed=seed)
        return w["buffs"], w["sigs"], w["truth"], \
            "acq_batched_uneven_L1CA_E1B_L5Q_L6D"
    raise SystemExit("unknown workload " + name)


def bench_acq(args, sdr, orc):
    from pocketsdr_amd import sdr_func
    buffs, sigs, truth, wname = build_workload(args.workload, args.seed)
    cells = sum(s["code_fft"].shape[0] * len(s["fds"]) for s in sigs)
    res = {}
    flags = sdr_func.SEARCH_BATCHED | (sdr_func.SEARCH_GRAPH if args.graph else 0)
    plan = sdr.SearchPlan(buffs, sigs, flags)
    log("batched plan built (%d cells); warmup %d steps %d" % (cells, args.warmup, args.steps))
    t_b = time_plan(plan, args.steps, args.warmup)
    det = {}
    idx_b = []
    for i, s in enumerate(sigs):
        cn0, ix, dop = plan.result(i)
        idx_b.append(ix.copy())
        found = sorted(s["prns"][q] for q in range(len(s["prns"])) if cn0[q] >= 38.0)
        det[s["name"]] = dict(found=found, truth=sorted(t["prn"] for t in truth[s["name"]]))
    log("batched: %.3f ms/step, detections %s" % (t_b * 1e3, det))
    # per-kernel HIP events on the plan's stream over timed steps
    plan.timing(True)
    for _ in range(args.steps):
        plan.run()
    ks = {}
    for n in ["fwd_col", "fwd_row", "inv_row", "inv_col", "peak", "fill"]:
        for i, s in enumerate(sigs):
            r = plan.kernel_ms("%s@%d" % (n, i))
            if r:
                ks["%s@%s" % (n, s["name"])] = dict(avg_ms=r[0], launches=r[1], elems=r[2], bytes=r[3],
                                                    M=s["M"], cls=n)
    plan.close()
    # dominant kernel: one instantiation (kernel class x signal group)
    dom = max(ks, key=lambda k: ks[k]["avg_ms"] * ks[k]["launches"])
    d = ks[dom]
    gbs = d["bytes"] / (d["avg_ms"] * 1e-3) / 1e9
    res["roofline"] = dict(bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                           frac=round(gbs / HBM_PEAK_GBS, 4), traffic=None, kernel=dom,
                           instantiation=kernel_instance(d["cls"], d["M"]),
                           alg_bytes_per_launch=round(d["bytes"]),
                           avg_launch_ms=round(d["avg_ms"], 5))
    res["kernels"] = {k: dict(avg_ms=round(v["avg_ms"], 5), launches=v["launches"],
                              gbs=round(v["bytes"] / (v["avg_ms"] * 1e-3) / 1e9, 1))
                      for k, v in ks.items()}
    if not args.skip_plain:
        plan = sdr.SearchPlan(buffs, sigs, sdr_func.SEARCH_PLAIN)
        t_p = time_plan(plan, max(1, args.steps // 4), 1)
        same = all(np.array_equal(plan.result(i)[1], idx_b[i]) for i in range(len(sigs)))
        plan.close()
        res["plain_fp32_cells_per_s"] = round(cells / t_p, 1)
        res["batched_over_plain"] = round(t_p / t_b, 3)
        res["plain_same_indices"] = bool(same)
        log("plain: %.3f ms/step (batched is %.2fx)" % (t_p * 1e3, t_p / t_b))
    res["detections"] = det
    if orc is not None and not args.skip_cpu:
        # bounded sample: the first cpu_prns PRNs of every signal, all bins and blocks
        t0 = time.perf_counter()
        n_cells = 0
        parts = []
        for s in sigs:
            nq = min(args.cpu_prns, s["code_fft"].shape[0])
            for q in range(nq):
                P = np.zeros((len(s["fds"]), s["M"]), np.float32)
                for k in range(s["n_blk"]):
                    orc.search_code(s["code_fft"][q], s["T"], buffs[s["buff"]],
                                    s["ix0"] + k * s["blk_step"], s["fs"], s["fi"], s["fds"], P)
            n_cells += nq * len(s["fds"])
            parts.append("%s %d PRNs x %d bins x %d blocks" % (s["name"], nq, len(s["fds"]), s["n_blk"]))
        t_cpu = time.perf_counter() - t0
        # cells/s of the full workload's mix: time per cell of each signal weighted by its cells
        res["cpu_baseline"] = dict(
            value=round(n_cells / t_cpu, 2), unit="cells/s", cores=1, kind="port",
            sample="; ".join(parts) + " (%.1f s): oracle C restatement of sdr_search_code with "
                   "the reference PocketFFT (%s), one thread, no 1 ms sleeps"
                   % (t_cpu, "loaded" if orc.REF_LOADED else "absent"))
        log("cpu: %.1f cells/s over %.1f s" % (n_cells / t_cpu, t_cpu))
    cfg = dict(workload=wname, cells=cells,
               signals=[dict(name=s["name"], fs=s["fs"], T=s["T"], M=s["M"], n_prn=s["code_fft"].shape[0],
                             n_bins=len(s["fds"]), n_blk=s["n_blk"]) for s in sigs],
               path="batched" + ("+hipgraph" if args.graph else ""))
    return cells / t_b, t_b, res, cfg


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


def copy_probe(mb):
    """A plain device-to-device copy of known size (torch's copy kernel), the
    calibration of the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mb << 20
    a = torch.empty(n, dtype=torch.uint8, device="cuda")
    b = torch.empty(n, dtype=torch.uint8, device="cuda")
    a.fill_(1)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps(dict(copy_probe_bytes=n, launches=3)), flush=True)


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
    ap.add_argument("--cpu-prns", type=int, default=4)
    ap.add_argument("--workload", default="multi", choices=["multi", "l1ca"])
    ap.add_argument("--trk-epochs", type=int, default=1000)
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a torch device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
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
    value, t_step, res, cfg = bench_acq(args, sdr, orc)
    if not args.skip_trk:
        res.update(bench_trk(args, sdr, workloads))
    line = dict(metric=METRIC, value=round(value, 1), unit="cells/s", n_gpus=1, steps=args.steps,
                warmup=args.warmup, ms_per_step=round(t_step * 1e3, 4), higher_is_better=True,
                scaling="weak", vs_baseline=None, dtype="f32",
                data="synthetic (seeded cpx8 IF, 2-bit, ~25% visible PRNs, C/N0 38-50 dB-Hz)",
                config=cfg)
    line.update(res)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
