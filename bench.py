"""Benchmark of the MI355X search / correlate hot path (one GPU).

Default workload: the batched uneven search of four signal groups in one
plan (GPS L1 C/A 1 ms @ 12 Msps, Galileo E1-B 4 ms @ 12 Msps, GPS L5-Q 1 ms
@ 24 Msps, QZSS L6D 4 ms @ 24 Msps; all PRNs of each system; Doppler bins of
1/(2T) over +/-5 kHz; T_AQC = 10 ms of data per stream as pocket_acq reads it,
app/pocket_acq/pocket_acq.c:152-192; 5070 search cells), seeded synthetic
2-bit IF.  A "step" is the whole search with the IF and code spectra resident
in HBM, replayed as one hipGraph.  value = search cells / s of the precision
form (FP16 words, MFMA DFT-matrix stages, FP32 accumulation).

Also on the same line: the FP32 batched path and the plain FP32 per-PRN GPU
path (the reference dataflow: one search_code per PRN and block, unfused) on
the same workload, the acquisition decisions of each form, the tracking
correlators (64 channels at 24 Msps, 44 taps, Msps*ch, with their int8-dot
roofline), the roofline of the dominant kernel from HIP events, a per-cell
HBM model and a per-signal FLOP model of the whole step, the CPU baseline
(the oracle restatement of sdr_search_code with the reference PocketFFT, one
thread, and the same sample on a pool of host threads), whose sample also
gives the peak-power error of the timed form against the FP64 oracle, the
front end (raw unpack, array combine) and L6D FFT tracking with CSK.
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
FP16_PEAK_TFLOPS = 2500.0       # dense FP16 / BF16 MFMA
FP32_PEAK_TFLOPS = 157.3        # FP32 vector
INT8_PEAK_TOPS = 5000.0         # dense int8 MFMA, the published int8 peak


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_plan(plan, steps, warmup):
    for _ in range(warmup):
        plan.run()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.run()
    return (time.perf_counter() - t0) / steps


MODES = {"fwd_col": (2, 0), "fwd_row": (0, 0), "inv_row": (3, 0), "inv_col": (0, 1)}


# fused cell inverse kernels of the FP16 form by split (M1, M2)
CELL = {(150, 160): "cell_kernel<160, 16, 10, 150, 10, 15, 50, 80, 1024>"}


def kernel_instance(cls, M, form="fp32", tiles=None):
    """Name of the pass kernel a (class, FFT length, form) launches, as
    rocprof lists it.  FP32: the four-step split M = M1 * M2 (rows M2 a
    multiple of 16 when possible, then the most balanced), column passes of
    length M1, row passes of length M2 (the plan's tiles when given).
    FP16 / BF16: the plan's tiles (M1, M2, column shape, row shape)."""
    if cls == "cell" and tiles and tuple(tiles[:2]) in CELL:
        return CELL[tuple(tiles[:2])]
    if cls not in MODES:
        return cls
    from pocketsdr_amd import sdr_func
    if form in ("fp16", "bf16") and tiles:
        M1, M2, vc, vr = tiles
        col = cls in ("fwd_col", "inv_col")
        L, v = (M1, vc) if col else (M2, vr)
        r0, r1, r2, G, nth = sdr_func.pass_shape(col, L, v)
        lm, sm = MODES[cls]
        return "mpass_kernel<%d, %d, %d, %d, %d, %d, %d, %d, %s, %d>" % (
            lm, sm, 0 if form == "fp16" else 1, L, r0, r1, r2, G, "false" if col else "true", nth)
    M1, M2 = M, 1
    if tiles:
        M1, M2 = tiles[0], tiles[1]
    elif M > 4096:
        cands = [(0 if (M // d) % 16 == 0 else 1, abs(d - M // d), d) for d in range(1, 4097)
                 if M % d == 0 and M // d <= 4096]
        M1 = min(cands)[2]
        M2 = M // M1
    L = M1 if cls in ("fwd_col", "inv_col") else M2
    lm, sm = MODES[cls]
    g = sdr_func.fft_pass_groups(L)
    if g:
        return "pass_kernel_ct<%d, %d, %d, %d>" % (lm, sm, L, g)
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
    if name == "l1cp":
        # one GPU's share of the 8-channel array: one RF channel, full size
        w = workloads.acq_l1cp(seed=seed, n_rf=1)
        return w["buffs"], w["sigs"], w["truth"], \
            "acq_weak_L1CP_48Msps_10ms_x20_241bins_32prn_1rf_of_8"
    raise SystemExit("unknown workload " + name)


FORM_FLAGS = {"fp32": 16, "fp16": 4 | 16, "bf16": 8 | 16}    # batched forms, tuned tiles
# the fastest tilings of each workload's FFT lengths on MI355X (every split
# and workgroup shape timed over 3 x 5 steps), fixed so that repeated runs
# and their counter profiles launch the same kernels; groups without an
# entry are timed by the plan tuner (SDR_SEARCH_TUNE) at plan creation
TILES = {"multi": "h24000=150/0/0,h96000=400/0/1,h48000=300/0/0,h192000=400/0/2,"
                  "f24000=150,f96000=200,f48000=200,f192000=200",
         "l1ca": "h24000=150/1/0,f24000=150",
         "l1cp": "h960000=1000/0/0"}
FORM_DTYPE = {"fp32": "f32", "fp16": "f16", "bf16": "bf16"}


def run_form(sdr, buffs, sigs, form, graph, steps, warmup):
    """Builds the batched plan of one arithmetic form and times it; returns
    (plan, s/step, per-signal (cn0, ix, dop))."""
    from pocketsdr_amd import sdr_func
    flags = FORM_FLAGS[form] | (sdr_func.SEARCH_GRAPH if graph else 0)
    plan = sdr.SearchPlan(buffs, sigs, flags)
    t = time_plan(plan, steps, warmup)
    res = [plan.result(i) for i in range(len(sigs))]
    return plan, t, res


def decisions(sigs, res, truth):
    """Per signal: PRNs over the 38 dB-Hz threshold, the planted PRNs, and
    the planted PRNs whose peak cell sits on the true code phase (2 samples
    plus the code Doppler over the blocks) and Doppler (one bin)."""
    det = {}
    for s, (cn0, ix, dop) in zip(sigs, res):
        found = sorted(s["prns"][q] for q in range(len(s["prns"])) if cn0[q] >= 38.0)
        located = []
        for t in truth[s["name"]]:
            q = s["prns"].index(t["prn"])
            dl = (ix[q][1] / s["fs"] - t["tau"] + s["T"] / 2) % s["T"] - s["T"] / 2
            tol = 2.0 / s["fs"] + abs(t["fd"]) / t["fc"] * (s["n_blk"] + 1) * s["T"]   # + code Doppler
            if abs(dl) < tol and abs(float(s["fds"][ix[q][0]]) - t["fd"]) <= 0.5 / s["T"]:
                located.append(t["prn"])
        det[s["name"]] = dict(found=found, truth=sorted(t["prn"] for t in truth[s["name"]]),
                              located=sorted(located))
    return det


def same_decisions(sigs, ra, rb):
    """Same detections, and the same (bin, lag) for every detection."""
    for s, (ca, ia, _), (cb, ib, _) in zip(sigs, ra, rb):
        da, db = ca >= 38.0, cb >= 38.0
        if not np.array_equal(da, db) or not np.array_equal(ia[da], ib[da]):
            return False
    return True


def bench_acq(args, sdr, orc):
    from pocketsdr_amd import sdr_func
    buffs, sigs, truth, wname = build_workload(args.workload, args.seed)
    cells = sum(s["code_fft"].shape[0] * len(s["fds"]) for s in sigs)
    # counter summary of the same workload (profiles/search_<workload>_traffic.md)
    res = {}
    form = args.form
    plan, t_b, r_main = run_form(sdr, buffs, sigs, form, args.graph, args.steps, args.warmup)
    log("%s batched: %.3f ms/step (%d cells)" % (form, t_b * 1e3, cells))
    forms = [sdr_func.FORM_FP32, sdr_func.FORM_PLAIN, sdr_func.FORM_FP16, sdr_func.FORM_BF16]
    names = ["fp32", "plain", "fp16", "bf16"]
    res["form_per_signal"] = {s["name"]: names[forms.index(plan.form(i))] for i, s in enumerate(sigs)}
    tiles = {s["name"]: plan.tiles(i) for i, s in enumerate(sigs)}
    res["tiles"] = {k: dict(M1=v[0], M2=v[1], col_shape=v[2], row_shape=v[3])
                    for k, v in tiles.items() if v}
    res["tiles_source"] = ("fixed: " + os.environ["PSA_TILES"]) if os.environ.get("PSA_TILES") \
        else "plan tuner (SDR_SEARCH_TUNE)"
    res["detections"] = decisions(sigs, r_main, truth)
    # per-kernel HIP events on the plan's stream over timed steps
    plan.timing(True)
    for _ in range(args.steps):
        plan.run()
    ks = {}
    for n in ["fwd_col", "fwd_row", "inv_row", "inv_col", "cell", "peak", "fill"]:
        for i, s in enumerate(sigs):
            r = plan.kernel_ms("%s@%d" % (n, i))
            if r:
                ks["%s@%s" % (n, s["name"])] = dict(avg_ms=r[0], launches=r[1], elems=r[2], bytes=r[3],
                                                    grid=plan.kernel_grid("%s@%d" % (n, i)),
                                                    M=s["M"], cls=n, form=res["form_per_signal"][s["name"]],
                                                    tiles=tiles[s["name"]])
    # roofline of the step: the algorithmic HBM bytes of the whole step (the
    # per-cell-step model, sum over every cell-step of every signal) over the
    # step time against the HBM peak; per signal the same over the signal's
    # own kernel time
    model = cell_model(sigs, form, t_b)
    cap = model["capacity"]
    counted = step_traffic(args.workload, form)
    res["roofline"] = dict(bound="hbm", achieved=cap["achieved_GBs"], peak=HBM_PEAK_GBS, unit="GB/s",
                           frac=cap["frac"], traffic=counted["bytes_per_step"] if counted else None,
                           scope="whole step (all kernels of one search step)",
                           alg_bytes_per_step=cap["bytes_per_step"],
                           traffic_over_alg=round(counted["bytes_per_step"] / cap["bytes_per_step"], 3)
                           if counted else None,
                           traffic_source=counted["source"] if counted else None)
    per_sig = {}
    for i, sg in enumerate(sigs):
        ms = sum(v["avg_ms"] * v["launches"] for k, v in ks.items() if k.endswith("@" + sg["name"])) / args.steps
        b = cap["per_signal"][sg["name"]]
        gbs = b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        per_sig[sg["name"]] = dict(alg_bytes=round(b), kernel_ms=round(ms, 4), achieved_GBs=round(gbs, 1),
                                   frac=round(gbs / HBM_PEAK_GBS, 4))
    res["roofline_per_signal"] = per_sig
    # the kernel instantiation (as rocprof names it) with the most device
    # time, priced with its own operand bytes (X, code spectra, Z, P): a
    # per-kernel figure, not the step's algorithmic model above
    inst_tot = {}
    for k, v in ks.items():
        inst = kernel_instance(v["cls"], v["M"], v["form"], v["tiles"])
        t = inst_tot.setdefault(inst, dict(ms=0.0, launches=0, bytes=0.0, parts=[], grids={}))
        t["ms"] += v["avg_ms"] * v["launches"]
        t["launches"] += v["launches"]
        t["bytes"] += v["bytes"] * v["launches"]
        t["parts"].append(k)
    inst = max(inst_tot, key=lambda i: inst_tot[i]["ms"])
    d = inst_tot[inst]
    avg_ms = d["ms"] / d["launches"]
    opb = d["bytes"] / d["launches"]
    res["dominant_kernel"] = dict(kernel=inst, launches_of=d["parts"], avg_launch_ms=round(avg_ms, 5),
                                  operand_bytes_per_launch=round(opb),
                                  operand_GBs=round(opb / (avg_ms * 1e-3) / 1e9, 1),
                                  share_of_step=round(d["ms"] / args.steps / (t_b * 1e3), 4))
    res["cell_model"] = {k: {kk: vv for kk, vv in v.items() if kk != "per_signal"} for k, v in model.items()}
    # algorithmic FLOPs per cell-step F = 5 M log2 M (1 + 1 / n_prn) + 8 M (the
    # forward FFT shared by a signal's PRNs, one inverse per PRN, product and
    # power) over each signal's own kernel time, against the dense FP16 MFMA
    # peak (the precision form's stages) or the FP32 vector peak
    fl = {}
    for i, sg in enumerate(sigs):
        M, nq = sg["M"], sg["code_fft"].shape[0]
        F = 5.0 * M * np.log2(M) * (1.0 + 1.0 / nq) + 8.0 * M
        nstep = nq * len(sg["fds"]) * sg["n_blk"]
        ms = sum(v["avg_ms"] * v["launches"] for k, v in ks.items() if k.endswith("@" + sg["name"])) / args.steps
        tf = F * nstep / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
        pk = FP32_PEAK_TFLOPS if res["form_per_signal"][sg["name"]] in ("fp32", "plain") else FP16_PEAK_TFLOPS
        fl[sg["name"]] = dict(flop_per_cell_step=round(F), cell_steps=nstep, kernel_ms=round(ms, 4),
                              tflops=round(tf, 2), peak_tflops=pk, frac=round(tf / pk, 5))
    res["flop_model"] = fl
    res["kernels"] = {k: dict(avg_ms=round(v["avg_ms"], 5), launches=v["launches"],
                              gbs=round(v["bytes"] / (v["avg_ms"] * 1e-3) / 1e9, 1))
                      for k, v in ks.items()}
    if form != "fp32" and not args.skip_fp32:
        p32, t32, r32 = run_form(sdr, buffs, sigs, "fp32", args.graph, args.steps, args.warmup)
        p32_tiles = [p32.tiles(i) for i in range(len(sigs))]
        p32.close()
        res["fp32_batched_tiles"] = {s["name"]: list(p32_tiles[i][:2]) for i, s in enumerate(sigs)}
        res["fp32_batched_cells_per_s"] = round(cells / t32, 1)
        res["fp32_batched_ms_per_step"] = round(t32 * 1e3, 4)
        res["same_decisions_as_fp32"] = same_decisions(sigs, r_main, r32)
        log("fp32 batched: %.3f ms/step; same decisions: %s" % (t32 * 1e3, res["same_decisions_as_fp32"]))
    if not args.skip_plain:
        pp = sdr.SearchPlan(buffs, sigs, sdr_func.SEARCH_PLAIN)
        t_p = time_plan(pp, max(1, args.steps // 4), 1)
        rp = [pp.result(i) for i in range(len(sigs))]
        pp.close()
        res["plain_fp32_cells_per_s"] = round(cells / t_p, 1)
        res["batched_over_plain"] = round(t_p / t_b, 3)
        res["same_decisions_as_plain"] = same_decisions(sigs, r_main, rp)
        log("plain: %.3f ms/step (batched %s is %.2fx)" % (t_p * 1e3, form, t_p / t_b))
    if orc is not None and not args.skip_cpu:
        res["cpu_baseline"], res["cpu_baseline_threads"], res["peak_power_err_vs_fp64"] = \
            cpu_leg(args, orc, plan, buffs, sigs)
    plan.close()
    cfg = dict(workload=wname, cells=cells,
               signals=[dict(name=s["name"], fs=s["fs"], T=s["T"], M=s["M"], n_prn=s["code_fft"].shape[0],
                             n_bins=len(s["fds"]), n_blk=s["n_blk"]) for s in sigs],
               path="batched" + ("+hipgraph" if args.graph else ""), form=form)
    return cells / t_b, t_b, res, cfg


def cpu_sample(args, sigs, prn_mult=1):
    """(PRNs, bins) per signal of the CPU sample: the first prn_mult x
    cpu_prns PRNs over all bins (or cpu_bins), the bins cut in proportion
    where the sample would exceed prn_mult x cpu_budget IF samples processed
    (M x bins x blocks x PRNs; 1e9 is about 10 s of one core)."""
    out = []
    for s in sigs:
        nq = min(prn_mult * args.cpu_prns, s["code_fft"].shape[0])
        nb = min(args.cpu_bins, len(s["fds"])) if args.cpu_bins else len(s["fds"])
        out.append([nq, nb])
    work = sum(nq * nb * s["n_blk"] * s["M"] for (nq, nb), s in zip(out, sigs))
    if work > prn_mult * args.cpu_budget:
        f = prn_mult * args.cpu_budget / work
        for o in out:
            o[1] = max(1, int(o[1] * f))
    return out


def cpu_leg(args, orc, plan, buffs, sigs):
    """CPU baseline: the oracle's sdr_search_code (reference PocketFFT, one
    thread) timed on a bounded sample, the first cpu_prns PRNs of every signal
    over all bins and blocks (fewer bins for the large workloads, cpu_sample).
    Untimed, on the same sample: the FP64 oracle's power at each PRN's peak
    cell against the timed GPU form's surface."""
    t0 = time.perf_counter()
    n_cells = 0
    parts, surf = [], []
    for i, (s, (nq, nb)) in enumerate(zip(sigs, cpu_sample(args, sigs))):
        for q in range(nq):
            P = np.zeros((nb, s["M"]), np.float32)
            for k in range(s["n_blk"]):
                orc.search_code(s["code_fft"][q], s["T"], buffs[s["buff"]],
                                s["ix0"] + k * s["blk_step"], s["fs"], s["fi"], s["fds"][:nb], P)
            surf.append((i, q, P))
        n_cells += nq * nb
        parts.append("%s %d PRNs x %d bins x %d blocks" % (s["name"], nq, nb, s["n_blk"]))
    t_cpu = time.perf_counter() - t0
    log("cpu: %.1f cells/s over %.1f s" % (n_cells / t_cpu, t_cpu))
    base = dict(value=round(n_cells / t_cpu, 2), unit="cells/s", cores=1, kind="port",
                sample="; ".join(parts) + " (%.1f s): oracle C restatement of sdr_search_code with "
                       "the reference PocketFFT (%s), one thread, no 1 ms sleeps"
                       % (t_cpu, "loaded" if orc.REF_LOADED else "absent"))
    # the receiver model: one host thread per task, all cores busy - four
    # times the PRNs of the sample above split into (signal, PRN, 8-bin)
    # tasks over a thread pool (the oracle's C calls run outside the
    # interpreter lock)
    from concurrent.futures import ThreadPoolExecutor
    nth = max(1, min(16, os.cpu_count() or 1))
    tasks, mt_parts = [], []
    for i, (s, (nq, nb)) in enumerate(zip(sigs, cpu_sample(args, sigs, 4))):
        tasks += [(i, q, b0, min(nb, b0 + 8)) for q in range(nq) for b0 in range(0, nb, 8)]
        mt_parts.append("%s %d PRNs x %d bins" % (s["name"], nq, nb))

    def task(tk):
        i, q, b0, b1 = tk
        s = sigs[i]
        P = np.zeros((b1 - b0, s["M"]), np.float32)
        for k in range(s["n_blk"]):
            orc.search_code(s["code_fft"][q], s["T"], buffs[s["buff"]], s["ix0"] + k * s["blk_step"],
                            s["fs"], s["fi"], s["fds"][b0:b1], P)
        return b1 - b0

    t0 = time.perf_counter()
    with ThreadPoolExecutor(nth) as ex:
        n_mt = sum(ex.map(task, tasks))
    t_mt = time.perf_counter() - t0
    log("cpu %d threads: %.1f cells/s over %.1f s" % (nth, n_mt / t_mt, t_mt))
    base_mt = dict(value=round(n_mt / t_mt, 2), unit="cells/s", cores=nth, kind="port",
                   sample="%s over all blocks as (signal, PRN, 8-bin) tasks on %d host "
                          "threads (%.1f s)" % (", ".join(mt_parts), nth, t_mt))
    # the FP64 check costs a full-length FP64 search per block and a transfer
    # of the PRN's surface: at most 5e7 / (M blocks) PRNs per signal (all of
    # the sample for the default workload, 2 for L1C-P)
    worst, n_chk = 0.0, 0
    for i, q, P in surf:
        s = sigs[i]
        if q >= max(1, int(5e7 / (s["M"] * s["n_blk"]))):
            continue
        n_chk += 1
        b, lag = np.unravel_index(int(np.argmax(P[:, :s["N"]])), (P.shape[0], s["N"]))
        ref = np.zeros((1, s["M"]), np.float64)
        for k in range(s["n_blk"]):
            orc.search_code_f64(s["code_fft"][q], s["T"], buffs[s["buff"]], s["ix0"] + k * s["blk_step"],
                                s["fs"], s["fi"], s["fds"][b:b + 1], ref)
        g = plan.power(i, q)
        worst = max(worst, abs(float(g[b, lag]) - ref[0, lag]) / ref[0, lag])
    return base, base_mt, dict(max_rel=float("%.3e" % worst), cells=n_chk,
                               note="peak cell of each sampled PRN, timed form vs FP64 oracle")


def step_traffic(workload, form):
    """Counted HBM bytes of one search step from the committed FETCH_SIZE /
    WRITE_SIZE summary of the same workload and form
    (profiles/search_<workload>_<form>_step_traffic.md, written from
    `bench.py --profile-step` runs): the line `| step total | <bytes> |`.
    None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "search_%s_%s_step_traffic.md" % (workload, form))
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.split("|")]
        if len(cells) > 3 and cells[1] == "step total":
            return dict(bytes_per_step=int(float(cells[2])), source="profiles/" + os.path.basename(path))
    return None


def bench_l1cp_full(args, sdr):
    """The weak-signal wide search at its full size on this GPU: GPS L1C-P on
    all 8 RF channels of the array (32 PRNs x 241 bins x 20 blocks of 10 ms at
    48 Msps each, M = 960000; 61,696 cells, 118.5 GB of power surfaces), FP16
    precision form, seeded synthetic 3-bit IF.  Roofline by the per-cell-step
    model (p = 2: M b_w = 3.84 MB does not fit one CU)."""
    from pocketsdr_amd import workloads
    w = workloads.acq_l1cp(seed=9)
    sigs = w["sigs"]
    cells = sum(s["code_fft"].shape[0] * len(s["fds"]) for s in sigs)
    tiles = os.environ.get("PSA_TILES", "")
    os.environ["PSA_TILES"] = (tiles + "," if tiles else "") + TILES["l1cp"]
    plan = sdr.SearchPlan(w["buffs"], sigs, FORM_FLAGS["fp16"])
    os.environ["PSA_TILES"] = tiles
    steps = max(1, args.l1cp_steps)
    t = time_plan(plan, steps, 1)
    res = [plan.result(i) for i in range(len(sigs))]
    plan.close()
    model = cell_model(sigs, "fp16", t)["capacity"]
    located = [len(set(decisions([s], [r], {s["name"]: w["truth"][s["name"]]})[s["name"]]["located"]))
               for s, r in zip(sigs, res)]
    log("l1cp full size: %.1f ms/step (%d cells)" % (t * 1e3, cells))
    return dict(workload="acq_weak_L1CP_48Msps_10ms_x20_241bins_32prn_8rf", cells=cells, steps=steps, warmup=1,
                ms_per_step=round(t * 1e3, 2), cells_per_s=round(cells / t, 1), form="fp16",
                roofline=dict(bound="hbm", achieved=model["achieved_GBs"], peak=HBM_PEAK_GBS, unit="GB/s",
                              frac=model["frac"], alg_bytes_per_step=model["bytes_per_step"]),
                located_per_rf=located, planted=len(w["truth"][sigs[0]["name"]]))


def bench_raw(args, sdr, orc):
    """Front end (write_buff): one second of the 8-channel L1 array's raw
    stream (Pocket SDR FE 8CH, RAW32, 2-bit I/Q, 48 Msps) unpacked to the 8
    cpx8 IF buffers, then combined into 4 array beams (sdr_arch_combine),
    inputs and outputs resident in HBM.  Algorithmic bytes per sample: unpack
    4 raw + 8 cpx8, combine 8 + 4 cpx8.  CPU: the oracle's write_buff on the
    first 1/10 s, one thread."""
    import torch
    fmt, N, fs = sdr.FMT_RAW32, 48_000_000, 48e6
    luts = [sdr.raw_gen_lut(fmt, k, 2, 2) for k in range(8)]
    plan = sdr.RawPlan(fmt, luts, [0] * 8)
    g = torch.Generator(device="cuda").manual_seed(args.seed)
    draw = torch.randint(0, 256, (N * 4,), dtype=torch.uint8, device="cuda", generator=g)
    dout = torch.empty((8, N), dtype=torch.uint8, device="cuda")
    ptrs = [dout[k].data_ptr() for k in range(8)]
    for _ in range(2):
        plan.unpack_dev(draw.data_ptr(), N, 0, ptrs)
    ks = []
    for i in range(max(3, args.steps)):
        plan.unpack_dev(draw.data_ptr(), N, i, ptrs)
        ks.append(plan.kernel_ms())
    kms = sorted(ks)[len(ks) // 2]
    gbs = 12.0 * N / (kms * 1e-3) / 1e9
    res = dict(format="RAW32 (8 ch, 2-bit I/Q)", samples=N, kernel_ms=round(kms, 4),
               msps_per_ch=round(N / (kms * 1e-3) / 1e6, 1), alg_bytes=12 * N,
               roofline=dict(bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                             frac=round(gbs / HBM_PEAK_GBS, 4)))
    # array channels: 4 beams combined from the 8 RF channels
    w = np.random.default_rng(args.seed).integers(-300, 300, (4, 8, 2)).astype(np.int16)
    beams = torch.empty((4, N), dtype=torch.uint8, device="cuda")
    bp = [beams[m].data_ptr() for m in range(4)]
    cms = sorted(sdr.arch_combine_dev(ptrs, N, w, bp) for _ in range(max(3, args.steps)))
    cms = cms[len(cms) // 2]
    cgbs = 12.0 * N / (cms * 1e-3) / 1e9
    res["arch_combine"] = dict(rf_channels=8, beams=4, samples=N, kernel_ms=round(cms, 4),
                               alg_bytes=12 * N,
                               roofline=dict(bound="hbm", achieved=round(cgbs, 1), peak=HBM_PEAK_GBS,
                                             unit="GB/s", frac=round(cgbs / HBM_PEAK_GBS, 4)))
    if orc is not None:
        n = N // 10
        raw = draw[:n * 4].cpu().numpy()
        t0 = time.perf_counter()
        orc.write_buff(fmt, raw, n, 0, [0] * 8, luts)
        dt = time.perf_counter() - t0
        res["cpu_msps_per_ch"] = round(n / dt / 1e6, 1)
        res["cpu_sample"] = "%d samples x 8 ch, oracle write_buff, one thread (%.2f s)" % (n, dt)
    plan.close()
    return res


def bench_trk(args, sdr, orc, w):
    """Tracking correlators (64 channels x 1000 epochs at 24 Msps, 44 taps):
    Msps*ch, the int8 roofline of the correlator kernel against the published
    dense int8 peak, and the CPU baseline: the oracle's open-loop tracking of
    src/sdr_ch.c:694-740 (sdr_mix_carr + sdr_corr_std per epoch, AVX2-exact
    integer dot products) on the first channels, one thread."""
    plan = sdr.TrkPlan(w["buff"], w["chans"], w["pos"], w["n_epoch"])
    t = time_plan(plan, max(1, args.steps // 2), 1)
    plan.timing(True)
    plan.run()
    ks = plan.kernel_ms("trk_corr")
    plan.close()
    n_ch, N = len(w["chans"]), w["chans"][0]["N"]
    msps_ch = n_ch * N * w["n_epoch"] / t / 1e6
    res = dict(correlator_msps_ch=round(msps_ch, 1), correlator_ms_per_step=round(t * 1e3, 3),
               correlator_config=dict(n_ch=n_ch, fs=24e6, n_epoch=w["n_epoch"], n_pos=len(w["pos"])),
               correlator_kernel_ms=round(ks[0], 4) if ks else None)
    if ks:
        # 2 n N int8 MACs per channel-epoch (I and Q), 2 ops each
        tops = 2.0 * 2 * len(w["pos"]) * N * n_ch * w["n_epoch"] / (ks[0] * 1e-3) / 1e12
        res["correlator_roofline"] = dict(bound="int8", achieved=round(tops, 2), peak=INT8_PEAK_TOPS,
                                          unit="TOPS", frac=round(tops / INT8_PEAK_TOPS, 5),
                                          note="v_dot4 int8 products on the vector ALUs; the peak is "
                                               "the published dense int8 MFMA rate")
    if orc is not None and not args.skip_cpu:
        nc = max(1, min(args.cpu_trk_channels, n_ch))
        t0 = time.perf_counter()
        for ch in w["chans"][:nc]:
            orc.trk_run(ch, ch["code"], w["buff"], w["n_epoch"], w["pos"])
        dt = time.perf_counter() - t0
        res["correlator_cpu_baseline"] = dict(
            value=round(nc * N * w["n_epoch"] / dt / 1e6, 2), unit="Msps*ch", cores=1, kind="port",
            sample="%d channels x %d epochs (%.1f s): oracle C restatement of track_sig's open loop "
                   "(sdr_mix_carr + sdr_corr_std, %d taps), one thread" % (nc, w["n_epoch"], dt, len(w["pos"])))
    return res


def bench_dropin(args, sdr, orc, w):
    """Per-call time of the drop-in entry points a receiver links
    (host buffers in and out, every call on the calling thread's stream):
    sdr_mix_carr + sdr_corr_std at N = 24000 with 44 taps (one channel-epoch
    of track_sig), sdr_search_code for L1 C/A at 12 Msps (21 bins, one
    block), and the oracle's per-call time of the same calls (one thread)."""
    ch = w["chans"][0]
    N = ch["N"]
    reps = max(20, args.steps * 4)
    out = {}

    def per_call(fn, n):
        fn()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        return (time.perf_counter() - t0) / n * 1e6

    iq = orc.mix_carr(w["buff"], 0, N, ch["fs"], ch["fi"] + ch["fd0"], ch["phi0"])
    out["corr_std"] = dict(
        N=N, taps=len(w["pos"]),
        gpu_us=round(per_call(lambda: sdr.corr_std(iq, ch["code"], N, ch["coff0"], w["pos"]), reps), 1),
        mix_carr_gpu_us=round(per_call(lambda: sdr.mix_carr(w["buff"], 0, N, ch["fs"], ch["fd0"], 0.3), reps), 1),
        cpu_us=round(per_call(lambda: orc.corr_std(iq, ch["code"], N, ch["coff0"], w["pos"]), reps), 1),
        mix_carr_cpu_us=round(per_call(lambda: orc.mix_carr(w["buff"], 0, N, ch["fs"], ch["fd0"], 0.3), reps), 1))
    from pocketsdr_amd import workloads
    a = workloads.acq_l1ca(seed=args.seed, n_prn=1, n_vis=1, t_int=1e-3)
    sg = a["sig"]
    cf = sg["code_fft"][0]
    out["search_code"] = dict(
        M=sg["M"], bins=len(sg["fds"]),
        gpu_us=round(per_call(lambda: sdr.search_code(cf, sg["T"], a["buff"], 0, sg["fs"], sg["fi"],
                                                       sg["fds"]), reps), 1),
        cpu_us=round(per_call(lambda: orc.search_code(cf, sg["T"], a["buff"], 0, sg["fs"], sg["fi"],
                                                       sg["fds"]), max(3, reps // 10)), 1))
    out["note"] = ("host pointers in and out (PCIe copies in every GPU call); CPU: the oracle's C "
                   "restatement, one thread")
    return out


L6_CLASSES = ("state", "fwd_col", "row", "fwd_row", "inv_row", "inv_col", "csk")


def bench_l6(args, sdr, orc, workloads):
    """QZSS L6D FFT tracking with CSK: n_ch channels at 24 Msps, 4 ms epochs
    (N = M = 96000), one second of IF (250 epochs) per run, IF and code
    spectra resident in HBM.  Algorithmic bytes per (channel, epoch) item:
    fwd column N IF bytes + 8M written, fused row pass 16M (read, write),
    inv column 8M read + the stored lag window (16 win, |lag| < win =
    255 R + max |pos| + 3, R = N / 10230), CSK 2 x 511 x 16 B; the code
    spectra (SDR_N_CODES per channel, 8M each) count once per run.  CPU: the
    oracle's L6 tracking loop (PocketFFT, one thread) over the epochs of one
    channel."""
    w = workloads.trk_l6(seed=args.seed + 2, fs=24e6, n_ch=args.l6_channels, n_epoch=args.l6_epochs)
    # FFT split M1 x M2 = 200 x 480: the measured best of the splits of 96000
    # for this workload (300 x 320, the default, is 7% slower)
    os.environ.setdefault("PSA_L6_SPLIT", "200")
    plan = sdr.L6Plan(w["buff"], w["chans"], w["pos"], w["n_epoch"])
    t = time_plan(plan, max(1, args.steps // 2), 1)
    syms, _ = plan.result()
    ok, seen = True, set()
    for c, ch in enumerate(w["chans"]):
        if ch["prn"] in w["truth"] and ch["prn"] not in seen:
            seen.add(ch["prn"])
            ok &= syms[c].tolist() == w["truth"][ch["prn"]]["syms"][:w["n_epoch"]].tolist()
    plan.timing(True)
    plan.run()
    ks = {k: plan.kernel_ms(k) for k in L6_CLASSES}
    plan.close()
    n_ch, N, ne = len(w["chans"]), w["chans"][0]["N"], w["n_epoch"]
    items = n_ch * ne
    win = int(np.floor(255 * N / (w["len_code"] // 2) + np.abs(w["pos"]).max())) + 3
    row = 16.0 * N if ks.get("row") else 32.0 * N
    per_item = N + 8.0 * N + row + 8.0 * N + 16.0 * win + 2 * 511 * 16
    alg = items * per_item + n_ch * 10 * 8.0 * N
    gbs = alg / t / 1e9
    res = dict(config=dict(n_ch=n_ch, fs=24e6, N=N, n_epoch=ne, n_pos=len(w["pos"]),
                           fft_split=os.environ["PSA_L6_SPLIT"],
                           signal="L6D (20460-entry TDM code, CSK 8 bit / 4 ms)"),
               ms_per_run=round(t * 1e3, 3), epochs_ch_per_s=round(items / t, 1),
               msps_ch=round(items * N / t / 1e6, 1), realtime_x=round(ne * 4e-3 / t, 1),
               symbols_decoded=bool(ok and len(seen) > 0),
               kernel_ms={k: round(v[0], 4) for k, v in ks.items() if v},
               alg_bytes_per_item=int(per_item), alg_bytes_spectra=n_ch * 10 * 8 * N,
               roofline=dict(bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                             frac=round(gbs / HBM_PEAK_GBS, 4)))
    # the same channels in the closed loop (sdr_l6_plan_new_loop): epochs in
    # order, the channels of one epoch in one batch of passes
    loops = [dict(len_code=w["len_code"], costas=0, cn0=40.0) for _ in w["chans"]]
    lp = sdr.L6Plan(w["buff"], w["chans"], w["pos"], w["n_epoch"], loops=loops)
    tl = time_plan(lp, max(1, args.steps // 4), 1)
    syms, _ = lp.result()
    st = lp.state()
    lp.timing(True)
    lp.run()
    lks = {k: lp.kernel_ms(k) for k in L6_CLASSES + ("loop",)}
    lp.close()
    ok, seen = True, set()
    for c, ch in enumerate(w["chans"]):
        if ch["prn"] in w["truth"] and ch["prn"] not in seen:
            seen.add(ch["prn"])
            ok &= syms[c].tolist() == w["truth"][ch["prn"]]["syms"][:w["n_epoch"]].tolist()
            ok &= not st[c, :, 5].any()
    res["closed_loop"] = dict(ms_per_run=round(tl * 1e3, 3), epochs_ch_per_s=round(items / tl, 1),
                              realtime_x=round(ne * 4e-3 / tl, 1), symbols_decoded=bool(ok and len(seen) > 0),
                              launches_per_epoch=5,
                              kernel_ms={k: round(v[0], 4) for k, v in lks.items() if v})
    if orc is not None:
        ch = w["chans"][0]
        n_e = ne
        t0 = time.perf_counter()
        orc.trk_l6_run(ch, w["len_code"], ch["code_fft"], w["buff"], n_e, w["pos"])
        dt = time.perf_counter() - t0
        res["cpu_epochs_ch_per_s"] = round(n_e / dt, 1)
        res["cpu_sample"] = "%d epochs of one channel, oracle L6 loop, one thread (%.2f s)" % (n_e, dt)
    return res


def bench_trk_loop(args, sdr, w):
    """Closed-loop tracking (FLL / PLL / DLL / C/N0 per epoch, one persistent
    workgroup per channel) on the correlator workload: 64 channels x 1000
    epochs at 24 Msps, 44 taps, every run from the acquisition states."""
    loops = [dict(len_code=1023, costas=1, cn0=40.0) for _ in w["chans"]]
    plan = sdr.TrkPlan(w["buff"], w["chans"], w["pos"], w["n_epoch"], loops=loops)
    t = time_plan(plan, max(1, args.steps // 4), 1)
    st = plan.state()
    plan.close()
    n_ch, N, ne = len(w["chans"]), w["chans"][0]["N"], w["n_epoch"]
    vis = {s["prn"] for s in w["truth"]}
    held = [bool(st[c, -1, 5] == 0) for c, ch in enumerate(w["chans"]) if ch["prn"] in vis]
    return dict(config=dict(n_ch=n_ch, fs=24e6, n_epoch=ne, n_pos=len(w["pos"])),
                ms_per_run=round(t * 1e3, 3), msps_ch=round(n_ch * N * ne / t / 1e6, 1),
                realtime_x=round(ne * 1e-3 / t, 1), satellites_held=sum(held), satellites=len(held))


def bench_trk_loop_sec(args, sdr, workloads):
    """Closed loop with the secondary-code sync and wipe-off: GPS L5-Q (NH20)
    on 32 channels x 2000 epochs at 12 Msps, 4 satellites in the IF started
    20 Hz and 0.1 chip off, 12 taps; the sync starts at 1.5 s."""
    w = workloads.trk_loop_sig("L5Q", seed=args.seed + 4, fs=12e6, n_ch=32, n_vis=4, n_epoch=2000,
                               nposx=8)
    plan = sdr.TrkPlan(w["buff"], w["chans"], w["pos"], w["n_epoch"], loops=w["loops"])
    t = time_plan(plan, max(1, args.steps // 4), 1)
    st = plan.state()
    P = plan.prompt()
    plan.close()
    n_ch, N, ne = len(w["chans"]), w["chans"][0]["N"], w["n_epoch"]
    vis = [c for c, ch in enumerate(w["chans"]) if ch["prn"] in w["truth"]]
    synced = [bool((P[c, 1520:].real > 0).all() and st[c, -1, 5] == 0) for c in vis]
    return dict(config=dict(signal="L5Q", n_ch=n_ch, fs=12e6, n_epoch=ne, n_pos=len(w["pos"]),
                            sec_code="NH20"),
                ms_per_run=round(t * 1e3, 3), msps_ch=round(n_ch * N * ne / t / 1e6, 1),
                realtime_x=round(ne * 1e-3 / t, 1), satellites_synced=sum(synced), satellites=len(vis))


ONCHIP_BYTES = 672 * 1024      # one CU: 160 KiB LDS + 512 KiB VGPR/AGPR


def cell_model(sigs, form, t_step):
    """Whole-step HBM bytes of the per-cell model: for one (signal, PRN, bin,
    block) cell-step B = M / (n_prn n_bins) (IF bytes) + b_cf M / (n_bins
    n_blk) (code spectrum) + 4 N / n_blk (final power surface) + (p - 1) 4 M
    b_w (one HBM round trip of each FFT when it does not fit one CU), with
    b_w = b_cf = 4 (FP16/BF16) or 8 (FP32).  'capacity': p = 1 when M b_w fits
    one CU's LDS + registers (a lower bound no multi-pass FFT reaches);
    'four_step': p = 2 for every length (the passes built here)."""
    bw = 8 if form == "fp32" else 4
    out = {}
    for name, rule in (("capacity", None), ("four_step", 2)):
        tot = 0.0
        per = {}
        for s in sigs:
            M, N = s["M"], s["M"] // 2
            nprn, nbin, nblk = s["code_fft"].shape[0], len(s["fds"]), s["n_blk"]
            p = rule or (1 if M * bw <= ONCHIP_BYTES else 2)
            B = M / (nprn * nbin) + bw * M / (nbin * nblk) + 4.0 * N / nblk + (p - 1) * 4.0 * M * bw
            per[s["name"]] = B * nprn * nbin * nblk
            tot += per[s["name"]]
        gbs = tot / t_step / 1e9
        out[name] = dict(bytes_per_step=round(tot), achieved_GBs=round(gbs, 1),
                         frac=round(gbs / HBM_PEAK_GBS, 4), per_signal=per)
    return out


def tiles_arg(spec):
    """PSA_TILES string from --tiles: given as is, or '@FILE' = the tilings a
    previous bench line used (tiles, fp32_batched_tiles)."""
    if not spec.startswith("@"):
        return spec
    line = json.loads(open(spec[1:]).read().strip().splitlines()[-1])
    sigs = {s["name"]: s["M"] for s in line["config"]["signals"]}
    ent = ["h%d=%d/%d/%d" % (sigs[n], t["M1"], t["col_shape"], t["row_shape"])
           for n, t in line.get("tiles", line.get("tuned_tiles", {})).items()]
    ent += ["f%d=%d" % (sigs[n], t[0]) for n, t in (line.get("fp32_batched_tiles") or {}).items() if t]
    return ",".join(ent)


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


def profile_step(args):
    """Only the timed form's plan of the workload and `steps` search steps
    (no graph, nothing else): the run the counter summaries of one step are
    taken from, dividing the step kernels' counts by `steps`."""
    import pocketsdr_amd  # noqa: F401
    from pocketsdr_amd import sdr_func as sdr
    sdr.func_init()
    buffs, sigs, _, wname = build_workload(args.workload, args.seed)
    plan = sdr.SearchPlan(buffs, sigs, FORM_FLAGS[args.form])
    for _ in range(args.steps):
        plan.run()
    plan.close()
    print(json.dumps(dict(profile_steps=args.steps, workload=wname, form=args.form)), flush=True)


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
    ap.add_argument("--skip-raw", action="store_true")
    ap.add_argument("--skip-l6", action="store_true")
    ap.add_argument("--skip-l1cp", action="store_true",
                    help="skip the full-size 8-channel L1C-P search leg")
    ap.add_argument("--l1cp-steps", type=int, default=2)
    ap.add_argument("--l6-channels", type=int, default=16)
    ap.add_argument("--l6-epochs", type=int, default=250)
    ap.add_argument("--cpu-prns", type=int, default=16)
    ap.add_argument("--workload", default="multi", choices=["multi", "l1ca", "l1cp"])
    ap.add_argument("--cpu-budget", type=float, default=1e9,
                    help="CPU sample bound in IF samples processed (about 10 s of one core at 1e9)")
    ap.add_argument("--cpu-bins", type=int, default=0,
                    help="CPU sample: only the first this many Doppler bins (0: all)")
    ap.add_argument("--form", default="fp16", choices=["fp16", "bf16", "fp32"],
                    help="arithmetic form of the timed batched search")
    ap.add_argument("--skip-fp32", action="store_true")
    ap.add_argument("--trk-epochs", type=int, default=1000)
    ap.add_argument("--cpu-trk-channels", type=int, default=4,
                    help="CPU baseline of the tracking correlators: this many channels over all epochs")
    ap.add_argument("--profile-step", action="store_true",
                    help="only build the timed plan and run --steps search steps (counter profiles)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a torch device copy of this many MiB (counter calibration)")
    ap.add_argument("--tiles", default="auto",
                    help="plan tilings: 'auto' (the workload's measured best, else tuned), 'tune' "
                         "(the plan tuner times every tiling), a PSA_TILES string "
                         "(e.g. h96000=300/0/0,f96000=200) or '@FILE' (those of a bench JSON line)")
    args = ap.parse_args()
    if args.tiles != "tune":
        os.environ["PSA_TILES"] = TILES.get(args.workload, "") if args.tiles == "auto" else tiles_arg(args.tiles)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.profile_step:
        profile_step(args)
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
        w_trk = workloads.trk_l1ca(seed=args.seed + 1, n_epoch=args.trk_epochs)
        res.update(bench_trk(args, sdr, orc, w_trk))
        if orc is not None:
            res["dropin_per_call"] = bench_dropin(args, sdr, orc, w_trk)
        res["closed_loop_tracking"] = bench_trk_loop(args, sdr, w_trk)
        res["closed_loop_secondary_code"] = bench_trk_loop_sec(args, sdr, workloads)
    if not args.skip_l1cp:
        res["l1cp_full"] = bench_l1cp_full(args, sdr)
    if not args.skip_raw:
        r = bench_raw(args, sdr, orc)
        res["arch_combine"] = r.pop("arch_combine")
        res["raw_unpack"] = r
    if not args.skip_l6:
        res["l6_tracking"] = bench_l6(args, sdr, orc, workloads)
    line = dict(metric=METRIC, value=round(value, 1), unit="cells/s", n_gpus=1, steps=args.steps,
                warmup=args.warmup, ms_per_step=round(t_step * 1e3, 4), higher_is_better=True,
                scaling="weak", vs_baseline=None, dtype=FORM_DTYPE[args.form],
                data="synthetic (seeded cpx8 IF, 2-bit, ~25% visible PRNs, C/N0 38-50 dB-Hz)",
                config=cfg)
    line.update(res)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
