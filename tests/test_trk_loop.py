"""Closed-loop tracking (src/sdr_ch.c:488-624, 694-763): FLL, PLL, DLL, C/N0
and the loss decision around the standard correlators.

CPU tests pin the oracle's loop with what the physics fixes: channels
started 40 Hz and 0.2 chip off a synthetic satellite pull in to its Doppler
and code phase and report its C/N0; a channel with no signal is declared
lost; before the first frequency update (epochs 0-1) the loop's correlations
equal the open-loop tracking restatement's.  GPU tests compare the
persistent-workgroup loop with the oracle: bit-identical until the first
discriminator (atan, log10: the GPU's and the host's libm may differ in the
last bit), then within stated tolerances, with the same lock and loss
decisions.

Secondary code and bump-jump (src/sdr_ch.c:453-486, 551-572): GPS L5-Q
satellites carrying the NH20 code are synchronised from T_NPULLIN = 1.5 s,
after which the wiped prompt keeps one sign (before, it follows NH20);
Galileo E1-B channels started on a BOC(1,1) side peak (half a chip off)
jump to the main peak at the first C/N0 window and stay there, while
without the VE / VL taps they hold the side peak.  The GPU loop runs the
same cases against the oracle.
"""
import numpy as np
import pytest

CHIP = 1.0 / 1.023e6


def loop_case(seed=4, n_epoch=2500, n_ch=3, n_vis=2, cn0=45.0, nposx=0):
    from pocketsdr_amd import workloads
    w = workloads.trk_l1ca(seed=seed, fs=4e6, n_ch=n_ch, n_epoch=n_epoch, n_vis=n_vis, nposx=nposx,
                           cn0=(cn0, cn0))
    truth = {s["prn"]: s for s in w["truth"]}
    chans = []
    for ch in w["chans"]:
        ch = dict(ch)
        tr = truth.get(ch["prn"])
        if tr is not None:                 # acquisition errors
            ch["fd0"] = tr["fd"] + 40.0
            ch["coff0"] = tr["tau"] + 0.2 * CHIP
        ch["phi0"] = 0.0
        chans.append(ch)
    loops = [dict(len_code=1023, costas=1, cn0=40.0) for _ in chans]
    return w, chans, loops, w["pos"], truth


def truth_track(tr, e, T=1e-3):
    t = (e + 1) * T
    fd = tr["fd"] + tr["drift"] * t
    coff = tr["tau"] - (tr["fd"] * t + 0.5 * tr["drift"] * t * t) / tr["fc"]
    return fd, coff


def test_oracle_loop_pull_in_and_loss(orc):
    w, chans, loops, pos, truth = loop_case()
    for c, ch in enumerate(chans):
        st, corr, P = orc.trkl_run(ch, 1023, 1, 40.0, ch["code"], w["buff"], w["n_epoch"], pos)
        tr = truth.get(ch["prn"])
        if tr is None:
            assert st[-1, 5] == 1                      # lost
            continue
        assert not st[:, 5].any() and st[-1, 4] == w["n_epoch"]
        for e in (999, 2499):                           # end of the FLL, end of the run
            fd, coff = truth_track(tr, e)
            assert abs(st[e, 0] - fd) < 2.0, (c, e, st[e, 0] - fd)
        fd, coff = truth_track(tr, 2499)
        assert abs(st[-1, 1] - coff) < 0.05 * CHIP, (c, (st[-1, 1] - coff) / CHIP)
        assert abs(st[-1, 3] - 45.0) < 3.0, st[-1, 3]


def test_oracle_loop_first_epochs_are_open_loop(orc):
    """Epochs 0 and 1 precede any frequency update (the FLL needs two
    prompts), so they equal the open-loop restatement at constant Doppler."""
    w, chans, loops, pos, truth = loop_case(n_epoch=4)
    for ch in chans:
        st, corr, P = orc.trkl_run(ch, 1023, 1, 40.0, ch["code"], w["buff"], 4, pos)
        ol = dict(ch, drift=0.0)
        rc, rC, rs = orc.trk_run(ol, ch["code"], w["buff"], 4, pos)
        assert np.array_equal(corr[:2].view(np.uint32), rc[:2].view(np.uint32))
        rP = orc.trk_stitch(rC, ch["N"])
        assert np.array_equal(P[:2].view(np.uint32), rP[:2].view(np.uint32))
        assert np.array_equal(st[:2, 1], rs[:2, 1]) and np.array_equal(st[:2, 2], rs[:2, 0])


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["plain", "exact"])
def test_trk_loop_vs_oracle(sdr, orc, variant):
    # the loop reads taps 0-3; the 8 plot taps come from the full-GPU pass;
    # "exact": -8 nibbles in the IF (the exact-sign correlator form)
    w, chans, loops, pos, truth = loop_case(nposx=8)
    assert len(pos) == 12
    if variant == "exact":
        w["buff"] = w["buff"].copy()
        w["buff"][::97] = 0x88
    plan = sdr.TrkPlan(w["buff"], chans, pos, w["n_epoch"], loops=loops)
    plan.run()
    st = plan.state()
    corr, _ = plan.result()
    P = plan.prompt()
    plan.close()
    for c, ch in enumerate(chans):
        rs, rc, rP = orc.trkl_run(ch, 1023, 1, 40.0, ch["code"], w["buff"], w["n_epoch"], pos)
        # before the first discriminator: bit-identical
        assert np.array_equal(corr[c, :2].view(np.uint32), rc[:2].view(np.uint32)), c
        assert np.array_equal(st[c, :2, :3], rs[:2, :3]), c
        # lock and loss decisions
        assert np.array_equal(st[c, :, 5], rs[:, 5]), c
        live = rs[:, 5] == 0
        assert np.array_equal(st[c, live, 4], rs[live, 4]), c
        # loop states: Doppler (Hz), code offset (s), C/N0 (dB-Hz); measured
        # on MI355X: the satellites' channels bit-identical over all 2500
        # epochs, the noise channel within 1.2e-13 Hz after a last-bit atan
        # difference at epoch 402
        assert np.max(np.abs(st[c, live, 0] - rs[live, 0])) < 1e-6, c
        assert np.max(np.abs(st[c, live, 1] - rs[live, 1])) < 1e-6 * CHIP, c
        assert np.max(np.abs(st[c, live, 3] - rs[live, 3])) < 1e-6, c
        if ch["prn"] in truth:
            a = np.abs(P[c, live]) ** 2
            b = np.abs(rP[live]) ** 2
            assert abs(a.mean() - b.mean()) < 1e-3 * b.mean(), c
            # the satellites' loop states are bit-identical, so are all taps
            # of every epoch (filled from the loop's recorded mixing states)
            assert np.array_equal(corr[c].view(np.uint32), rc.view(np.uint32)), c
        # epochs after a loss of lock correlate nothing
        assert not np.any(corr[c, ~live]), c


@pytest.mark.gpu
def test_trk_loop_rerun_and_open_loop_unchanged(sdr, orc):
    """A second run restarts from the initial states (same results); the
    open-loop plan has no loop state."""
    w, chans, loops, pos, truth = loop_case(n_epoch=600)
    plan = sdr.TrkPlan(w["buff"], chans, pos, w["n_epoch"], loops=loops)
    plan.run()
    s1 = plan.state()
    plan.run()
    s2 = plan.state()
    plan.close()
    assert np.array_equal(s1, s2)
    ol = sdr.TrkPlan(w["buff"], chans, pos, w["n_epoch"])
    with pytest.raises(ValueError):
        ol.state()
    ol.close()


def sec_case(n_epoch=2200, nposx=0):
    from pocketsdr_amd import workloads
    return workloads.trk_loop_sig("L5Q", seed=6, fs=12e6, n_ch=3, n_vis=2, n_epoch=n_epoch, nposx=nposx)


def bump_case(nposx=0):
    # channel 0 starts half a chip late, channel 1 half a chip early: each
    # on a side peak of the BOC(1,1) correlation
    from pocketsdr_amd import workloads
    w = workloads.trk_loop_sig("E1B", seed=7, fs=12e6, n_ch=2, n_vis=2, n_epoch=750, coff_err=0.5,
                               nposx=nposx)
    ch1 = w["chans"][1]
    ch1["coff0"] = w["truth"][ch1["prn"]]["tau"] - 0.5 * 4e-3 / 4092
    return w


def run_oracle(orc, w, c, **kw):
    ch, lp = w["chans"][c], w["loops"][c]
    return orc.trkl_run(ch, lp["len_code"], 1, lp["cn0"], ch["code"], w["buff"], w["n_epoch"], w["pos"], **kw)


def test_oracle_loop_secondary_code(orc):
    from pocketsdr_amd import codes
    w = sec_case()
    nh = codes.sec_code_NH20()
    e = np.arange(1100, 1500)
    for c in range(2):                                   # the two satellites
        lp = w["loops"][c]
        st, corr, P = run_oracle(orc, w, c, sec_code=lp["sec_code"])
        st0, corr0, P0 = run_oracle(orc, w, c)
        assert not st[:, 5].any() and st[-1, 3] > 40.0
        # before T_NPULLIN the code is on the prompt: its signs are NH20 at
        # one phase; the runs are the same until the sync
        s0 = np.sign(P0.real[e])
        assert max(abs(np.sum(s0 * nh[(e + k) % 20])) for k in range(20)) == len(e)
        assert np.array_equal(st[:1499], st0[:1499]) and np.array_equal(P[:1499], P0[:1499])
        # synchronised within the first code period after 1.5 s: the wiped
        # prompt and taps have one sign (the sync polarity makes it positive)
        assert (P.real[1519:] > 0).all() and (corr[1519:, 0].real > 0).all()
        assert not (P0.real[1519:] > 0).all()
        # the wipe is a sign per epoch: |P| and the loop's tracking unchanged
        assert np.array_equal(np.abs(P), np.abs(P0))
        assert np.array_equal(st[:, :4], st0[:, :4])


def test_oracle_loop_bump_jump(orc):
    w = bump_case()
    T, chip = 4e-3, 4e-3 / 4092
    for c in range(2):
        ch = w["chans"][c]
        tr = w["truth"][ch["prn"]]
        ref = tr["tau"] - tr["fd"] * (np.arange(1, w["n_epoch"] + 1) * T) / ch["fc_sig"]
        st, _, _ = run_oracle(orc, w, c, bump_step=w["loops"][c]["bump_step"])
        st0, _, _ = run_oracle(orc, w, c)
        err = (st[:, 1] - ref) / chip
        err0 = (st0[:, 1] - ref) / chip
        # one jump of one code entry (half a chip) at the first C/N0
        # window (lock 125 = epoch 124), towards the main peak
        jumps = np.nonzero(np.abs(np.diff(st[:, 1])) > 0.3 * chip)[0]
        assert jumps.tolist() == [123], jumps
        assert abs(abs(err[123] - err[124]) - 0.5) < 0.01
        assert abs(err[-1]) < 0.05 and abs(err0[-1]) > 0.45, (err[-1], err0[-1])
        assert st[-1, 3] > st0[-1, 3] + 4.0                 # the main peak's power


def compare_loop(sdr, orc, w, kw_fn):
    plan = sdr.TrkPlan(w["buff"], w["chans"], w["pos"], w["n_epoch"], loops=w["loops"])
    plan.run()
    st = plan.state()
    corr, _ = plan.result()
    P = plan.prompt()
    plan.close()
    out = []
    for c, ch in enumerate(w["chans"]):
        rs, rc, rP = run_oracle(orc, w, c, **kw_fn(c))
        assert np.array_equal(corr[c, :2].view(np.uint32), rc[:2].view(np.uint32)), c
        assert np.array_equal(st[c, :, 5], rs[:, 5]), c
        live = rs[:, 5] == 0
        assert np.array_equal(st[c, live, 4], rs[live, 4]), c
        assert np.max(np.abs(st[c, live, 0] - rs[live, 0])) < 1e-6, c
        assert np.max(np.abs(st[c, live, 1] - rs[live, 1])) < 1e-9 * 1e-3, c
        assert np.max(np.abs(st[c, live, 3] - rs[live, 3])) < 1e-6, c
        if ch["prn"] in w["truth"]:
            # until the loop states first differ (a last-bit atan / log10
            # difference), the wiped prompts and every tap are bit-identical;
            # after it, the prompt power agrees
            d = np.nonzero((st[c][:, :3] != rs[:, :3]).any(1))[0]
            e1 = int(d[0]) if len(d) else w["n_epoch"]
            assert e1 >= 2, c
            assert np.array_equal(P[c, :e1].view(np.uint32), rP[:e1].view(np.uint32)), c
            assert np.array_equal(corr[c, :e1].view(np.uint32), rc[:e1].view(np.uint32)), c
            a, b = np.abs(P[c, live]) ** 2, np.abs(rP[live]) ** 2
            assert abs(a.mean() - b.mean()) < 1e-3 * b.mean(), c
            print(f"channel {c}: loop states bit-identical for {e1} of {w['n_epoch']} epochs")
        assert not np.any(corr[c, ~live]), c
        out.append((st[c], P[c]))
    return out


@pytest.mark.gpu
def test_trk_loop_secondary_code_vs_oracle(sdr, orc):
    w = sec_case(nposx=8)
    out = compare_loop(sdr, orc, w, lambda c: dict(sec_code=w["loops"][c]["sec_code"]))
    for c in range(2):
        assert (out[c][1].real[1519:] > 0).all(), c


@pytest.mark.gpu
def test_trk_loop_bump_jump_vs_oracle(sdr, orc):
    w = bump_case(nposx=4)
    assert len(w["pos"]) == 10
    out = compare_loop(sdr, orc, w, lambda c: dict(bump_step=w["loops"][c]["bump_step"]))
    chip = 4e-3 / 4092
    for c in range(2):
        tr = w["truth"][w["chans"][c]["prn"]]
        ref = tr["tau"] - tr["fd"] * w["n_epoch"] * 4e-3 / w["chans"][c]["fc_sig"]
        assert abs(out[c][0][-1, 1] - ref) < 0.05 * chip, c


def test_trk_loop_rejects_bad_loop_parameters():
    """Checked before any device work: a secondary code longer than the
    prompt history (SDR_N_HIST = 5000), bump-jump without the VE / VL
    taps."""
    from pocketsdr_amd import sdr_func
    w = loop_case(n_epoch=4, n_ch=1, n_vis=1)
    chans = w[1]
    pos4 = w[3][:4]
    base = dict(len_code=1023, costas=1, cn0=40.0)
    bad = [(dict(base, sec_code=np.ones(5001, np.int8)), w[3]),
           (dict(base, bump_step=1e-6), pos4)]
    for lp, pos in bad:
        with pytest.raises(ValueError):
            sdr_func.TrkPlan(w[0]["buff"], chans, pos, 4, loops=[lp])
