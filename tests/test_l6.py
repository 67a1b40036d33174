"""QZSS L6D / L6E FFT tracking and CSK (src/sdr_ch.c:627-664, 694-720).

CPU tests pin the oracle's CSK restatement with answers the mathematics
fixes (a delta correlation peak at d chips decodes to 255 - d, linear
interpolation of a linear correlation is exact, the first of two equal peaks
wins, the reference loop's NaN behaviour) and its tracking loop with an
end-to-end decode of planted code-shift symbols.  GPU tests compare the HIP
plan with the oracle: CSK symbols and taps bit-exact on the same correlation,
correlations within the FP32 FFT tolerance of the other correlator tests,
symbols equal.
"""
import numpy as np
import pytest

LEN_L6 = 20460


def delta_corr(N, lags, amps):
    c = np.zeros(N, np.complex64)
    for m, a in zip(lags, amps):
        c[m % N] += a
    return c


# ------------------------------------------------------------ oracle pins
@pytest.mark.parametrize("d", [0, 1, 37, 200, 255, -1, -8, -255])
def test_oracle_csk_delta(orc, d):
    """A correlation peak exactly at d chips (R = 2 samples per chip with
    N = len_code) decodes to shift d and symbol (255 - d) mod 256."""
    N = LEN_L6
    corr = delta_corr(N, [2 * d], [0.5 - 0.25j])
    ix, sym, taps = orc.csk(corr, LEN_L6, [0.0, 1.0, -1.0])
    assert ix == d and sym == (255 - d) & 0xFF
    # taps: the peak itself and the zero samples either side of it
    assert taps[0] == np.complex64(0.5 - 0.25j)
    assert np.allclose(taps[1:], 0.0)


def test_oracle_csk_fractional_R(orc):
    """Non-integer samples per chip (N = 24000, R = 2.346...): a peak at the
    sample nearest d R still decodes to d."""
    N = 24000
    R = N / (LEN_L6 // 2)
    for d in (3, 100, -77, 254):
        corr = delta_corr(N, [int(round(d * R))], [1.0])
        assert orc.csk(corr, LEN_L6, [0.0])[0] == d


def test_oracle_csk_linear_interp(orc):
    """corr linear in the lag m (|m| < N / 2): the interpolated taps equal
    the line at x exactly (up to FP32 rounding) and the power peaks at the
    far end of the scanned range."""
    N = 24000
    R = N / (LEN_L6 // 2)
    m = np.arange(N, dtype=np.float64)
    m[m >= N // 2] -= N
    beta, m0 = 1e-3 * (0.6 + 0.8j), 50.0
    corr = (beta * (m + m0)).astype(np.complex64)
    pos = np.array([0.0, -1.7, 2.3, -120.0, 0.25])
    ix, sym, taps = orc.csk(corr, LEN_L6, pos)
    assert ix == 255 and sym == 0
    want = beta * (255 * R + pos + m0)
    assert np.allclose(taps, want, rtol=1e-6, atol=0)


def test_oracle_csk_first_max_wins(orc):
    N = LEN_L6
    corr = delta_corr(N, [2 * -5, 2 * 7], [1.0, 1.0])
    assert orc.csk(corr, LEN_L6, [])[0] == -5
    corr = delta_corr(N, [2 * 7, 2 * 9], [1.0, 1.0j])
    assert orc.csk(corr, LEN_L6, [])[0] == 7


def test_oracle_csk_zero_and_nan(orc):
    N = LEN_L6
    ix, sym, _ = orc.csk(np.zeros(N, np.complex64), LEN_L6, [])
    assert (ix, sym) == (0, 255)
    # the reference loop skips P <= P_max; a NaN power fails that test, is
    # taken, and the next shift (P = 0, not <= NaN) is taken after it
    corr = delta_corr(N, [2 * 17], [np.nan])
    assert orc.csk(corr, LEN_L6, [])[0] == 18


def test_oracle_l6_decodes_planted_symbols(orc):
    from pocketsdr_amd import workloads
    fn = lambda c, T, coff, fs, N: orc.gen_code_fft(c, T, coff, fs, N, 0)
    w = workloads.trk_l6(seed=3, fs=6e6, n_ch=3, n_epoch=8, n_vis=2, code_fft_fn=fn)
    for c, ch in enumerate(w["chans"][:2]):
        syms, taps, _, st = orc.trk_l6_run(ch, w["len_code"], ch["code_fft"], w["buff"],
                                           w["n_epoch"], w["pos"])
        assert syms.tolist() == w["truth"][ch["prn"]]["syms"][:w["n_epoch"]].tolist()
        # the open loop is the standard tracking loop's
        _, _, st2 = orc.trk_run(dict(ch), np.zeros((10 * ch["N"], 2), np.int8), w["buff"],
                                w["n_epoch"], [0.0])
        assert np.array_equal(st[:, :2], st2[:, :2])
        assert np.all((st[:, 3] >= 0) & (st[:, 3] <= 9))


def test_oracle_l6_composition(orc):
    """Each epoch is sdr_corr_fft of the carrier-mixed samples from ring index
    e N + i at phase phi + fc i / fs against bank j, followed by CSK."""
    from pocketsdr_amd import workloads
    fn = lambda c, T, coff, fs, N: orc.gen_code_fft(c, T, coff, fs, N, 0)
    w = workloads.trk_l6(seed=5, fs=6e6, n_ch=1, n_epoch=3, n_vis=1, code_fft_fn=fn)
    ch = w["chans"][0]
    syms, taps, corr, st = orc.trk_l6_run(ch, w["len_code"], ch["code_fft"], w["buff"],
                                          w["n_epoch"], w["pos"], keep_corr=True)
    N, fs = ch["N"], ch["fs"]
    for e in range(w["n_epoch"]):
        phi, i, j = st[e, 0], int(st[e, 2]), int(st[e, 3])
        fc = ch["fi"] + ch["fd0"]
        IQ = orc.mix_carr(w["buff"], (e * N + i) % len(w["buff"]), N, fs, fc, phi + fc * i / fs)
        ref = orc.corr_fft(IQ, ch["code_fft"][j])
        assert np.array_equal(corr[e].view(np.uint32), ref.view(np.uint32))
        ix, sym, tp = orc.csk(ref, w["len_code"], w["pos"])
        assert sym == syms[e] and np.array_equal(tp.view(np.uint32), taps[e].view(np.uint32))


# ------------------------------------------------------------ GPU parity
def csk_cases(rng):
    N = 24000
    R = N / (LEN_L6 // 2)
    out = []
    noise = (rng.standard_normal((6, N)) + 1j * rng.standard_normal((6, N))).astype(np.complex64) * 0.01
    for k, d in enumerate((0, 5, -9, 255, -255, 128)):
        c = noise[k].copy()
        c[int(round(d * R)) % N] += 0.3 + 0.1j
        out.append((N, c))
    out.append((N, noise[0]))                                  # noise only
    out.append((LEN_L6, delta_corr(LEN_L6, [-10, 14], [1.0, 1.0])))   # exact tie
    out.append((LEN_L6, np.zeros(LEN_L6, np.complex64)))       # no power
    out.append((LEN_L6, delta_corr(LEN_L6, [34], [np.nan])))   # NaN power
    return out


@pytest.mark.gpu
def test_csk_bitexact(sdr, orc):
    rng = np.random.default_rng(11)
    pos = np.array([0.0, -2.3, 2.3, -120.0, 0.5, -0.5, 7.25])
    cases = csk_cases(rng)
    for N in sorted({n for n, _ in cases}):
        batch = np.stack([c for n, c in cases if n == N])
        syms, taps = sdr.csk(batch, LEN_L6, pos)
        for k in range(len(batch)):
            ix, sym, tp = orc.csk(batch[k], LEN_L6, pos)
            assert syms[k] == sym, (N, k)
            assert np.array_equal(taps[k].view(np.uint32), tp.view(np.uint32)), (N, k)


def rel(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30)


def l6_workload(orc, fs, n_ch, n_epoch, n_vis, buff_epochs=None, seed=3):
    from pocketsdr_amd import workloads
    fn = lambda c, T, coff, fs_, N: orc.gen_code_fft(c, T, coff, fs_, N, 0)
    return workloads.trk_l6(seed=seed, fs=fs, n_ch=n_ch, n_epoch=n_epoch, n_vis=n_vis,
                            code_fft_fn=fn, buff_epochs=buff_epochs)


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["6M", "12M-chunks", "wrap", "unfused"])
def test_l6_plan(sdr, orc, variant, monkeypatch):
    fs = 12e6 if variant == "12M-chunks" else 6e6
    if variant == "12M-chunks":
        monkeypatch.setenv("PSA_L6_CHUNK", "7")     # several batches and a remainder
    if variant == "unfused":
        monkeypatch.setenv("PSA_L6_UNFUSED", "1")   # separate forward and inverse row passes
    w = l6_workload(orc, fs, n_ch=4, n_epoch=9, n_vis=2,
                    buff_epochs=(5 if variant == "wrap" else None))
    buff = w["buff"]
    if variant == "wrap":
        buff = np.concatenate([buff, buff[:1234]])   # ring not a multiple of N
    plan = sdr.L6Plan(buff, w["chans"], w["pos"], w["n_epoch"], keep_corr=True)
    plan.run()
    syms, taps = plan.result()
    for c, ch in enumerate(w["chans"]):
        rs, rt, rc, _ = orc.trk_l6_run(ch, w["len_code"], ch["code_fft"], buff, w["n_epoch"],
                                       w["pos"], keep_corr=True)
        for e in range(w["n_epoch"]):
            got = plan.corr(c, e)
            assert rel(got, rc[e]) < 1e-5, (c, e)
            if syms[c, e] != rs[e]:
                # a noise-only correlation may hold two near-equal peaks: the
                # GPU's choice must then be as strong as the oracle's
                R = ch["N"] / (w["len_code"] // 2)
                ir, _, _ = orc.csk(rc[e], w["len_code"], [])
                ig = 255 - int(syms[c, e])
                ig = ig - 256 if ig > 255 else ig
                pr, pg = np.abs(orc.csk(rc[e], w["len_code"], [0.0, (ig - ir) * R])[2])
                assert ch["prn"] not in w["truth"] and pg >= pr * (1 - 1e-4), (c, e)
            else:
                assert rel(taps[c, e], rt[e]) < 1e-4, (c, e)
    plan.close()
    # without kept correlations only the lag window CSK reads is stored:
    # the same symbols and taps, bit for bit
    plan2 = sdr.L6Plan(buff, w["chans"], w["pos"], w["n_epoch"])
    plan2.run()
    s2, t2 = plan2.result()
    assert np.array_equal(s2, syms)
    assert np.array_equal(t2.view(np.uint32), taps.view(np.uint32))
    plan2.close()


@pytest.mark.gpu
def test_l6_plan_decodes(sdr, orc):
    w = l6_workload(orc, 24e6, n_ch=6, n_epoch=12, n_vis=3, seed=8)
    plan = sdr.L6Plan(w["buff"], w["chans"], w["pos"], w["n_epoch"])
    plan.run()
    syms, _ = plan.result()
    seen = set()
    for c, ch in enumerate(w["chans"]):
        if ch["prn"] in w["truth"] and ch["prn"] not in seen:
            seen.add(ch["prn"])
            assert syms[c].tolist() == w["truth"][ch["prn"]]["syms"][:w["n_epoch"]].tolist(), c
    assert len(seen) == 3
    with pytest.raises(ValueError):
        plan.corr(0, 0)                              # made without keep_corr
    plan.close()


@pytest.mark.gpu
def test_l6_plan_device_ring(sdr, orc):
    """sdr_l6_plan_new_dev on an IF ring in device memory: the same symbols
    and taps as the plan on the host ring, also after new samples are
    written into the device ring between runs."""
    import torch
    w = l6_workload(orc, 6e6, n_ch=3, n_epoch=6, n_vis=2, seed=12)
    buff = np.asarray(w["buff"], np.uint8)
    dring = torch.from_numpy(buff.copy()).cuda()
    pd = sdr.L6Plan((dring.data_ptr(), len(buff)), w["chans"], w["pos"], w["n_epoch"], dev_buff=True)
    for data in (buff, np.roll(buff, 4321)):
        dring.copy_(torch.from_numpy(data))
        pd.run()
        ph = sdr.L6Plan(data, w["chans"], w["pos"], w["n_epoch"])
        ph.run()
        s1, t1 = pd.result()
        s2, t2 = ph.result()
        assert np.array_equal(s1, s2) and np.array_equal(t1.view(np.uint32), t2.view(np.uint32))
        ph.close()
    pd.close()


@pytest.mark.gpu
def test_l6_plan_full_size(sdr, orc):
    """The L6 configuration of the bench at full size (16 channels x 250
    epochs at 24 Msps, N = 96000): every planted symbol decoded, and one
    visible channel's symbols and taps against the oracle."""
    from pocketsdr_amd import workloads
    w = workloads.trk_l6(seed=3, fs=24e6, n_ch=16, n_epoch=250)
    plan = sdr.L6Plan(w["buff"], w["chans"], w["pos"], w["n_epoch"])
    plan.run()
    syms, taps = plan.result()
    plan.close()
    seen = set()
    for c, ch in enumerate(w["chans"]):
        if ch["prn"] in w["truth"] and ch["prn"] not in seen:
            seen.add(ch["prn"])
            assert syms[c].tolist() == w["truth"][ch["prn"]]["syms"][:w["n_epoch"]].tolist(), c
    assert len(seen) == 4
    ch = w["chans"][0]
    rs, rt, _, _ = orc.trk_l6_run(ch, w["len_code"], ch["code_fft"], w["buff"], w["n_epoch"], w["pos"])
    assert np.array_equal(syms[0], rs)
    assert rel(taps[0], rt) < 1e-4


# ------------------------------------------------------------ closed loop
CHIP_L6 = 4e-3 / 10230


def l6_loop_case(orc, fs=6e6, n_ch=3, n_epoch=500, n_vis=2, seed=3, fd_err=15.0, coff_err=0.1):
    """Channels of the visible satellites started fd_err Hz and coff_err
    chips off; loops as start_track sets them for L6D (plain PLL)."""
    w = l6_workload(orc, fs, n_ch, n_epoch, n_vis, seed=seed)
    seen = set()
    for ch in w["chans"]:
        tr = w["truth"].get(ch["prn"])
        if tr is not None and ch["prn"] not in seen:
            seen.add(ch["prn"])
            ch["fd0"] = tr["fd"] + fd_err
            ch["coff0"] = tr["tau"] + coff_err * CHIP_L6
    w["loops"] = [dict(len_code=LEN_L6, costas=0, cn0=40.0) for _ in w["chans"]]
    return w


def test_oracle_l6_loop_pull_in(orc):
    """The L6 branch inside the loop (src/sdr_ch.c:705-720, 741-763): the
    satellites' channels pull in from 15 Hz and 0.1 chip off and decode
    every planted symbol; before the first frequency update (epochs 0-1)
    the loop equals the open-loop restatement."""
    w = l6_loop_case(orc)
    for c, ch in enumerate(w["chans"][:2]):
        tr = w["truth"][ch["prn"]]
        st, syms, taps = orc.l6l_run(ch, LEN_L6, 40.0, ch["code_fft"], w["buff"], w["n_epoch"], w["pos"])
        assert not st[:, 5].any() and st[-1, 4] == w["n_epoch"]
        assert syms.tolist() == tr["syms"][:w["n_epoch"]].tolist()
        assert abs(st[-1, 0] - tr["fd"]) < 1.0 and abs(st[249, 0] - tr["fd"]) < 1.0
        coff = tr["tau"] - tr["fd"] * w["n_epoch"] * 4e-3 / ch["fc_sig"]
        assert abs(st[-1, 1] - coff) < 0.1 * CHIP_L6
        s0, t0, _, st0 = orc.trk_l6_run(dict(ch, drift=0.0), LEN_L6, ch["code_fft"], w["buff"], 2, w["pos"])
        assert np.array_equal(syms[:2], s0) and np.array_equal(taps[:2].view(np.uint32), t0.view(np.uint32))
        assert np.array_equal(st[:2, 2], st0[:, 0]) and np.array_equal(st[:2, 1], st0[:, 1])


def test_l6_loop_rejects_bad_loop_parameters():
    from pocketsdr_amd import sdr_func
    from oracle import oracle as orc
    w = l6_loop_case(orc, n_ch=1, n_epoch=2, n_vis=1)
    bad = [(dict(len_code=LEN_L6, cn0=40.0, sec_code=np.ones(20, np.int8)), w["pos"]),
           (dict(len_code=LEN_L6, cn0=40.0, bump_step=1e-7), w["pos"][:4]),
           (dict(len_code=LEN_L6, cn0=40.0), w["pos"][:3])]
    for lp, pos in bad:
        with pytest.raises((ValueError, TypeError)):
            sdr_func.L6Plan(w["buff"], w["chans"], pos, 2, loops=[lp])


@pytest.mark.gpu
@pytest.mark.parametrize("fs,n_ch,n_epoch", [(6e6, 3, 500), (24e6, 10, 300)])
def test_l6_loop_vs_oracle(sdr, orc, fs, n_ch, n_epoch, monkeypatch):
    w = l6_loop_case(orc, fs=fs, n_ch=n_ch, n_epoch=n_epoch, n_vis=2 if n_ch == 3 else 4)
    plan = sdr.L6Plan(w["buff"], w["chans"], w["pos"], w["n_epoch"], loops=w["loops"])
    plan.run()
    st = plan.state()
    syms, taps = plan.result()
    # reruns (replays of the captured loop) start from the initial states:
    # the same states, symbols and taps every time, and the same as a timed
    # run, which launches the kernels one by one
    for rep in range(3):
        plan.timing(rep == 1)
        plan.run()
        assert np.array_equal(plan.state(), st), rep
        s2, t2 = plan.result()
        assert np.array_equal(s2, syms) and np.array_equal(t2.view(np.uint32), taps.view(np.uint32)), rep
    plan.close()
    seen = set()
    for c, ch in enumerate(w["chans"]):
        rs, rsym, rtaps = orc.l6l_run(ch, LEN_L6, 40.0, ch["code_fft"], w["buff"], w["n_epoch"], w["pos"])
        assert not np.any(syms[c, st[c, :, 5] == 1]) and not np.any(taps[c, st[c, :, 5] == 1]), c
        if ch["prn"] not in w["truth"] or ch["prn"] in seen:
            # a channel on noise: its CSK peak is the largest of 511 noise
            # powers, which the last-bit differences of the correlations
            # move, so only the epochs before the first frequency update
            # (the FLL's at epoch 1) agree
            assert np.array_equal(syms[c, :2], rsym[:2]), c
            assert np.array_equal(st[c, :1, :3], rs[:1, :3]), c
            continue
        seen.add(ch["prn"])
        # a satellite: the same lock counts, no loss, every planted symbol
        assert not rs[:, 5].any() and not st[c, :, 5].any(), c
        assert np.array_equal(st[c, :, 4], rs[:, 4]), c
        assert np.array_equal(syms[c], rsym), c
        assert syms[c].tolist() == w["truth"][ch["prn"]]["syms"][:w["n_epoch"]].tolist(), c
        # the correlations differ from the oracle's PocketFFT in the last
        # bits (FP32 FFTs of other orders), so the loops drift apart by
        # rounding: Doppler (Hz), code offset (chip), C/N0 (dB-Hz)
        assert np.max(np.abs(st[c, :, 0] - rs[:, 0])) < 1e-2, c
        assert np.max(np.abs(st[c, :, 1] - rs[:, 1])) < 1e-4 * CHIP_L6, c
        assert np.max(np.abs(st[c, :, 3] - rs[:, 3])) < 1e-2, c
        # the taps follow: the loops' carrier phases differ by rounding, which
        # rotates the taps (measured 5.8e-4 of the largest tap at 6 Msps)
        d = np.abs(taps[c] - rtaps)
        assert np.max(d) <= 2e-3 * np.max(np.abs(rtaps)), c
        dp = np.abs(np.abs(taps[c]) - np.abs(rtaps))
        assert np.max(dp) <= 1e-3 * np.max(np.abs(rtaps)), c
    assert len(seen) == (2 if n_ch == 3 else 4)
    # epochs 0-1 precede any loop update: the loop plan's symbols and taps
    # equal the open-loop plan's (the same GPU passes) bit for bit
    ol = sdr.L6Plan(w["buff"], w["chans"], w["pos"], 2)
    ol.run()
    s0, t0 = ol.result()
    ol.close()
    assert np.array_equal(syms[:, :2], s0) and np.array_equal(taps[:, :2].view(np.uint32), t0.view(np.uint32))
