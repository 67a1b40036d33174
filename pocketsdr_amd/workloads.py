"""Seeded synthetic workloads of the search / correlate hot path.

acq_l1ca      GPS L1 C/A acquisition as pocket_acq runs it
              (app/pocket_acq/pocket_acq.c:152-192): complex IQ cpx8 at fs,
              1 ms coherent with code zero-padded to M = 2N, T_AQC = 10 ms of
              non-coherent blocks, Doppler +/-5 kHz in 1/(2T) steps.
trk_l1ca      open-loop tracking correlators on one IF stream: per channel
              E/P/L at +/-0.5 chip, the N tap at -120 samples and 40 extra taps
              over +/-2 chips (src/sdr_ch.c:228-231, 804-814).
trk_l6        QZSS L6D FFT tracking with CSK: 4 ms epochs, 20460-entry TDM
              code, each code period of a visible satellite delayed by
              255 - symbol chips (src/sdr_ch.c:636-664, 705-720).
trk_loop_sig  closed-loop channels of GPS L1 C/A, GPS L5-Q (NH20 secondary
              code) or Galileo E1-B (BOC(1,1), the VE / VL taps of
              bump-jump) started off the satellites' Doppler and code phase
              (src/sdr_ch.c:221-244, 453-486, 552-572).
"""
import numpy as np

from . import codes, synth

L1_FREQ = 1575.42e6

# G2 delays of PRN 33-64 (IS-GPS-200 extended table, same values as the
# reference's src/sdr_code.c:159-161)
_L1CA_G2_DELAY_33_64 = (863, 950, 947, 948, 950, 67, 103, 91, 19, 679, 225, 625, 946, 638, 161,
                        1001, 554, 280, 710, 709, 775, 864, 558, 220, 397, 55, 898, 759, 367, 299,
                        1018, 729)


def l1ca_code(prn):
    if prn <= 32:
        return codes.gen_code_L1CA(prn)
    if prn > 64:
        raise ValueError("PRN 1-64 only")
    c1 = codes.gen_code_L1CA(1)  # ensures the shift registers exist
    del c1
    d = _L1CA_G2_DELAY_33_64[prn - 33]
    idx = (np.arange(1023) + 1023 - d) % 1023
    return (codes._G1 * codes._G2[idx]).astype(np.int8)


def dop_bins(T, dop, max_dop):
    """sdr_dop_bins arithmetic (src/sdr_func.c:758-769) in float32"""
    step = np.float32(0.5 / T)
    n = int(2.0 * float(np.float32(max_dop)) / float(step)) + 1
    return (np.float32(dop) - np.float32(max_dop) + np.arange(n, dtype=np.float32) * step).astype(np.float32)


def acq_l1ca(seed=1, fs=12e6, n_prn=32, n_vis=8, t_int=10e-3, max_dop=5000.0,
             cn0=(38.0, 50.0), bits=2, code_fft_fn=None):
    """Returns dict(buff, sig, truth).  code_fft_fn(code, T, fs, N) gives the
    conjugated, zero-padded code spectrum (default: the library's GPU FFT)."""
    rng = np.random.default_rng(seed)
    T = 1e-3
    N = int(fs * T)
    M = 2 * N
    n_samp = int(fs * (t_int + T))
    prns = list(range(1, n_prn + 1))
    vis = sorted(rng.choice(prns, n_vis, replace=False).tolist())
    cd = {p: l1ca_code(p) for p in prns}
    sats = synth.random_sats(rng, vis, cd, T, L1_FREQ, cn0, (-max_dop, max_dop))
    buff = synth.synth_if(fs, n_samp, sats, seed, bits=bits)
    fn = code_fft_fn or (lambda c, T_, fs_, N_: codes.gen_code_fft(c, T_, 0.0, fs_, N_, N_))
    code_fft = np.stack([fn(cd[p], T, fs, N) for p in prns]).astype(np.complex64)
    fds = dop_bins(T, 0.0, max_dop)
    n_blk = len(range(0, n_samp - 2 * N + 1, N))
    sig = dict(buff=0, M=M, N=N, T=T, fs=fs, fi=0.0, code_fft=code_fft, fds=fds,
               n_blk=n_blk, ix0=0, blk_step=N, prns=prns)
    return dict(buff=buff, sig=sig, truth=sats)


def trk_positions(fs, T=1e-3, len_code=1023, sp_corr=1.0, nposx=40, width_chip=4.0):
    """P, E, L, N (src/sdr_ch.c:224-231) + nposx extra taps (src/sdr_ch.c:804-814)"""
    sc = T / len_code * fs
    sp = sp_corr * sc
    pos = [0.0, -0.5 * sp, 0.5 * sp, -120.0]
    width = width_chip / 1.023e6
    for i in range(nposx):
        pos.append(float(i - (nposx - 1) // 2) / nposx * width * fs)
    return np.array(pos, np.float64)


def trk_l1ca(seed=2, fs=24e6, n_ch=64, n_epoch=1000, n_vis=8, cn0=(38.0, 50.0),
             drift=(-2.0, 2.0), bits=2, buff_epochs=None, nposx=40):
    """Returns dict(buff, chans, pos, n_epoch).  The IF holds n_vis of the
    channels' PRNs; their channels start on the true code phase / Doppler."""
    rng = np.random.default_rng(seed)
    T = 1e-3
    N = int(fs * T)
    n_buf = N * (buff_epochs or n_epoch)
    prns = list(range(1, n_ch + 1))
    vis = sorted(rng.choice(prns, min(n_vis, n_ch), replace=False).tolist())
    cd = {p: l1ca_code(p) for p in prns}
    sats = synth.random_sats(rng, vis, cd, T, L1_FREQ, cn0, (-5000.0, 5000.0), drift)
    truth = {s["prn"]: s for s in sats}
    buff = synth.synth_if(fs, n_buf, sats, seed, bits=bits)
    chans = []
    for p in prns:
        if p in truth:
            s = truth[p]
            fd0, dr, coff0, phi0 = s["fd"], s["drift"], s["tau"], 0.0
        else:
            fd0 = float(rng.uniform(-5000, 5000))
            dr = float(rng.uniform(*drift))
            coff0 = float(rng.uniform(0, T))
            phi0 = float(rng.uniform(0, 1))
        chans.append(dict(fs=fs, fi=0.0, fc_sig=L1_FREQ, T=T, N=N, fd0=fd0, drift=dr,
                          phi0=phi0, coff0=coff0, code=codes.code_bank(cd[p], T, fs, N), prn=p))
    return dict(buff=buff, chans=chans, pos=trk_positions(fs, T, nposx=nposx), n_epoch=n_epoch,
                truth=sats)


L6_FREQ = 1278.75e6


def trk_l6(seed=3, fs=24e6, n_ch=16, n_epoch=250, n_vis=4, cn0=(45.0, 50.0), nposx=8,
           code_fft_fn=None, buff_epochs=None):
    """Returns dict(buff, chans, pos, n_epoch, len_code, truth).  Channel c
    tracks L6D PRN 193 + c % 10; the first n_vis PRNs are in the IF with
    random CSK symbols (truth[prn]["syms"], one per 4 ms code period) and
    their channels start on the true code phase and Doppler (code phase in
    the middle half of the period, so the open-loop code offset never
    wraps).  code_fft_fn(code, T, coff, fs, N) makes one N-point conjugated
    code spectrum (default: codes.gen_code_fft)."""
    rng = np.random.default_rng(seed)
    T, len_code = 4e-3, 20460
    N = int(fs * T)
    fn = code_fft_fn or (lambda c, T_, coff, fs_, N_: codes.gen_code_fft(c, T_, coff, fs_, N_, 0))
    n_buf = N * (buff_epochs or n_epoch + 1)
    prns = [193 + c % 10 for c in range(n_ch)]
    cd = {p: codes.gen_code_L6D(p) for p in sorted(set(prns))}
    vis = sorted(set(prns))[:n_vis]
    n_per = n_buf // N + 2
    sats = []
    for p in vis:
        syms = rng.integers(0, 256, n_per)
        sats.append(dict(prn=p, code=cd[p], T=T, fc=L6_FREQ, tau=float(rng.uniform(0.25, 0.75) * T),
                         fd=float(rng.uniform(-4000, 4000)), drift=0.0,
                         cn0=float(rng.uniform(*cn0)), phi=float(rng.uniform(0, 1)),
                         syms=syms, shift=2 * (255 - syms)))
    truth = {s_["prn"]: s_ for s_ in sats}
    buff = synth.synth_if(fs, n_buf, sats, seed, bits=2)
    banks = {p: np.stack([fn(cd[p], T, -k / fs / 10, fs, N) for k in range(10)]).astype(np.complex64)
             for p in cd}
    chans = []
    seen = set()
    for c, p in enumerate(prns):
        if p in truth and p not in seen:
            s_ = truth[p]
            fd0, coff0, phi0 = s_["fd"], s_["tau"], 0.0
        else:
            fd0 = float(rng.uniform(-4000, 4000))
            coff0 = float(rng.uniform(0, T))
            phi0 = float(rng.uniform(0, 1))
        seen.add(p)
        chans.append(dict(fs=fs, fi=0.0, fc_sig=L6_FREQ, T=T, N=N, len_code=len_code, fd0=fd0,
                          drift=0.0, phi0=phi0, coff0=coff0, code_fft=banks[p], prn=p))
    return dict(buff=buff, chans=chans, pos=trk_positions(fs, T, len_code // 2, 0.25, nposx),
                n_epoch=n_epoch, len_code=len_code, truth=truth)


# ---------------------------------------------------------------------------
# batched uneven search: four signals on three IF streams
# ---------------------------------------------------------------------------
MULTI_SIGNALS = (
    # name, stream, fs, code period, prns, generator, carrier, n_visible
    ("L1CA", 0, 12e6, 1e-3, list(range(1, 33)), "L1CA", L1_FREQ, 8),
    ("E1B", 0, 12e6, 4e-3, list(range(1, 37)), "E1B", L1_FREQ, 9),
    ("L5Q", 1, 24e6, 1e-3, list(range(1, 33)), "L5Q", 1176.45e6, 8),
    ("L6D", 2, 24e6, 4e-3, list(range(193, 203)), "L6D", 1278.75e6, 3),
)


def sig_code(gen, prn):
    if gen == "L1CA":
        return l1ca_code(prn)
    return getattr(codes, "gen_code_" + gen)(prn)


def acq_multi(seed=5, t_int=10e-3, max_dop=5000.0, cn0=(38.0, 50.0), bits=2, code_fft_fn=None,
              signals=MULTI_SIGNALS, prn_limit=None):
    """GPS L1 C/A and Galileo E1-B on one 12 Msps L1 stream, GPS L5-Q and
    QZSS L6D on 24 Msps streams; T_AQC = 10 ms plus one code period per
    stream (pocket_acq's buffer length), Doppler bins 1/(2T) over +/-5 kHz."""
    rng = np.random.default_rng(seed)
    fn = code_fft_fn or (lambda c, T, fs, N: codes.gen_code_fft(c, T, 0.0, fs, N, N))
    streams = {}
    for name, st, fs, T, prns, gen, fc, nv in signals:
        streams.setdefault(st, dict(fs=fs, Tmax=0.0, sats=[]))
        streams[st]["Tmax"] = max(streams[st]["Tmax"], T)
    sigs, truth = [], {}
    for name, st, fs, T, prns, gen, fc, nv in signals:
        if prn_limit:
            prns = prns[:prn_limit]
            nv = min(nv, len(prns))
        cd = {p: sig_code(gen, p) for p in prns}
        vis = sorted(rng.choice(prns, nv, replace=False).tolist())
        sats = synth.random_sats(rng, vis, cd, T, fc, cn0, (-max_dop, max_dop))
        streams[st]["sats"] += sats
        truth[name] = sats
        N = int(fs * T)
        n_samp = int(fs * (t_int + streams[st]["Tmax"]))
        n_blk = len(range(0, int(fs * (t_int + T)) - 2 * N + 1, N))
        code_fft = np.stack([fn(cd[p], T, fs, N) for p in prns]).astype(np.complex64)
        sigs.append(dict(name=name, buff=st, M=2 * N, N=N, T=T, fs=fs, fi=0.0, code_fft=code_fft,
                         fds=dop_bins(T, 0.0, max_dop), n_blk=n_blk, ix0=0, blk_step=N, prns=prns))
    buffs = []
    for st in sorted(streams):
        d = streams[st]
        n_samp = int(d["fs"] * (t_int + d["Tmax"]))
        buffs.append(synth.synth_if(d["fs"], n_samp, d["sats"], seed * 10 + st, bits=bits))
    return dict(buffs=buffs, sigs=sigs, truth=truth)


# ---------------------------------------------------------------------------
# weak-signal wide search: GPS L1C-P on an antenna array
# ---------------------------------------------------------------------------
def acq_l1cp(seed=9, n_rf=8, fs=48e6, n_prn=32, n_vis=8, n_blk=20, max_dop=6000.0,
             cn0=(25.0, 35.0), bits=3, code_fft_fn=None, prn_limit=None, bin_limit=None):
    """GPS L1C pilot (10 ms, 10230 chips TMBOC, 122760 replica entries,
    src/sdr_code.c:884-913) at 48 Msps complex 3-bit on n_rf RF channels of
    one antenna array: every channel sees the same satellites (code phase,
    Doppler, C/N0) with its own carrier phase and noise.  10 ms coherent
    (N = 480000, M = 960000) x n_blk non-coherent blocks, Doppler bins of
    1/(2T) = 50 Hz over +/-6 kHz (241 bins).  One signal group per RF channel;
    the groups share one code-spectrum array.  prn_limit / bin_limit cut the
    PRN list / take the bin_limit bins around each visible satellite's
    Doppler (reduced sizes for tests)."""
    rng = np.random.default_rng(seed)
    T = 10e-3
    N = int(round(fs * T))
    M = 2 * N
    n_samp = (n_blk + 1) * N
    prns = list(range(1, n_prn + 1))[:prn_limit] if prn_limit else list(range(1, n_prn + 1))
    vis = sorted(rng.choice(prns, min(n_vis, len(prns)), replace=False).tolist())
    cd = {p: codes.gen_code_L1CP(p) for p in prns}
    sats = synth.random_sats(rng, vis, cd, T, L1_FREQ, cn0, (-max_dop, max_dop))
    fn = code_fft_fn or (lambda c, T_, fs_, N_: codes.gen_code_fft(c, T_, 0.0, fs_, N_, N_))
    code_fft = np.stack([fn(cd[p], T, fs, N) for p in prns]).astype(np.complex64)
    fds = dop_bins(T, 0.0, max_dop)
    if bin_limit:
        keep = set()
        for s in sats:
            c = int(np.argmin(np.abs(fds - s["fd"])))
            keep.update(range(max(0, c - bin_limit // 2), min(len(fds), c - bin_limit // 2 + bin_limit)))
        fds = fds[sorted(keep)]
    sigs = []
    rf_sats = []
    for r in range(n_rf):
        rf_sats.append([dict(s, phi=float(rng.uniform(0, 1))) for s in sats])
        sigs.append(dict(name="L1CP@rf%d" % r, buff=r, M=M, N=N, T=T, fs=fs, fi=0.0, code_fft=code_fft,
                         fds=fds, n_blk=n_blk, ix0=0, blk_step=N, prns=prns))
    # the RF streams are independent: synthesised on a pool of host threads
    # (numpy releases the interpreter lock in its array work)
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max(1, min(n_rf, 8))) as ex:
        buffs = list(ex.map(lambda r: synth.synth_if(fs, n_samp, rf_sats[r], seed * 100 + r, bits=bits),
                            range(n_rf)))
    return dict(buffs=buffs, sigs=sigs, truth={s["name"]: sats for s in sigs})


# ---------------------------------------------------------------------------
# real IF data
# ---------------------------------------------------------------------------
def read_if(path, IQ, n=None):
    """int8 IF samples to the cpx8 buffer as sdr_read_data does
    (src/sdr_func.c:451-510): I-sampling (IQ 1) takes SDR_CPX8(raw[i], 0),
    IQ-sampling (IQ 2) SDR_CPX8(raw[2i], -raw[2i+1]) (Q polarity flipped);
    n samples (None: all)."""
    raw = np.fromfile(path, dtype=np.int8).astype(np.int64)
    if IQ == 1:
        I, Q = raw, np.zeros_like(raw)
    else:
        I, Q = raw[0::2], -raw[1::2]
    if n is not None:
        I, Q = I[:n], Q[:n]
    return ((I & 0xF) | ((Q & 0xF) << 4)).astype(np.uint8)



L5_FREQ = 1176.45e6

# signal: code generator, code period (s), chips per period, carrier (Hz),
# PRNs, secondary code, BOC (VE / VL taps and bump-jump)
_LOOP_SIGS = {
    "L1CA": (l1ca_code, 1e-3, 1023, L1_FREQ, range(1, 33), None, False),
    "L5Q": (codes.gen_code_L5Q, 1e-3, 10230, L5_FREQ, range(1, 33), codes.sec_code_NH20, False),
    "E1B": (codes.gen_code_E1B, 4e-3, 4092, L1_FREQ, range(1, 37), None, True),
}


def trk_loop_sig(sig="L5Q", seed=6, fs=12e6, n_ch=3, n_epoch=2200, n_vis=2, cn0=(45.0, 45.0),
                 fd_err=20.0, coff_err=0.1, sp_corr=0.25, nposx=0, bump=None, pos_n=-120.0):
    """Closed-loop case: dict(buff, chans, loops, pos, n_epoch, truth, sig).
    The IF holds the first n_vis channels' satellites (with the signal's
    secondary code, one chip per code period); their channels start fd_err
    Hz and coff_err chips off.  Taps (src/sdr_ch.c:227-244): P, E and L at
    -/+ sp_corr / 2 chip, the noise tap at pos_n samples, for a BOC signal
    (bump: default on) VE and VL at -/+ half a chip, then nposx extra taps
    over +/-2 chips.  The loops carry the code length (entries), Costas, a
    starting C/N0 of 40 dB-Hz, the secondary code and the bump-jump step of
    one code entry (src/sdr_ch.c:551-561)."""
    gen, T, chips, fc, prn_range, sec_fn, boc = _LOOP_SIGS[sig]
    bump = boc if bump is None else bump
    rng = np.random.default_rng(seed)
    N = int(round(fs * T))
    prns = list(prn_range)[:n_ch]
    cd = {p: np.asarray(gen(p), np.int8) for p in prns}
    len_code = len(cd[prns[0]])
    sec = None if sec_fn is None else sec_fn()
    sats = synth.random_sats(rng, prns[:n_vis], cd, T, fc, cn0, (-4000.0, 4000.0))
    for s_ in sats:
        if sec is not None:
            s_["sec"] = sec
    truth = {s_["prn"]: s_ for s_ in sats}
    buff = synth.synth_if(fs, N * n_epoch, sats, seed, bits=2)
    chip = T / chips
    chans, loops = [], []
    for p in prns:
        tr = truth.get(p)
        if tr is not None:
            fd0, coff0 = tr["fd"] + fd_err, tr["tau"] + coff_err * chip
        else:
            fd0, coff0 = float(rng.uniform(-4000, 4000)), float(rng.uniform(0, T))
        chans.append(dict(fs=fs, fi=0.0, fc_sig=fc, T=T, N=N, fd0=fd0, drift=0.0, phi0=0.0,
                          coff0=coff0, code=codes.code_bank(cd[p], T, fs, N), prn=p))
        loops.append(dict(len_code=len_code, costas=1, cn0=40.0, sec_code=sec,
                          bump_step=T / len_code if bump else 0.0))
    sc = chip * fs                                      # samples per chip
    pos = [0.0, -0.5 * sp_corr * sc, 0.5 * sp_corr * sc, pos_n]
    if bump:
        pos += [-0.5 * sc, 0.5 * sc]
    for i in range(nposx):
        pos.append(float(i - (nposx - 1) // 2) / nposx * 4.0 * chip * fs)
    return dict(buff=buff, chans=chans, loops=loops, pos=np.array(pos, np.float64), n_epoch=n_epoch,
                truth=truth, sig=sig)
