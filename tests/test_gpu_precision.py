"""Precision form of the batched search (FP16 / BF16 words, MFMA DFT-matrix
stages, FP32 accumulation): peak power against the FP64 oracle and the
acquisition decisions against the FP32 batched path.

Tolerances: peak power within 1e-2 relative of the FP64 restatement of
sdr_search_code (src/sdr_func.c:664-687); the whole power surface within
2e-2 (FP16) / 5e-2 (BF16) normwise of the FP32 path; detections (C/N0 >= 38
dB-Hz), code offset and Doppler bin of every detection identical to FP32.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CN0_THRES = 38.0


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


@pytest.fixture(scope="module")
def multi_weak():
    """L1CA / E1B / L5Q / L6D, 4 PRNs each, C/N0 swept 30-50 dB-Hz."""
    from oracle import oracle as orc
    from pocketsdr_amd import workloads
    return workloads.acq_multi(seed=11, prn_limit=4, cn0=(30.0, 50.0),
                               code_fft_fn=lambda c, T, fs, N: orc.gen_code_fft(c, T, 0.0, fs, N, N))


def peak_f64(orc, w, s, q, b):
    """FP64 oracle power of PRN q over the blocks at Doppler bin b (all lags)."""
    P = np.zeros((1, s["M"]), np.float64)
    for k in range(s["n_blk"]):
        orc.search_code_f64(s["code_fft"][q], s["T"], w["buffs"][s["buff"]], s["ix0"] + k * s["blk_step"],
                            s["fs"], s["fi"], s["fds"][b:b + 1], P)
    return P[0]


# non-default splits and workgroup shapes, fixed through PSA_TILES
FIXED = {24000: (150, 1, 1), 96000: (200, 1, 1), 48000: (200, 0, 1), 192000: (200, 1, 0)}


@pytest.mark.parametrize("form", ["fp16", "bf16", "fp16-tuned", "fp16-fixed"])
def test_precision_form(sdr, orc, multi_weak, form, monkeypatch):
    w = multi_weak
    flag = sdr.SEARCH_BF16 if form == "bf16" else sdr.SEARCH_FP16
    if form.endswith("tuned"):
        flag |= sdr.SEARCH_TUNE        # tiles chosen by timing the candidates
    if form.endswith("fixed"):
        monkeypatch.setenv("PSA_TILES", ",".join("h%d=%d/%d/%d" % ((M,) + t) for M, t in FIXED.items()))
    else:
        monkeypatch.delenv("PSA_TILES", raising=False)
    want = sdr.FORM_BF16 if form == "bf16" else sdr.FORM_FP16
    surf_tol = 5e-2 if form == "bf16" else 2e-2
    p32 = sdr.SearchPlan(w["buffs"], w["sigs"], sdr.SEARCH_BATCHED)
    ph = sdr.SearchPlan(w["buffs"], w["sigs"], flag | sdr.SEARCH_GRAPH)
    p32.run()
    for rep in range(2):                 # graph replays
        ph.run()
    worst = 0.0
    for i, s in enumerate(w["sigs"]):
        assert ph.form(i) == want, s["name"]
        M1, M2, vc, vr = ph.tiles(i)
        assert M1 * M2 == s["M"] and 0 <= vc < 4 and 0 <= vr < 4
        if form.endswith("fixed"):
            # (M = 24000 runs the fused cell kernel in FP16: one split, no
            # workgroup shapes to choose)
            want_t = (150, 0, 0) if s["M"] == 24000 else FIXED[s["M"]]
            assert (M1, vc, vr) == want_t, (s["name"], ph.tiles(i))
        c32, ix32, d32 = p32.result(i)
        ch, ixh, dh = ph.result(i)
        det32 = c32 >= CN0_THRES
        deth = ch >= CN0_THRES
        assert np.array_equal(det32, deth), (s["name"], c32, ch)
        for q in range(s["code_fft"].shape[0]):
            P32 = p32.power(i, q)
            Ph = ph.power(i, q)
            assert np.all(np.isfinite(Ph))
            assert rel(Ph, P32) < surf_tol, (s["name"], q, rel(Ph, P32))
            b, lag = ix32[q]
            ref = peak_f64(orc, w, s, q, b)
            err = abs(float(Ph[b, lag]) - ref[lag]) / ref[lag]
            worst = max(worst, err)
            assert err < 1e-2, (s["name"], q, err)
            if det32[q]:
                assert ixh[q].tolist() == ix32[q].tolist(), (s["name"], q)
                assert abs(dh[q] - d32[q]) < 5.0
                assert abs(ch[q] - c32[q]) < 0.05
        for sat in w["truth"][s["name"]]:
            q = s["prns"].index(sat["prn"])
            if sat["cn0"] >= 44.0:
                assert deth[q], (s["name"], sat["prn"], ch[q])
    print("%s worst peak-power error vs FP64: %.2e" % (form, worst))
    p32.close()
    ph.close()


# column lengths of the MFMA splits (the row length is M / M1)
SPLITS = {96000: (200, 300, 400), 48000: (150, 200, 300), 192000: (200, 400)}


@pytest.mark.parametrize("name", ["E1B", "L5Q", "L6D"])
def test_every_shape(sdr, orc, multi_weak, name, monkeypatch):
    """Every workgroup shape (radix split, workgroup size) of every split of
    the group's FFT length, fixed through PSA_TILES: power surface within
    2e-2 of FP32, peak power of each PRN within 1e-2 of FP64, the FP32 peak
    cell of each detection."""
    w = multi_weak
    i = [k for k, s in enumerate(w["sigs"]) if s["name"] == name][0]
    s = w["sigs"][i]
    M = s["M"]
    monkeypatch.delenv("PSA_TILES", raising=False)
    p32 = sdr.SearchPlan(w["buffs"], [s], sdr.SEARCH_BATCHED)
    p32.run()
    c32, ix32, _ = p32.result(0)
    P32 = [p32.power(0, q) for q in range(len(s["prns"]))]
    ref = [peak_f64(orc, w, s, q, ix32[q][0])[ix32[q][1]] for q in range(len(s["prns"]))]
    p32.close()
    done = []
    for M1 in SPLITS[M]:
        for vc in range(4):
            for vr in range(4):
                monkeypatch.setenv("PSA_TILES", "h%d=%d/%d/%d" % (M, M1, vc, vr))
                ph = sdr.SearchPlan(w["buffs"], [s], sdr.SEARCH_FP16)
                if ph.tiles(0) != (M1, M // M1, vc, vr):     # no such shape
                    ph.close()
                    continue
                ph.run()
                ch, ixh, _ = ph.result(0)
                for q in range(len(s["prns"])):
                    Ph = ph.power(0, q)
                    assert rel(Ph, P32[q]) < 2e-2, (M1, vc, vr, q, rel(Ph, P32[q]))
                    b, lag = ix32[q]
                    assert abs(float(Ph[b, lag]) - ref[q]) / ref[q] < 1e-2, (M1, vc, vr, q)
                    if c32[q] >= CN0_THRES:
                        assert ixh[q].tolist() == ix32[q].tolist(), (M1, vc, vr, q)
                ph.close()
                done.append((M1, vc, vr))
    assert len(done) >= 2 * len(SPLITS[M]), done
    print("%s: %d shapes" % (name, len(done)))


@pytest.mark.parametrize("name,M,tiles", [("L6D", 192000, (400, 0, 1)), ("E1B", 96000, (400, 0, 1)),
                                          ("L5Q", 48000, (300, 0, 0))])
def test_paired_blocks_bitexact(sdr, multi_weak, monkeypatch, name, M, tiles):
    """The power passes of 400 (16 x 25) and 300 (15 x 20) with 32 columns
    take two blocks per launch (L6D and E1-B: both blocks, L5-Q: five
    pairs): P, peaks, C/N0 and Doppler bit-identical to one block per
    launch (PSA_NO_PAIR)."""
    w = multi_weak
    s = [x for x in w["sigs"] if x["name"] == name][0]
    assert s["n_blk"] >= 2
    monkeypatch.setenv("PSA_TILES", "h%d=%d/%d/%d" % ((M,) + tiles))
    monkeypatch.delenv("PSA_NO_PAIR", raising=False)
    pp = sdr.SearchPlan(w["buffs"], [s], sdr.SEARCH_FP16 | sdr.SEARCH_GRAPH)
    monkeypatch.setenv("PSA_NO_PAIR", "1")
    p1 = sdr.SearchPlan(w["buffs"], [s], sdr.SEARCH_FP16 | sdr.SEARCH_GRAPH)
    for rep in range(2):
        pp.run()
        p1.run()
        assert pp.tiles(0) == (tiles[0], M // tiles[0], tiles[1], tiles[2])
        for q in range(len(s["prns"])):
            assert np.array_equal(pp.power(0, q), p1.power(0, q)), q
        for x, y in zip(pp.result(0), p1.result(0)):
            assert np.array_equal(x, y)
    pp.close()
    p1.close()


def test_precision_plain_flag_ignored(sdr, multi_weak):
    """PLAIN keeps the FP32 reference dataflow whatever the precision flags."""
    w = multi_weak
    p = sdr.SearchPlan(w["buffs"][:1], [w["sigs"][0]], sdr.SEARCH_PLAIN | sdr.SEARCH_FP16)
    assert p.form(0) == sdr.FORM_PLAIN
    p.close()
