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
        assert M1 * M2 == s["M"] and vc in (0, 1) and vr in (0, 1)
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


def test_precision_plain_flag_ignored(sdr, multi_weak):
    """PLAIN keeps the FP32 reference dataflow whatever the precision flags."""
    w = multi_weak
    p = sdr.SearchPlan(w["buffs"][:1], [w["sigs"][0]], sdr.SEARCH_PLAIN | sdr.SEARCH_FP16)
    assert p.form(0) == sdr.FORM_PLAIN
    p.close()
