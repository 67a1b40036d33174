"""Weak-signal wide search shape: GPS L1C-P, 10 ms coherent at 48 Msps
(N = 480000, FFT length M = 960000), 3-bit IF on two RF channels of an
antenna array (two signal groups over two IF streams).  Reduced sizes: three
PRNs, the three Doppler bins around each visible satellite, two blocks.

FP32 batched against the oracle's sdr_search_code (src/sdr_func.c:664-687):
power surfaces within 1e-5 normwise, identical peak (bin, lag).  FP16 MFMA
form: peak power within 1e-2 of the FP64 oracle and the same peak cell as
FP32 for every visible satellite.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


@pytest.fixture(scope="module")
def l1cp():
    from oracle import oracle as orc
    from pocketsdr_amd import workloads
    return workloads.acq_l1cp(seed=9, n_rf=2, n_blk=2, prn_limit=3, n_vis=2, bin_limit=3, cn0=(40.0, 45.0),
                              code_fft_fn=lambda c, T, fs, N: orc.gen_code_fft(c, T, 0.0, fs, N, N))


def test_l1cp_fp32_vs_oracle(sdr, orc, l1cp):
    w = l1cp
    plan = sdr.SearchPlan(w["buffs"], w["sigs"], sdr.SEARCH_GRAPH)
    plan.run()
    for i, s in enumerate(w["sigs"]):
        assert plan.form(i) == sdr.FORM_FP32
        cn0, ix, dop = plan.result(i)
        for q in range(s["code_fft"].shape[0]):
            R = np.zeros((len(s["fds"]), s["M"]), np.float32)
            for k in range(s["n_blk"]):
                orc.search_code(s["code_fft"][q], s["T"], w["buffs"][s["buff"]], s["ix0"] + k * s["blk_step"],
                                s["fs"], s["fi"], s["fds"], R)
            P = plan.power(i, q)
            assert rel(P, R[:, :s["N"]]) < 1e-5, (s["name"], q, rel(P, R[:, :s["N"]]))
            rc, rix = orc.corr_max(R, s["M"], s["N"], len(s["fds"]), s["T"])
            assert ix[q].tolist() == rix.tolist(), (s["name"], q)
        for sat in w["truth"][s["name"]]:
            q = s["prns"].index(sat["prn"])
            lag = ix[q][1] / s["fs"]
            dl = (lag - sat["tau"] + s["T"] / 2) % s["T"] - s["T"] / 2
            # 2 samples plus the code Doppler accumulated over the blocks
            tol = 2.0 / s["fs"] + abs(sat["fd"]) / sat["fc"] * (s["n_blk"] + 1) * s["T"]
            assert abs(dl) < tol, (s["name"], sat["prn"], dl)
            assert abs(dop[q] - sat["fd"]) < 50.0
    plan.close()


# FP16 tilings of M = 960000 (M1 / column shape / row shape): the default
# (columns of 960 in 32-column workgroups, rows of 1000 as 10 x 10 x 10), the
# 16-column workgroups paired on one XCD, rows of 1000 as 25 x 40, and
# columns of 1000 as 25 x 40 (16- and 32-column workgroups) with rows of 960
# (8- and 16-row workgroups)
TILINGS = [None, "960/1/1", "960/0/1", "1000/1/0", "1000/0/1", "1000/0/0", "1000/0/2", "1000/0/3"]


@pytest.mark.parametrize("tiles", TILINGS)
def test_l1cp_fp16_vs_fp64(sdr, orc, l1cp, tiles, monkeypatch):
    w = l1cp
    if tiles:
        monkeypatch.setenv("PSA_TILES", "h960000=" + tiles)
    else:
        monkeypatch.delenv("PSA_TILES", raising=False)
    p32 = sdr.SearchPlan(w["buffs"], w["sigs"], sdr.SEARCH_BATCHED)
    ph = sdr.SearchPlan(w["buffs"], w["sigs"], sdr.SEARCH_FP16 | sdr.SEARCH_GRAPH)
    p32.run()
    ph.run()
    worst = 0.0
    for i, s in enumerate(w["sigs"]):
        assert ph.form(i) == sdr.FORM_FP16
        if tiles:
            M1, vc, vr = (int(x) for x in tiles.split("/"))
            assert ph.tiles(i) == (M1, s["M"] // M1, vc, vr)
        _, ix32, _ = p32.result(i)
        _, ixh, _ = ph.result(i)
        for q in range(s["code_fft"].shape[0]):
            b, lag = ix32[q]
            ref = np.zeros((1, s["M"]), np.float64)
            for k in range(s["n_blk"]):
                orc.search_code_f64(s["code_fft"][q], s["T"], w["buffs"][s["buff"]], s["ix0"] + k * s["blk_step"],
                                    s["fs"], s["fi"], s["fds"][b:b + 1], ref)
            Ph = ph.power(i, q)
            err = abs(float(Ph[b, lag]) - ref[0, lag]) / ref[0, lag]
            worst = max(worst, err)
            assert err < 1e-2, (s["name"], q, err)
            assert rel(Ph, p32.power(i, q)) < 2e-2
        for sat in w["truth"][s["name"]]:
            q = s["prns"].index(sat["prn"])
            assert ixh[q].tolist() == ix32[q].tolist(), (s["name"], sat["prn"])
    print("L1CP fp16 (%s) worst peak-power error vs FP64: %.2e" % (tiles, worst))
    p32.close()
    ph.close()


@pytest.fixture(scope="module")
def l1cp3():
    from oracle import oracle as orc
    from pocketsdr_amd import workloads
    return workloads.acq_l1cp(seed=4, n_rf=1, n_blk=3, prn_limit=2, n_vis=1, bin_limit=2, cn0=(40.0, 45.0),
                              code_fft_fn=lambda c, T, fs, N: orc.gen_code_fft(c, T, 0.0, fs, N, N))


@pytest.mark.parametrize("tiles", ["960/0/0", "1000/0/0"])
def test_l1cp_paired_blocks_bitexact(sdr, l1cp3, tiles, monkeypatch):
    """The power pass of the 32-column shapes takes two blocks per launch
    (the first block's powers kept in registers, P read and written once per
    pair; three blocks: one pair and one single): P, peaks, C/N0 and Doppler
    bit-identical to one block per launch (PSA_NO_PAIR)."""
    w = l1cp3
    monkeypatch.setenv("PSA_TILES", "h960000=" + tiles)
    monkeypatch.delenv("PSA_NO_PAIR", raising=False)
    pp = sdr.SearchPlan(w["buffs"], w["sigs"], sdr.SEARCH_FP16 | sdr.SEARCH_GRAPH)
    monkeypatch.setenv("PSA_NO_PAIR", "1")
    p1 = sdr.SearchPlan(w["buffs"], w["sigs"], sdr.SEARCH_FP16 | sdr.SEARCH_GRAPH)
    for rep in range(2):
        pp.run()
        p1.run()
        for i, s in enumerate(w["sigs"]):
            assert pp.tiles(i)[0] == int(tiles.split("/")[0])
            for q in range(s["code_fft"].shape[0]):
                assert np.array_equal(pp.power(i, q), p1.power(i, q)), (i, q)
            ra, rb = pp.result(i), p1.result(i)
            for x, y in zip(ra, rb):
                assert np.array_equal(x, y)
    pp.close()
    p1.close()
