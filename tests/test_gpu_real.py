"""Acquisition on real signals: the reference's own captures (slices in
tests/golden, see make_sample_slices.py) searched on the GPU as pocket_acq
searches them (app/pocket_acq/pocket_acq.c:152-192: blocks of one code period
over T_AQC, Doppler bins of 1/(2T) over +/-5 kHz, IF 3 MHz for the
I-sampled L1 file).

One plan over the L1 capture holds GPS L1 C/A (32 PRNs), Galileo E1-B (36)
and GPS L1C-P (PRNs 11, 18 and 1); a second plan over the L6 capture holds
QZSS L6D (PRNs 193-199).  FP32 batched against the oracle for every PRN: the
power surface within 1e-5 normwise, the same peak (bin, lag), C/N0 within
1e-3 relative, Doppler within 0.5 Hz.  FP16 precision form: the same
detections (C/N0 >= 38 dB-Hz) on the same Doppler bins, code phase within
one sample (the same sample for signals of 42 dB-Hz or more).  The drop-in
sdr_search_code on the L1 C/A satellites in view: within 1e-5 of the oracle.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


def group(orc, name, prns, buf, fs, fi, T, t_aqc):
    from pocketsdr_amd import workloads
    N = int(round(fs * T))
    nb = int(round(fs * (t_aqc + T)))
    cf = np.stack([orc.gen_code_fft(workloads.sig_code(name, p), T, 0.0, fs, N, N) for p in prns])
    return dict(name=name, buff=0, M=2 * N, N=N, T=T, fs=fs, fi=fi, code_fft=cf.astype(np.complex64),
                fds=workloads.dop_bins(T, 0.0, 5000.0), n_blk=len(range(0, nb - 2 * N + 1, N)), ix0=0,
                blk_step=N, prns=list(prns))


@pytest.fixture(scope="module")
def captures(orc):
    from pocketsdr_amd import workloads
    meta = json.load(open(os.path.join(GOLD, "slices.json")))
    l1n, l6n = "L1_20211125_004000_12MHz_I_30ms.bin", "L6_20211125_004000_12MHz_IQ_14ms.bin"
    l1 = workloads.read_if(os.path.join(GOLD, l1n), meta[l1n]["IQ"])
    l6 = workloads.read_if(os.path.join(GOLD, l6n), meta[l6n]["IQ"])
    fs, fi = meta[l1n]["fs"], meta[l1n]["fi"]
    sig_l1 = [group(orc, "L1CA", range(1, 33), l1, fs, fi, 1e-3, 0.010),
              group(orc, "E1B", range(1, 37), l1, fs, fi, 4e-3, 0.010),
              group(orc, "L1CP", (11, 18, 1), l1, fs, fi, 10e-3, 0.020)]
    sig_l6 = [group(orc, "L6D", range(193, 200), l6, meta[l6n]["fs"], meta[l6n]["fi"], 4e-3, 0.010)]
    return [(l1, sig_l1), (l6, sig_l6)]


def oracle_P(orc, buf, s, q):
    R = np.zeros((len(s["fds"]), s["M"]), np.float32)
    for k in range(s["n_blk"]):
        orc.search_code(s["code_fft"][q], s["T"], buf, s["ix0"] + k * s["blk_step"], s["fs"], s["fi"], s["fds"], R)
    return R


def test_real_signals(sdr, orc, captures):
    found = {}
    for buf, sigs in captures:
        p32 = sdr.SearchPlan([buf], sigs, sdr.SEARCH_BATCHED | sdr.SEARCH_GRAPH)
        ph = sdr.SearchPlan([buf], sigs, sdr.SEARCH_FP16 | sdr.SEARCH_GRAPH)
        p32.run()
        ph.run()
        for i, s in enumerate(sigs):
            cn0, ix, dop = p32.result(i)
            for q in range(len(s["prns"])):
                R = oracle_P(orc, buf, s, q)
                assert rel(p32.power(i, q), R[:, :s["N"]]) < 1e-5, (s["name"], s["prns"][q])
                rc, rix = orc.corr_max(R, s["M"], s["N"], len(s["fds"]), s["T"])
                assert ix[q].tolist() == rix.tolist(), (s["name"], s["prns"][q])
                assert abs(cn0[q] - rc) <= 1e-3 * abs(rc), (s["name"], s["prns"][q])
                assert abs(dop[q] - orc.fine_dop(R, s["M"], s["fds"], rix)) <= 0.5
            c16, i16, _ = ph.result(i)
            det = cn0 >= 38.0
            assert np.array_equal(c16 >= 38.0, det), s["name"]
            for q in np.flatnonzero(det):
                assert i16[q][0] == ix[q][0] and abs(int(i16[q][1]) - int(ix[q][1])) <= 1, (s["name"], q)
                if cn0[q] >= 42.0:
                    assert i16[q][1] == ix[q][1], (s["name"], s["prns"][q])
            found[s["name"]] = {s["prns"][q]: (round(float(cn0[q]), 2), float(dop[q])) for q in np.flatnonzero(det)}
        p32.close()
        ph.close()
    print("found: %s" % found)
    # the satellites in view (the oracle's detections on these captures)
    assert {5, 13, 15, 20, 30} <= set(found["L1CA"])
    assert {3, 8, 13, 15} <= set(found["E1B"])
    assert {11, 18} <= set(found["L1CP"])
    assert {193, 196, 199} <= set(found["L6D"])
    for p in (11, 18):      # L1C-P at the Doppler of the same satellite's L1 C/A
        assert abs(found["L1CP"][p][1] - found["L1CA"][p][1]) <= 50.0


def test_real_signals_dropin(sdr, orc, captures):
    buf, sigs = captures[0]
    s = sigs[0]
    for p in (5, 13, 15, 20, 30):
        q = s["prns"].index(p)
        P = None
        for k in range(s["n_blk"]):
            P = sdr.search_code(s["code_fft"][q], s["T"], buf, k * s["blk_step"], s["fs"], s["fi"], s["fds"], P)
        R = oracle_P(orc, buf, s, q)
        assert rel(P, R) < 1e-5, p
        cn0, ix = sdr.corr_max(P, s["M"], s["N"], len(s["fds"]), s["T"])
        rc, rix = orc.corr_max(R, s["M"], s["N"], len(s["fds"]), s["T"])
        assert ix.tolist() == rix.tolist() and abs(cn0 - rc) <= 1e-3 * abs(rc), p
