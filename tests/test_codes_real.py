"""Spreading-code generators pinned by real signals: the reference's own
captures (sample/L1_20211125_004000_12MHz_I.bin, L6_20211125_004000_12MHz_IQ
.bin; slices in tests/golden, cut by make_sample_slices.py) searched with the
oracle's sdr_search_code loop as pocket_acq runs it
(app/pocket_acq/pocket_acq.c:152-192, T_AQC = 10 ms plus one code period).
A wrong chip, register tap, table entry or modulation leaves no correlation
peak, so a generator that finds the satellites in the sky is right:

* Galileo E1-B (memory codes, BOC(1,1)): E3, E8, E13, E15 at 44-47 dB-Hz;
* GPS L1C-P (Weil codes, insertion, TMBOC): PRN 11 and 18, at the Doppler of
  their own L1 C/A signals in the same capture (one 50 Hz bin);
* QZSS L6D (Kasami-type registers, TDM): PRN 193, 196, 199 at 42-47 dB-Hz
  with the near-zero Doppler of the quasi-zenith orbits;
* GPS L1 C/A (IS-GPS-200 table, pinned also by the ICD's first chips):
  PRN 5, 13, 15, 20, 30;
and PRNs of the same systems that are not in view stay below 38 dB-Hz.
Also: the slices are the recorded bytes (SHA-256 of slices.json).
"""
import hashlib
import json
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def slice_buf(name):
    from pocketsdr_amd import workloads
    meta = json.load(open(os.path.join(GOLD, "slices.json")))[name]
    return workloads.read_if(os.path.join(GOLD, name), meta["IQ"]), meta


def search(orc, code, buf, fs, fi, T, t_aqc=0.010):
    from pocketsdr_amd import workloads
    N = int(round(fs * T))
    M = 2 * N
    cf = orc.gen_code_fft(code, T, 0.0, fs, N, N)
    fds = workloads.dop_bins(T, 0.0, 5000.0)
    b = buf[:int(round(fs * (t_aqc + T)))]
    P = np.zeros((len(fds), M), np.float32)
    for i in range(0, len(b) - 2 * N + 1, N):
        orc.search_code(cf, T, b, i, fs, fi, fds, P)
    cn0, ix = orc.corr_max(P, M, N, len(fds), T)
    return cn0, ix, orc.fine_dop(P, M, fds, ix)


def test_slices_are_the_recorded_bytes():
    meta = json.load(open(os.path.join(GOLD, "slices.json")))
    for name, m in meta.items():
        data = open(os.path.join(GOLD, name), "rb").read()
        assert len(data) == m["bytes"] and hashlib.sha256(data).hexdigest() == m["sha256"]


@pytest.mark.parametrize("sig,T,visible,absent", [
    ("L1CA", 1e-3, [5, 13, 15, 20, 30], [1, 4]),
    ("E1B", 4e-3, [3, 8, 13, 15], [1, 4]),
])
def test_l1_capture(orc, sig, T, visible, absent):
    from pocketsdr_amd import workloads
    buf, meta = slice_buf("L1_20211125_004000_12MHz_I_30ms.bin")
    for prn in visible + absent:
        cn0, ix, dop = search(orc, workloads.sig_code(sig, prn), buf, meta["fs"], meta["fi"], T)
        if prn in visible:
            assert cn0 >= 42.0, (sig, prn, cn0)
        else:
            assert cn0 < 38.0, (sig, prn, cn0)


def test_l1cp_at_the_l1ca_doppler(orc):
    from pocketsdr_amd import workloads
    buf, meta = slice_buf("L1_20211125_004000_12MHz_I_30ms.bin")
    for prn in (11, 18):
        _, _, dop_ca = search(orc, workloads.sig_code("L1CA", prn), buf, meta["fs"], meta["fi"], 1e-3)
        cn0, ix, dop = search(orc, workloads.sig_code("L1CP", prn), buf, meta["fs"], meta["fi"], 10e-3,
                              t_aqc=0.020)
        assert cn0 >= 37.0, (prn, cn0)
        assert abs(dop - dop_ca) <= 50.0, (prn, dop, dop_ca)
    cn0, _, _ = search(orc, workloads.sig_code("L1CP", 1), buf, meta["fs"], meta["fi"], 10e-3, t_aqc=0.020)
    assert cn0 < 36.0


def test_l6_capture(orc):
    from pocketsdr_amd import workloads
    buf, meta = slice_buf("L6_20211125_004000_12MHz_IQ_14ms.bin")
    for prn in (193, 196, 199, 194):
        cn0, ix, dop = search(orc, workloads.sig_code("L6D", prn), buf, meta["fs"], meta["fi"], 4e-3)
        if prn == 194:
            assert cn0 < 36.0
        else:
            assert cn0 >= 42.0 and abs(dop) < 500.0, (prn, cn0, dop)
