"""Fused cell inverse of the FP16 search form (one workgroup per PRN and
Doppler bin runs the code-product row pass, the column pass and the power
accumulation of every block in LDS).  Its stages are those of the two-pass
inverse with the same roundings, so the power surfaces must equal the
two-pass form's (PSA_NO_CELL=1) bit for bit; the peak cells and Doppler are
then equal and C/N0 differs only by the grouping of the FP64 power sum.
Cases: the GPS L1 C/A search (12 Msps, 1 ms, 21 bins, 10 blocks, 32 PRNs),
one block, a bin count that leaves most XCD rounds partial, fewer lags than
M / 2, and a surface against the FP64 oracle.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def plans(sdr, buffs, sigs, flags):
    """(cell plan, two-pass plan), both run."""
    pc = sdr.SearchPlan(buffs, sigs, flags)
    os.environ["PSA_NO_CELL"] = "1"
    try:
        pt = sdr.SearchPlan(buffs, sigs, flags)
    finally:
        del os.environ["PSA_NO_CELL"]
    pc.run()
    pt.run()
    return pc, pt


def check_equal(pc, pt, sigs):
    for i, s in enumerate(sigs):
        assert pc.tiles(i)[:2] == (150, 160)
        for q in range(s["code_fft"].shape[0]):
            a, b = pc.power(i, q), pt.power(i, q)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (i, q)
        (c1, i1, d1), (c2, i2, d2) = pc.result(i), pt.result(i)
        assert np.array_equal(i1, i2) and np.array_equal(d1, d2), i
        assert np.allclose(c1, c2, rtol=1e-6, atol=0), i


@pytest.mark.parametrize("graph", [False, True])
def test_cell_l1ca_full(sdr, graph):
    from pocketsdr_amd import workloads
    w = workloads.acq_l1ca(seed=31)
    s = w["sig"]
    assert s["M"] == 24000 and len(s["fds"]) == 21 and s["n_blk"] == 10
    pc, pt = plans(sdr, [w["buff"]], [s], sdr.SEARCH_FP16 | (sdr.SEARCH_GRAPH if graph else 0))
    check_equal(pc, pt, [s])
    # a second run on the same plan gives the same surfaces (P is stored on
    # the first block, not accumulated over runs)
    P0 = pc.power(0, 5).copy()
    pc.run()
    assert np.array_equal(pc.power(0, 5), P0)
    pc.close()
    pt.close()


@pytest.mark.parametrize("nbin,nblk,N", [(1, 1, 12000), (3, 2, 12000), (21, 3, 11001), (13, 1, 7)])
def test_cell_shapes(sdr, nbin, nblk, N):
    rng = np.random.default_rng(nbin * 100 + nblk)
    M, fs, T = 24000, 12e6, 1e-3
    buff = rng.integers(0, 256, 3 * M, dtype=np.uint8)
    # spectra of |.| <= N (the form scales them by 2^-ceil(log2 N))
    cf = (rng.standard_normal((5, M)) + 1j * rng.standard_normal((5, M))) * (0.2 * N ** 0.5)
    fds = (np.arange(nbin) - nbin // 2).astype(np.float32) * 500.0
    s = dict(buff=0, M=M, N=N, T=T, fs=fs, fi=1.5e5, code_fft=cf.astype(np.complex64), fds=fds, n_blk=nblk,
             ix0=1234, blk_step=M // 2 + 7, prns=list(range(5)))
    pc, pt = plans(sdr, [buff], [s], sdr.SEARCH_FP16)
    check_equal(pc, pt, [s])
    pc.close()
    pt.close()


def test_cell_vs_oracle(sdr, orc):
    """FP16 cell surface of a visible satellite within 1e-2 of the FP64 oracle
    at its peak cell, and the same peak cell as the oracle's FP32 search."""
    from pocketsdr_amd import workloads
    w = workloads.acq_l1ca(seed=32, n_vis=3, t_int=3e-3,
                           code_fft_fn=lambda c, T, fs, N: orc.gen_code_fft(c, T, 0.0, fs, N, N))
    s = w["sig"]
    p = sdr.SearchPlan([w["buff"]], [s], sdr.SEARCH_FP16)
    assert p.tiles(0)[:2] == (150, 160)
    p.run()
    cn0, ix, _ = p.result(0)
    for t in w["truth"]:
        q = s["prns"].index(t["prn"])
        R = np.zeros((len(s["fds"]), s["M"]), np.float64)
        R32 = np.zeros((len(s["fds"]), s["M"]), np.float32)
        for k in range(s["n_blk"]):
            orc.search_code_f64(s["code_fft"][q], s["T"], w["buff"], k * s["blk_step"], s["fs"], 0.0, s["fds"], R)
            orc.search_code(s["code_fft"][q], s["T"], w["buff"], k * s["blk_step"], s["fs"], 0.0, s["fds"], R32)
        rc, rix = orc.corr_max(R32, s["M"], s["N"], len(s["fds"]), s["T"])
        if rc >= 38.0:
            assert ix[q].tolist() == rix.tolist(), t["prn"]
        b, lag = ix[q]
        P = p.power(0, q)
        assert abs(P[b, lag] - R[b, lag]) <= 1e-2 * R[b, lag], t["prn"]
        assert np.linalg.norm(P - R[:, :s["N"]]) <= 2e-2 * np.linalg.norm(R[:, :s["N"]])
    p.close()
