"""Parity of the HIP library (through its C ABI) with the pinned CPU oracle.

Integer / byte work (mixing, correlators) must be bit-exact; FP32 FFT work
must agree within the tolerances written in each test (the reference's FFT is
PocketFFT, a different FP32 summation order).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CSCALE = np.float32(1.0) / np.float32(11.2)


def rel(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30)


# ---------------------------------------------------------------- mixing
@pytest.mark.parametrize("case", [
    (1000, 0, 999, 12e6, 3000.0, 0.0),
    (1000, 10, 1, 12e6, -4999.0, 0.7),
    (5000, 4000, 3001, 24e6, 1234.56, 0.123456789),    # ring wrap
    (24000, 333, 24000, 24e6, -2750.25, 0.99),          # wrap, full length
    (4, 3, 3, 4.0, 0.0, 0.0),                            # reference test case
])
def test_mix_carr_bitexact(sdr, orc, case):
    buffN, ix, N, fs, fc, phi = case
    rng = np.random.default_rng(buffN + ix)
    buff = rng.integers(0, 256, buffN).astype(np.uint8)   # includes -8 nibbles
    got = sdr.mix_carr(buff, ix, N, fs, fc, phi)
    ref = orc.mix_carr(buff, ix, N, fs, fc, phi)
    assert np.array_equal(got, ref)


def test_mix_known_answers(sdr):
    # test/src/test_sdr_func.c:170-192
    buff = np.array([0xE1, 0x12, 0x3F, 0xF0], np.uint8)
    assert sdr.mix_carr(buff, 0, 2, 4.0, 0.0, 0.0).tolist() == [[11, -22], [22, 11]]
    assert sdr.mix_carr(buff, 3, 3, 4.0, 0.0, 0.0)[:2].tolist() == [[0, -11], [11, -22]]


# ---------------------------------------------------------------- cpx_mul, FFT
@pytest.mark.parametrize("N", [1, 4, 103, 4096])
def test_cpx_mul_bitexact(sdr, orc, N):
    rng = np.random.default_rng(N)
    a = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    b = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    assert np.array_equal(sdr.cpx_mul(a, b, 0.37).view(np.uint32), orc.cpx_mul(a, b, 0.37).view(np.uint32))


@pytest.mark.parametrize("N", [1, 2, 8, 12, 30, 150, 4000, 4096, 24000, 48000, 96000, 192000, 960000])
def test_cpx_fft(sdr, orc, N):
    rng = np.random.default_rng(N)
    x = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    for d in (0, 1):
        ref = orc.fft_f64(x.astype(np.complex128), d)
        got = sdr.cpx_fft(x, d)
        # FP32 FFT: normwise relative error bound 5e-7 * log2(N) + 1e-7
        assert rel(got, ref) < 5e-7 * max(np.log2(N), 1) + 1e-7


def test_cpx_fft_bad_args(sdr):
    with pytest.raises(ValueError):
        sdr.cpx_fft(np.ones(0, np.complex64), 0)        # empty input
    with pytest.raises(ValueError):
        sdr.cpx_fft(np.ones(8, np.complex64), 2)        # bad direction


# ---------------------------------------------------------------- correlators
def corr_case(N, seed, minus128=True):
    rng = np.random.default_rng(seed)
    IQ = rng.integers(-112, 113, (N, 2)).astype(np.int8)
    if minus128:
        IQ[rng.integers(0, N, 7)] = -128
    code = rng.integers(-1, 2, (10 * N, 2)).astype(np.int8)
    return rng, IQ, code


@pytest.mark.parametrize("N", [37, 1001, 4000, 24000])
@pytest.mark.parametrize("cpx", [False, True])
def test_corr_std_bitexact(sdr, orc, N, cpx):
    rng, IQ, code = corr_case(N, N + 7 * cpx)
    if not cpx:
        code[:, 1] = code[:, 0]
    pos = np.array([0.0, -0.37 * N, 0.21 * N, -0.9 * N, 0.55, -0.95, 0.8 * N] +
                   list(rng.uniform(-0.45 * N, 0.45 * N, 30)))
    for coff in (0.0, 7.25, N - 1.6, 0.5 * N + 0.05):
        got, gC = sdr.corr_std(IQ, code, N, coff, pos, cpx_code=cpx)
        ref, rC = orc.corr_std(IQ, code, N, coff, pos, cpx_code=cpx)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (N, coff)
        assert np.array_equal(gC.view(np.uint32), rC.view(np.uint32))


@pytest.mark.parametrize("N", [119999, 200000])
def test_corr_std_large_sums(sdr, orc, N):
    """IQ matched to the bank (+-112 against +-1 chips), so the prompt's sums
    reach 112 N: beyond 2^24 at N = 200000, where the float adds of the
    scalar tail round and must follow the reference's order; N = 119999 is
    the longest block whose tail joins the integer sums."""
    rng = np.random.default_rng(N)
    code = (2 * rng.integers(0, 2, (10 * N, 1)) - 1).astype(np.int8).repeat(2, axis=1)
    IQ = (112 * code[:N]).astype(np.int8)
    IQ[rng.integers(0, N, 50)] = rng.integers(-112, 113, (50, 2))
    pos = np.array([0.0, -11.7, 11.7, 0.5, -3.0, 17.0, 1.25])
    for coff in (0.0, 0.31, 13.0):
        got, gC = sdr.corr_std(IQ, code, N, coff, pos)
        ref, rC = orc.corr_std(IQ, code, N, coff, pos)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (N, coff)
        assert np.array_equal(gC.view(np.uint32), rC.view(np.uint32))
        if coff == 0.0:
            # the prompt's integer sum (corr N 11.2): past 2^24 at N = 200000
            assert (abs(float(ref[0].real)) * N * 11.2 > 2 ** 24) == (N > 150000)


def test_corr_std_dropin_banks(sdr, orc):
    """The drop-in on the tracking correlator: real-code banks with I != Q,
    five banks in turn (more than the per-thread device cache holds, so some
    are rebuilt), 107 taps (three passes of the workgroup's tap slots), the
    -128 corner in some calls only."""
    N = 24000
    rng = np.random.default_rng(77)
    banks = [rng.integers(-1, 2, (10 * N, 2)).astype(np.int8) for _ in range(5)]
    pos = np.concatenate([[0.0, -11.7, 11.7, -120.0], rng.uniform(-400.0, 400.0, 103)])
    for it in range(12):
        bank = banks[it % 5]
        IQ = rng.integers(-112, 113, (N, 2)).astype(np.int8)
        if it % 3 == 0:
            IQ[rng.integers(0, N, 5)] = -128
        coff = float(rng.uniform(0, N))
        got, gC = sdr.corr_std(IQ, bank, N, coff, pos)
        ref, rC = orc.corr_std(IQ, bank, N, coff, pos)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), it
        assert np.array_equal(gC.view(np.uint32), rC.view(np.uint32)), it


def test_corr_std_cpx(sdr, orc):
    N = 4000
    rng = np.random.default_rng(11)
    buff = (rng.integers(-7, 8, 5000) + 1j * rng.integers(-7, 8, 5000)).astype(np.complex64)
    code = rng.choice([-1.0, 1.0], N).astype(np.float32)
    pos = np.array([0.0, -3.0, 3.0, 10.0])
    got = sdr.corr_std_cpx(buff, 4500, N, 4e6, 1500.0, 0.25, code, pos)
    # composition in the oracle (src/sdr_func.c:1060-1072)
    b8 = ((buff.real.astype(np.int64) & 0xF) | ((buff.imag.astype(np.int64) & 0xF) << 4)).astype(np.uint8)
    idx = (4500 + np.arange(N)) % 5000
    IQ = orc.mix_carr(b8[idx], 0, N, 4e6, 1500.0, 0.25)
    bank = np.zeros((10 * N, 2), np.int8)
    bank[:N, 0] = bank[:N, 1] = code.astype(np.int8)
    ref, _ = orc.corr_std(IQ, bank, N, 0.0, pos)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("M", [8000, 24000, 48000])
def test_corr_fft(sdr, orc, M):
    rng = np.random.default_rng(M)
    IQ = rng.integers(-112, 113, (M, 2)).astype(np.int8)
    c = np.zeros(M, np.complex64)
    c[:M // 2] = rng.choice([-1, 1], M // 2)
    code_fft = np.conj(orc.fft_f32(c, 0))
    got = sdr.corr_fft(IQ, code_fft)
    ref = orc.corr_fft(IQ, code_fft)
    assert rel(got, ref) < 1e-5


def test_corr_fft_zero_code(sdr):
    IQ = np.array([[1, 0], [2, -1], [3, -2], [4, -3]], np.int8)
    assert np.all(np.abs(sdr.corr_fft(IQ, np.zeros(4, np.complex64))) < 1e-6)


def test_search_code_zero_code(sdr):
    # test/src/test_sdr_func.c:244-248
    buff = np.array([0x01, 0xF2, 0xE3, 0xD4], np.uint8)
    P = np.ones((1, 4), np.float32)
    sdr.search_code(np.zeros(4, np.complex64), 1.0, buff, 0, 4.0, 0.0, np.zeros(1, np.float32), P)
    assert np.allclose(P, 1.0, atol=1e-6)


@pytest.mark.parametrize("case", [(4e6, 0, 5), (12e6, 7000, 21)])
def test_search_code(sdr, orc, case):
    from pocketsdr_amd import workloads
    fs, ix, nb = case
    N = int(fs * 1e-3)
    M = 2 * N
    rng = np.random.default_rng(int(fs) + ix)
    buff = rng.integers(0, 256, 3 * N + 500).astype(np.uint8)
    code = workloads.l1ca_code(5)
    code_fft = orc.gen_code_fft(code, 1e-3, 0.0, fs, N, N)
    fds = workloads.dop_bins(1e-3, 0.0, 500.0 * (nb // 2))
    P0 = rng.random((nb, M)).astype(np.float32)
    got = sdr.search_code(code_fft, 1e-3, buff, ix, fs, 1000.0, fds, P0.copy())
    ref = orc.search_code(code_fft, 1e-3, buff, ix, fs, 1000.0, fds, P0.copy())
    # FP32 FFT correlator: power surface within 1e-5 normwise, 1e-3 per element
    # of the peak, same argmax
    assert rel(got - P0, ref - P0) < 1e-5
    assert np.argmax(got) == np.argmax(ref)


# ---------------------------------------------------------------- plans
@pytest.fixture(scope="module")
def small_acq():
    from oracle import oracle as orc
    from pocketsdr_amd import workloads
    return workloads.acq_l1ca(seed=3, fs=4e6, n_prn=6, n_vis=3, t_int=4e-3, max_dop=2000.0,
                              cn0=(45, 50),
                              code_fft_fn=lambda c, T, fs, N: orc.gen_code_fft(c, T, 0.0, fs, N, N))


def oracle_search(orc, w):
    s = w["sig"]
    out = []
    for q in range(s["code_fft"].shape[0]):
        P = np.zeros((len(s["fds"]), s["M"]), np.float32)
        for k in range(s["n_blk"]):
            orc.search_code(s["code_fft"][q], s["T"], w["buff"], s["ix0"] + k * s["blk_step"],
                            s["fs"], s["fi"], s["fds"], P)
        out.append(P)
    return out


@pytest.mark.parametrize("flags", [0, 1, 2])
def test_search_plan(sdr, orc, small_acq, flags):
    w = small_acq
    s = w["sig"]
    ref = oracle_search(orc, w)
    plan = sdr.SearchPlan([w["buff"]], [s], flags)
    for rep in range(2):       # graph replays must match the eager result
        plan.run()
        cn0, ix, dop = plan.result(0)
        for q in range(s["code_fft"].shape[0]):
            P = plan.power(0, q)
            R = ref[q][:, :s["N"]]
            assert rel(P, R) < 1e-5
            rc, rix = orc.corr_max(ref[q], s["M"], s["N"], len(s["fds"]), s["T"])
            assert ix[q].tolist() == rix.tolist()
            assert abs(cn0[q] - rc) <= 1e-3 * abs(rc) + 1e-3
            rd = orc.fine_dop(ref[q], s["M"], s["fds"], rix)
            assert abs(dop[q] - rd) < 0.5
    # the planted satellites are found at their code phase and Doppler
    for sat in w["truth"]:
        q = s["prns"].index(sat["prn"])
        assert cn0[q] > 38.0
        lag = ix[q][1] / s["fs"]
        dl = (lag - sat["tau"] + s["T"] / 2) % s["T"] - s["T"] / 2
        assert abs(dl) < 1.5 / s["fs"]
        assert abs(dop[q] - sat["fd"]) < 250.0
    plan.close()


@pytest.mark.parametrize("variant", ["plain", "exact", "wrap", "iq", "long"])
def test_trk_plan_bitexact(sdr, orc, variant):
    from pocketsdr_amd import workloads
    fs = 24e6 if variant == "long" else 4e6
    w = workloads.trk_l1ca(seed=4, fs=fs, n_ch=5, n_epoch=12, n_vis=2, nposx=8,
                           buff_epochs=(7 if variant == "wrap" else None))
    if variant == "iq":
        # distinct I and Q code planes (the general bank layout)
        rng = np.random.default_rng(9)
        for ch in w["chans"]:
            ch["code"] = ch["code"].copy()
            ch["code"][:, 1] = rng.integers(-1, 2, len(ch["code"])).astype(np.int8)
    buff = w["buff"].copy()
    if variant == "wrap":
        buff = np.concatenate([buff, buff[:777]])        # ring length not a multiple of N
    if variant == "exact":
        buff[::97] = 0x88                                 # -8 nibbles -> -128 products
    plan = sdr.TrkPlan(buff, w["chans"], w["pos"], w["n_epoch"])
    plan.run()
    corr, Cc = plan.result()
    P = plan.prompt()
    for c, ch in enumerate(w["chans"]):
        rc, rC, _ = orc.trk_run(ch, ch["code"], buff, w["n_epoch"], w["pos"])
        assert np.array_equal(corr[c].view(np.uint32), rc.view(np.uint32)), c
        assert np.array_equal(Cc[c].view(np.uint32), rC.view(np.uint32)), c
        rP = orc.trk_stitch(rC, ch["N"])
        assert np.array_equal(P[c].view(np.uint32), rP.view(np.uint32)), c
    plan.close()


@pytest.fixture(scope="module")
def multi_acq():
    from oracle import oracle as orc
    from pocketsdr_amd import workloads
    return workloads.acq_multi(seed=7, prn_limit=3, cn0=(44, 50),
                               code_fft_fn=lambda c, T, fs, N: orc.gen_code_fft(c, T, 0.0, fs, N, N))


@pytest.mark.parametrize("tune", [False, True])
def test_search_plan_multi(sdr, orc, multi_acq, tune):
    """Uneven batched search: L1CA (M 24000), E1B (96000), L5Q (48000) and
    L6D (192000) in one plan, against the oracle's sdr_search_code; with the
    tuner choosing each group's FFT split."""
    w = multi_acq
    plan = sdr.SearchPlan(w["buffs"], w["sigs"], sdr.SEARCH_GRAPH | (sdr.SEARCH_TUNE if tune else 0))
    for i, s in enumerate(w["sigs"]):
        M1, M2, _, _ = plan.tiles(i)
        assert M1 * M2 == s["M"]
    plan.run()
    for i, s in enumerate(w["sigs"]):
        cn0, ix, dop = plan.result(i)
        for q in range(s["code_fft"].shape[0]):
            R = np.zeros((len(s["fds"]), s["M"]), np.float32)
            for k in range(s["n_blk"]):
                orc.search_code(s["code_fft"][q], s["T"], w["buffs"][s["buff"]],
                                s["ix0"] + k * s["blk_step"], s["fs"], s["fi"], s["fds"], R)
            P = plan.power(i, q)
            assert rel(P, R[:, :s["N"]]) < 1e-5, (s["name"], q)
            rc, rix = orc.corr_max(R, s["M"], s["N"], len(s["fds"]), s["T"])
            assert ix[q].tolist() == rix.tolist(), (s["name"], q)
            assert abs(cn0[q] - rc) <= 1e-3 * abs(rc) + 1e-3
        for sat in w["truth"][s["name"]]:
            q = s["prns"].index(sat["prn"])
            assert cn0[q] > 38.0, (s["name"], sat["prn"], cn0[q])
    plan.close()


@pytest.mark.parametrize("n_pos", [1, 3, 4])
def test_trk_plan_few_taps(sdr, orc, n_pos):
    """At most half as many taps as the workgroup's waves (the closed loop
    correlates four): each tap's words are split over several waves and the
    parts' sums combined; bit-exact as with many taps, on a ring whose length
    is not a multiple of N (epochs split at the wrap)."""
    from pocketsdr_amd import workloads
    w = workloads.trk_l1ca(seed=5, fs=24e6, n_ch=4, n_epoch=10, n_vis=2, nposx=8, buff_epochs=7)
    buff = np.concatenate([w["buff"], w["buff"][:777]])
    pos = w["pos"][:n_pos]
    plan = sdr.TrkPlan(buff, w["chans"], pos, w["n_epoch"])
    plan.run()
    corr, Cc = plan.result()
    plan.close()
    for c, ch in enumerate(w["chans"]):
        rc, rC, _ = orc.trk_run(ch, ch["code"], buff, w["n_epoch"], pos)
        assert np.array_equal(corr[c].view(np.uint32), rc.view(np.uint32)), c
        assert np.array_equal(Cc[c].view(np.uint32), rC.view(np.uint32)), c


def test_trk_plan_full_size(sdr, orc):
    """The tracking configuration of the bench at full size (64 channels x
    1000 epochs at 24 Msps, 44 taps): three channels, the first, a middle and
    the last, bit-exact against the oracle over all epochs, and the prompt
    power of every visible satellite's channel well above the others'."""
    from pocketsdr_amd import workloads
    w = workloads.trk_l1ca(seed=2, n_epoch=1000)
    plan = sdr.TrkPlan(w["buff"], w["chans"], w["pos"], w["n_epoch"])
    plan.run()
    corr, Cc = plan.result()
    P = plan.prompt()
    plan.close()
    for c in (0, 37, len(w["chans"]) - 1):
        ch = w["chans"][c]
        rc, rC, _ = orc.trk_run(ch, ch["code"], w["buff"], w["n_epoch"], w["pos"])
        assert np.array_equal(corr[c].view(np.uint32), rc.view(np.uint32)), c
        assert np.array_equal(Cc[c].view(np.uint32), rC.view(np.uint32)), c
    vis = {s["prn"] for s in w["truth"]}
    pw = np.mean(np.abs(P[:, 1:]) ** 2, axis=1)
    on = [pw[c] for c, ch in enumerate(w["chans"]) if ch["prn"] in vis]
    off = [pw[c] for c, ch in enumerate(w["chans"]) if ch["prn"] not in vis]
    assert min(on) > 3 * max(off)
