"""Pins the CPU oracle (oracle/sdr_oracle.c) before it is trusted as the
checker of the HIP library.

Each oracle function is checked against (a) the known answers of the
reference's own unit tests (test/src/test_sdr_func.c, test/src/test_sdr_code.c),
(b) the reference PocketFFT built from the reference sources (oracle/_ref),
when present, and (c) what the mathematics fixes: brute-force DFTs, the
time-domain definition of the FFT correlator, closed forms for delta codes,
the textbook L1 C/A first-chip table, and the documented semantics of the
AVX2 instruction the reference's correlator uses.

Pinning summary (function -> checks):
  mix table / mix_carr  : ref known answers (test_sdr_func.c:170-192); complex
                          product closed form; carrier quadrature values
  dot / corr_std        : brute-force split-sum formulation; _mm256_sign_epi8
                          definition for the -128 corner; self-correlation
                          closed form
  cpx_mul               : ref known answers (test_sdr_func.c:78-82); FP64 product
  fft_f64 / fft_f32     : brute-force DFT; PocketFFT (oracle/_ref); inverse
  corr_fft              : ref zero-code answer (test_sdr_func.c:235-243);
                          time-domain correlation definition; delta-code shift
  search_code           : ref zero-code answer (test_sdr_func.c:244-248);
                          time-domain definition summed over bins
  corr_max, fine_dop,
  dop_bins              : ref known answers (test_sdr_func.c:255-280); closed forms
  res_code              : ref known answers (test_sdr_code.c:117-136)
  gen_code_L1CA         : IS-GPS-200 first-10-chip octal table, balance
  trk_run               : composition of pinned mix_carr + corr_std with the
                          open-loop NCO recurrence checked in closed form
"""
import numpy as np
import pytest

CSCALE = np.float32(1.0) / np.float32(11.2)


def pack(I, Q):
    return ((np.asarray(I) & 0xF) | ((np.asarray(Q) & 0xF) << 4)).astype(np.uint8)


# ---------------------------------------------------------------- mixing
def test_mix_known_answers(orc):
    # test/src/test_sdr_func.c:170-192
    buff = pack(np.array([1, 2, -1, 0]), np.array([-2, 1, 3, -1]))
    IQ = orc.mix_carr(buff, 0, 2, 4.0, 0.0, 0.0)
    assert IQ.tolist() == [[11, -22], [22, 11]]
    IQ = orc.mix_carr(buff, 3, 3, 4.0, 0.0, 0.0)
    assert IQ[:2].tolist() == [[0, -11], [11, -22]]


def test_mix_table_closed_form(orc):
    tbl = orc.mix_tbl()
    # carrier quadrature points: round(11.2 e^{-2 pi i k/256})
    j1 = 0x01  # cpx8 (1, 0)
    carr = tbl[(j1 << 8) + np.arange(256)]
    assert carr[0].tolist() == [11, 0]
    assert carr[64].tolist() == [0, -11]
    assert carr[128].tolist() == [-11, 0]
    assert carr[192].tolist() == [0, 11]
    assert carr[32].tolist() == [8, -8]
    mag = np.hypot(carr[:, 0], carr[:, 1])
    assert np.all(np.abs(mag - 11.2) < 0.75)
    # every entry is the complex product cpx8(j) * carrier, wrapped to int8
    j = np.arange(256)
    sI = ((j << 4) & 0xFF).astype(np.uint8).view(np.int8) >> 4
    sQ = (j.astype(np.uint8).view(np.int8)) >> 4
    s = (sI.astype(np.int64) + 1j * sQ.astype(np.int64))
    c = carr[:, 0].astype(np.int64) + 1j * carr[:, 1].astype(np.int64)
    prod = s[:, None] * c[None, :]
    wrapI = ((np.real(prod).astype(np.int64) + 128) % 256 - 128).reshape(-1)
    wrapQ = ((np.imag(prod).astype(np.int64) + 128) % 256 - 128).reshape(-1)
    assert np.array_equal(tbl[:, 0], wrapI)
    assert np.array_equal(tbl[:, 1], wrapQ)
    # the only wrapping entries involve a -8 component
    over = (np.abs(np.real(prod)) > 127) | (np.abs(np.imag(prod)) > 127)
    rows = np.nonzero(over.any(axis=1))[0]
    assert all(sI[r] == -8 or sQ[r] == -8 for r in rows)


def test_mix_phase_and_wrap(orc):
    rng = np.random.default_rng(1)
    buff = rng.integers(0, 256, 1000).astype(np.uint8)
    fs, fc, phi = 12e6, 4321.5, 0.37
    IQ = orc.mix_carr(buff, 100, 500, fs, fc, phi)
    tbl = orc.mix_tbl()
    p = np.uint64(int((phi - np.floor(phi)) * 2.0 ** 32))
    s = np.uint64(np.uint32(np.int64(int(fc / fs * 2.0 ** 32)) & 0xFFFFFFFF))
    ph = ((p + s * np.arange(500, dtype=np.uint64)) & np.uint64(0xFFFFFFFF)) >> np.uint64(24)
    idx = (buff[100:600].astype(np.int64) << 8) + ph.astype(np.int64)
    assert np.array_equal(IQ, tbl[idx])
    # ring split: the second segment restarts from phi + step * n
    IQ2 = orc.mix_carr(buff, 900, 300, fs, fc, phi)
    assert np.array_equal(IQ2[:100], orc.mix_carr(buff[900:], 0, 100, fs, fc, phi))
    from fractions import Fraction
    phi2 = float(Fraction(fc / fs) * 100 + Fraction(phi))    # fma(n, step, phi)
    assert np.array_equal(IQ2[100:], orc.mix_carr(buff, 0, 200, fs, fc, phi2))


# ---------------------------------------------------------------- correlator
def sign_epi8(a, b):
    """_mm256_sign_epi8 per its definition: b < 0 -> -a (two's complement,
    so -(-128) = -128), b == 0 -> 0, else a."""
    a = np.asarray(a, np.int64)
    b = np.asarray(b, np.int64)
    neg = ((-a + 128) % 256) - 128
    return np.where(b < 0, neg, np.where(b == 0, 0, a))


def brute_corr_std(IQ, code, N, coff, pos, cpx=False):
    IQ = IQ.astype(np.int64)
    code = code.astype(np.int64)
    c1s, c2s = [], []
    for p0 in pos:
        p = coff + p0
        j = int(np.floor(p))
        k = int((p - j) * 10)
        if j < 0:
            j += N
        elif j >= N:
            j -= N
        bank = code[k * N:(k + 1) * N]
        res = []
        for d, c in ((IQ[:j], bank[N - j:]), (IQ[j:], bank[:N - j])):
            V = len(d) // 16 * 16
            if cpx:
                vr = sign_epi8(d[:V, 0], c[:V, 0]) + sign_epi8(d[:V, 1], c[:V, 1])
                vi = sign_epi8(d[:V, 1], c[:V, 0]) + sign_epi8(d[:V, 0], -c[:V, 1])
                tr = d[V:, 0] * c[V:, 0] + d[V:, 1] * c[V:, 1]
                ti = d[V:, 1] * c[V:, 0] - d[V:, 0] * c[V:, 1]
            else:
                vr = sign_epi8(d[:V, 0], c[:V, 0])
                vi = sign_epi8(d[:V, 1], c[:V, 1])
                tr = d[V:, 0] * c[V:, 0]
                ti = d[V:, 1] * c[V:, 1]
            a = np.float32(int(vr.sum()))
            b = np.float32(int(vi.sum()))
            for x, y in zip(tr, ti):
                a = np.float32(a + np.float32(int(x)))
                b = np.float32(b + np.float32(int(y)))
            res.append((np.float32(a * CSCALE), np.float32(b * CSCALE)))
        c1s.append(res[0])
        c2s.append(res[1])
    dot = np.float32(0)
    # taps 0-2 decide the polarity; fewer taps contribute nothing (the
    # reference reads uninitialised entries there, src/sdr_func.c:994-997)
    for i in range(min(3, len(pos))):
        t = np.float32(np.float64(c1s[i][0]) * np.float64(c2s[i][0]) +
                       np.float64(np.float32(c1s[i][1] * c2s[i][1])))
        dot = np.float32(dot + t)
    sign = np.float32(-1.0 if dot < 0 else 1.0)
    corr = np.array([complex(np.float32(np.float32(a[0] + sign * b[0]) / np.float32(N)),
                             np.float32(np.float32(a[1] + sign * b[1]) / np.float32(N)))
                     for a, b in zip(c1s, c2s)], np.complex64)
    return corr, np.array([complex(*c1s[0]), complex(*c2s[0])], np.complex64)


@pytest.mark.parametrize("N", [37, 160, 1001])
@pytest.mark.parametrize("cpx", [False, True])
def test_corr_std_bruteforce(orc, N, cpx):
    rng = np.random.default_rng(N + cpx)
    IQ = rng.integers(-112, 113, (N, 2)).astype(np.int8)
    IQ[rng.integers(0, N, 5)] = -128      # the sign_epi8 corner
    code = rng.integers(-1, 2, (10 * N, 2)).astype(np.int8)
    if not cpx:
        code[:, 1] = code[:, 0]
    pos = np.array([0.0, -3.3, 3.3, -0.9 * N, 0.55, -0.95, N + 2.5 - 5.0])
    for coff in (0.0, 7.25, N - 1.6, 2.0):
        got, gC = orc.corr_std(IQ, code, N, coff, pos, cpx_code=cpx)
        ref, rC = brute_corr_std(IQ, code, N, coff, pos, cpx)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
        assert np.array_equal(gC.view(np.uint32), rC.view(np.uint32))


@pytest.mark.parametrize("n", [1, 2])
def test_corr_std_fewer_than_three_taps(orc, n):
    """With n < 3 taps the polarity sum takes the missing taps as zero."""
    N = 1001
    rng = np.random.default_rng(n)
    IQ = rng.integers(-112, 113, (N, 2)).astype(np.int8)
    code = rng.integers(-1, 2, (10 * N, 2)).astype(np.int8)
    code[:, 1] = code[:, 0]
    pos = np.array([0.0, -3.3])[:n]
    flips = 0
    for coff in rng.uniform(0, N, 40):
        got, gC = orc.corr_std(IQ, code, N, coff, pos)
        ref, rC = brute_corr_std(IQ, code, N, coff, pos)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
        assert np.array_equal(gC.view(np.uint32), rC.view(np.uint32))
        flips += int(abs(got[0] - (gC[0] + gC[1]) / np.float32(N)) > 1e-9)
    assert 0 < flips < 40        # both polarities occur


def test_corr_std_self_correlation(orc):
    # IQ = (a c, b c) against its own code at tap 0 -> (a, b) * CSCALE exactly
    N = 2000
    rng = np.random.default_rng(5)
    c = rng.choice([-1, 1], N).astype(np.int8)
    code = np.tile(np.stack([c, c], 1), (10, 1))
    IQ = np.stack([c * 33, c * -50], 1).astype(np.int8)
    # taps with j = 0 make corr1 = 0, so the polarity heuristic keeps sign +1
    corr, _ = orc.corr_std(IQ, code, N, 0.0, np.array([0.0, 0.2, 0.4]))
    re = np.float32(np.float32(np.float32(33 * N) * CSCALE) / np.float32(N))
    im = np.float32(np.float32(np.float32(-50 * N) * CSCALE) / np.float32(N))
    assert corr[0] == np.complex64(complex(re, im))
    assert abs(corr[0] - complex(33 / 11.2, -50 / 11.2)) < 1e-5


# ---------------------------------------------------------------- cpx_mul / FFT
def test_cpx_mul_known_answers(orc):
    # test/src/test_sdr_func.c:71-82
    a = np.array([complex(i + 1, -i) for i in range(4)], np.complex64)
    b = np.full(4, 2 + 1j, np.complex64)
    c = orc.cpx_mul(a, b, 0.5)
    assert abs(c[0] - (1.0 + 0.5j)) < 1e-6
    assert abs(c[1] - (2.5 + 0.0j)) < 1e-6
    rng = np.random.default_rng(2)
    a = (rng.standard_normal(103) + 1j * rng.standard_normal(103)).astype(np.complex64)
    b = (rng.standard_normal(103) + 1j * rng.standard_normal(103)).astype(np.complex64)
    c = orc.cpx_mul(a, b, 0.25)
    ref = a.astype(np.complex128) * b.astype(np.complex128) * 0.25
    assert np.max(np.abs(c - ref)) < 1e-5


def dft(x, sign):
    n = len(x)
    k = np.arange(n)
    W = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
    return W @ x


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 12, 15, 30, 49, 100, 160])
def test_fft_f64_bruteforce(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose(orc.fft_f64(x, 0), dft(x, -1), atol=1e-9 * n)
    assert np.allclose(orc.fft_f64(x, 1), dft(x, +1), atol=1e-9 * n)


@pytest.mark.parametrize("n", [8, 150, 4000, 24000])
def test_fft_f32(orc, n):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    ref = orc.fft_f64(x.astype(np.complex128), 0)
    y = orc.fft_f32(x, 0)
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 2e-6
    z = orc.fft_f32(y, 1) / n
    assert np.linalg.norm(z - x) / np.linalg.norm(x) < 2e-6


def test_ref_pocketfft_loaded_when_built(orc):
    import os
    if os.path.exists(orc.REF_FFT):
        assert orc.REF_LOADED


# ---------------------------------------------------------------- FFT correlator
def test_corr_fft_zero_code(orc):
    # test/src/test_sdr_func.c:235-243
    IQ = np.array([[1, 0], [2, -1], [3, -2], [4, -3]], np.int8)
    out = orc.corr_fft(IQ, np.zeros(4, np.complex64))
    assert np.all(np.abs(out) < 1e-6)


def test_corr_fft_time_domain(orc):
    # IFFT(FFT(x) conj(FFT(c))) / M^2 = (1/M) sum_m x[m + j] c[m]
    M = 120
    rng = np.random.default_rng(3)
    IQ = rng.integers(-100, 100, (M, 2)).astype(np.int8)
    c = np.zeros(M)
    c[:M // 2] = rng.choice([-1, 1], M // 2)
    code_fft = np.conj(orc.fft_f32(c.astype(np.complex64), 0))
    got = orc.corr_fft(IQ, code_fft)
    x = (IQ[:, 0] + 1j * IQ[:, 1]) * float(CSCALE)
    ref = np.array([np.sum(np.roll(x, -j) * c) for j in range(M)]) / M
    assert np.max(np.abs(got - ref)) < 1e-5 * np.max(np.abs(ref))
    # delta code at d: corr[j] = x[j + d] / M
    d = 17
    cd = np.zeros(M, np.complex64)
    cd[d] = 1
    got = orc.corr_fft(IQ, np.conj(orc.fft_f32(cd, 0)))
    assert np.allclose(got, np.roll(x, -d) / M, atol=1e-6)


def test_search_code_zero_code(orc):
    # test/src/test_sdr_func.c:244-248
    buff = pack(np.array([1, 2, 3, 4]), np.array([0, -1, -2, -3]))
    P = np.ones((1, 4), np.float32)
    orc.search_code(np.zeros(4, np.complex64), 1.0, buff, 0, 4.0, 0.0, np.zeros(1, np.float32), P)
    assert np.allclose(P, 1.0, atol=1e-6)


def test_search_code_time_domain(orc):
    M, fs, fi = 200, 2.0e5, 1234.0
    rng = np.random.default_rng(4)
    buff = rng.integers(0, 256, 700).astype(np.uint8)
    c = np.zeros(M)
    c[:M // 2] = rng.choice([-1, 1], M // 2)
    code_fft = np.conj(orc.fft_f32(c.astype(np.complex64), 0))
    fds = np.array([-1000, 0, 1000], np.float32)
    P = orc.search_code(code_fft, 1e-3, buff, 550, fs, fi, fds)   # wraps the ring
    code_fft64 = np.conj(orc.fft_f64(c.astype(np.complex128), 0))
    Pd = orc.search_code_f64(code_fft64, 1e-3, buff, 550, fs, fi, fds)
    for b, fd in enumerate(fds):
        IQ = orc.mix_carr(buff, 550, M, fs, fi + float(fd), 0.0)
        x = (IQ[:, 0] + 1j * IQ[:, 1]) * float(CSCALE)
        C = np.array([np.sum(np.roll(x, -j) * c) for j in range(M)]) / M
        ref = np.abs(C) ** 2
        assert np.max(np.abs(P[b] - ref)) < 1e-4 * ref.max()
        assert np.max(np.abs(Pd[b] - ref)) < 1e-9 * ref.max()


# ---------------------------------------------------------------- peak helpers
def test_peak_helpers_known_answers(orc):
    # test/src/test_sdr_func.c:255-280
    P = np.array([1, 1, 1, 1, 1, 1, 4, 1, 1, 1, 1, 1], np.float32)
    fds = np.array([-500, 0, 500], np.float32)
    cn0, ix = orc.corr_max(P, 4, 4, 3, 0.001, np.array([-1, -1]))
    assert cn0 > 0 and ix.tolist() == [1, 2]
    mean = P.mean()
    assert abs(cn0 - 10 * np.log10((4 - mean) / mean / 0.001)) < 1e-4
    assert abs(orc.fine_dop(P, 4, fds, ix)) < 1e-6
    assert abs(orc.fine_dop(P, 4, fds, np.array([0, 0])) + 500.0) < 1e-6
    cn0, _ = orc.corr_max(np.zeros(4, np.float32), 4, 4, 1, 0.001)
    assert cn0 == 0.0
    bins = orc.dop_bins(0.001, 0.0, 1000.0)
    assert bins.tolist() == [-1000.0, -500.0, 0.0, 500.0, 1000.0]
    assert len(orc.dop_bins(0.004, 0.0, 5000.0)) == 81
    assert len(orc.dop_bins(0.010, 0.0, 6000.0)) == 241


def test_fine_dop_parabola(orc):
    fds = np.arange(-2000, 2001, 500).astype(np.float32)
    x0 = 123.25
    y = (1e4 - 0.001 * (fds.astype(np.float64) - x0) ** 2).astype(np.float32)
    P = np.zeros((len(fds), 3), np.float32)
    P[:, 1] = y
    d = orc.fine_dop(P, 3, fds, np.array([4, 1]))
    assert abs(d - x0) < 0.5


def test_corr_max_first_max_wins(orc):
    P = np.zeros((3, 5), np.float32)
    P[1, 3] = 7
    P[2, 0] = 7
    _, ix = orc.corr_max(P, 5, 5, 3, 0.001)
    assert ix.tolist() == [1, 3]
    _, ix = orc.corr_max(P, 5, 3, 3, 0.001)   # lags >= Nmax ignored
    assert ix.tolist() == [2, 0]


# ---------------------------------------------------------------- codes
def test_res_code_known_answers(orc):
    # test/src/test_sdr_code.c:117-136
    out = orc.res_code(np.array([1, -1, 1, -1], np.int8), 1.0, 0.0, 4.0, 4, 2)
    assert out[:, 0].tolist() == [1, -1, 1, -1, 0, 0]
    assert out[:, 1].tolist() == [1, -1, 1, -1, 0, 0]


L1CA_OCTAL = ["1440", "1620", "1710", "1744", "1133", "1455", "1131", "1454", "1626", "1504",
              "1642", "1750", "1764", "1772", "1775", "1776", "1156", "1467", "1633", "1715",
              "1746", "1763", "1063", "1706", "1743", "1761", "1770", "1774", "1127", "1453",
              "1625", "1712"]


@pytest.mark.parametrize("prn", range(1, 33))
def test_l1ca_first_chips(orc, prn):
    # IS-GPS-200 Table 3-Ia: first 10 chips in octal (chip -1 <-> logic 1)
    c = orc.gen_code_L1CA(prn)
    bits = "".join("1" if v == -1 else "0" for v in c[:10])
    assert oct(int(bits, 2))[2:] == L1CA_OCTAL[prn - 1]
    assert np.sum(c == -1) == 512


def test_product_code_generator_matches_oracle(orc):
    from pocketsdr_amd import codes
    for prn in (1, 7, 32):
        assert np.array_equal(codes.gen_code_L1CA(prn), orc.gen_code_L1CA(prn))
    c = orc.gen_code_L1CA(3)
    assert np.array_equal(codes.res_code(c, 1e-3, 0.0, 12e6, 12000, 12000),
                          orc.res_code(c, 1e-3, 0.0, 12e6, 12000, 12000))
    assert np.array_equal(codes.code_bank(c, 1e-3, 24e6, 24000), orc.code_bank(c, 1e-3, 24e6, 24000))


# ---------------------------------------------------------------- tracking
def test_trk_run_composition(orc):
    fs, N = 4e6, 4000
    rng = np.random.default_rng(6)
    buff = rng.integers(0, 256, N * 5 + 123).astype(np.uint8)
    c = orc.gen_code_L1CA(5)
    bank = orc.code_bank(c, 1e-3, fs, N)
    ch = dict(fs=fs, fi=1000.0, fc_sig=1575.42e6, T=1e-3, N=N, fd0=-2500.0, drift=1.5,
              phi0=0.25, coff0=0.3e-3)
    pos = np.array([0.0, -2.0, 2.0, -120.0, 0.5])
    corr, Cc, state = orc.trk_run(ch, bank, buff, 7, pos)
    # recurrence in closed form (no wrap of coff here)
    t = np.arange(1, 8) * 1e-3
    fd = -2500.0 + 1.5 * t
    phi = 0.25
    coff = 0.3e-3
    for e in range(7):
        coff -= fd[e] / 1575.42e6 * 1e-3                   # fused in the oracle
        phi += (1000.0 + fd[e]) * 1e-3
        phi -= np.floor(phi)
        assert abs(state[e, 0] - phi) < 1e-12 and abs(state[e, 1] - coff) < 1e-15
        IQ = orc.mix_carr(buff, (e * N) % len(buff), N, fs, state[e, 2], state[e, 0])
        r, rC = orc.corr_std(IQ, bank, N, state[e, 1] * fs, pos)
        assert np.array_equal(r, corr[e]) and np.array_equal(rC, Cc[e])


def test_trk_stitch_known_answers(orc):
    """Prompt history (src/sdr_ch.c:736-740): this epoch's correlation before
    the code wrap plus the previous epoch's after it, over N; zero before the
    first epoch, so a -0 first half becomes +0 (-0 + 0 = +0)."""
    Cc = np.array([[1 + 2j, 3 + 4j], [5 + 6j, 7 - 8j], [-0.5 + 1j, 0j]], np.complex64)
    P = orc.trk_stitch(Cc, 4)
    assert P.tolist() == [0.25 + 0.5j, 2 + 2.5j, 1.625 - 1.75j]
    z = np.array([[complex(-0.0, -0.0), 0j]], np.complex64)
    Pz = orc.trk_stitch(z, 8)
    assert not np.signbit(Pz.real[0]) and not np.signbit(Pz.imag[0])


def test_trk_stitch_full_period(orc):
    """For a channel on the true code phase the stitched prompt is one whole
    code period's correlation: it equals the sum over that period's samples
    (previous epoch's tail, this epoch's head) of mixed sample x code, / N."""
    from pocketsdr_amd import workloads
    w = workloads.trk_l1ca(seed=4, fs=4e6, n_ch=2, n_epoch=6, n_vis=2, nposx=0)
    ch = w["chans"][0]
    N, fs = ch["N"], ch["fs"]
    _, Cc, st = orc.trk_run(ch, ch["code"], w["buff"], w["n_epoch"], [0.0])
    P = orc.trk_stitch(Cc, N)
    bank = ch["code"].reshape(10, N, 2)[..., 0].astype(np.float64)
    for e in range(1, w["n_epoch"]):
        parts = []
        for ee, half in ((e - 1, 1), (e, 0)):
            IQ = orc.mix_carr(w["buff"], (ee * N) % len(w["buff"]), N, fs, st[ee, 2], st[ee, 0])
            x = (IQ[:, 0] + 1j * IQ[:, 1]).astype(np.complex128) / 11.2
            p = st[ee, 1] * fs
            j = int(np.floor(p))
            k = int((p - j) * 10)
            code = bank[k]
            if half == 1:
                parts.append(np.dot(x[j:], code[:N - j]))
            else:
                parts.append(np.dot(x[:j], code[N - j:]))
        want = sum(parts) / N
        assert abs(P[e] - want) <= 1e-5 * abs(want) + 1e-7, e


@pytest.mark.parametrize("N", [4093, 50021])
def test_fft_f64_large_prime(orc, N):
    """The FP64 transform of a prime length against numpy's FFT (a textbook
    library routine): the generic radix's coefficient index q m mod r must
    not overflow for r > 46340."""
    rng = np.random.default_rng(N)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    for d in (0, 1):
        ref = np.fft.fft(x) if d == 0 else np.fft.ifft(x) * N
        got = orc.fft_f64(x, d)
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-12

