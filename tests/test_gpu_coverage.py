"""GPU parity of the Python-facing FFT correlator, of the GPU code spectra
that feed every search workload, and the device-memory balance of the
search plans.

* sdr_corr_fft_cpx (src/sdr_func.c:1105-1113 with mix_carr_cpx, :839-848):
  float complex samples cast to int8, packed to cpx8, mixed, FFT-correlated;
  checked against the same composition in the oracle.
* sdr_gen_code_fft (src/sdr_code.c:2589-2608) as the workloads compute it on
  the library's FFT, against the oracle's restatement (reference PocketFFT
  when built) and the FP64 oracle, for every signal of the search configs.
* FFT lengths with prime factors above 5 (the reference's PocketFFT computes
  every length, src/sdr_func.c:236-264): sdr_cpx_fft against the FP64
  oracle, sdr_corr_fft and sdr_search_code against the oracle, e.g. the
  16368-point correlator of a 1 ms code at fs = 8.184 MHz (2^4 3 11 31).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30)


def cpx8_of(buff):
    """SDR_CPX8((int8_t)re, (int8_t)im) of src/pocket_sdr.h:105: C's
    float-to-int conversion truncates toward zero, then the low nibbles."""
    re = np.trunc(buff.real).astype(np.int64)
    im = np.trunc(buff.imag).astype(np.int64)
    return ((re & 0xF) | ((im & 0xF) << 4)).astype(np.uint8)


@pytest.mark.parametrize("case", [
    # len_buff, ix, N (FFT length), fs, fc, phi
    (8000, 0, 8000, 4e6, 1500.0, 0.25),
    (9000, 7000, 8000, 4e6, -2250.0, 0.9),      # ring wrap inside the block
    (30000, 5, 24000, 12e6, 3125.5, 0.0),
])
def test_corr_fft_cpx(sdr, orc, case):
    len_buff, ix, N, fs, fc, phi = case
    rng = np.random.default_rng(N + ix)
    # fractional values of both signs (truncation), the -8 nibble included
    buff = (rng.uniform(-8.99, 7.99, len_buff) + 1j * rng.uniform(-8.99, 7.99, len_buff)).astype(np.complex64)
    code = rng.choice([-1, 1], 1023).astype(np.int8)
    code_fft = orc.gen_code_fft(code, N / 2 / fs, 0.0, fs, N // 2, N // 2)
    got = sdr.corr_fft_cpx(buff, ix, N, fs, fc, phi, code_fft)
    idx = (ix + np.arange(N)) % len_buff
    IQ = orc.mix_carr(cpx8_of(buff[idx]), 0, N, fs, fc, phi)
    ref = orc.corr_fft(IQ, code_fft)
    # FP32 FFT correlator against the oracle on the reference PocketFFT
    assert rel(got, ref) < 1e-5
    assert np.argmax(np.abs(got)) == np.argmax(np.abs(ref))


# signal, sampling rate, code period: the search configurations' groups
CODE_CASES = [("L1CA", 1, 12e6, 1e-3), ("E1B", 3, 12e6, 4e-3), ("L5Q", 7, 24e6, 1e-3),
              ("L6D", 194, 24e6, 4e-3), ("L1CP", 11, 48e6, 10e-3)]


@pytest.mark.parametrize("case", CODE_CASES, ids=[c[0] for c in CODE_CASES])
def test_gen_code_fft_gpu(orc, sdr, case):
    from pocketsdr_amd import codes, workloads
    name, prn, fs, T = case
    code = workloads.sig_code(name, prn)
    N = int(round(fs * T))
    got = codes.gen_code_fft(code, T, 0.0, fs, N, N)
    assert got.shape == (2 * N,) and got.dtype == np.complex64
    ref32 = orc.gen_code_fft(code, T, 0.0, fs, N, N)
    ref64 = orc.gen_code_fft_f64(code, T, 0.0, fs, N, N)
    M = 2 * N
    # FP32 FFT: normwise 5e-7 log2 M + 1e-7 against FP64, and against the
    # FP32 oracle twice that (both carry an FP32 rounding error)
    assert rel(got, ref64) < 5e-7 * np.log2(M) + 1e-7
    assert rel(got, ref32) < 1e-6 * np.log2(M) + 2e-7
    # DC bin: the sum of the resampled code (exact in FP32 for these sizes)
    assert abs(got[0] - ref64[0]) <= 1e-3 * max(1.0, abs(ref64[0]))


def test_search_plan_free_releases_memory(sdr):
    """Every device buffer a plan allocates is freed with it: the bytes held by
    live plans (the library's own accounting of each allocation) return to
    zero after each of ten create / run / free cycles of a precision-form
    plan, and are the same for each plan while it lives."""
    from pocketsdr_amd import workloads
    w = workloads.acq_multi(seed=3, prn_limit=4)
    base = sdr.search_plan_bytes()
    held = set()
    for _ in range(10):
        p = sdr.SearchPlan(w["buffs"], w["sigs"], sdr.SEARCH_FP16)
        p.run()
        held.add(sdr.search_plan_bytes() - base)
        p.close()
        assert sdr.search_plan_bytes() == base
    assert len(held) == 1 and held.pop() > 0


# lengths: primes (50021 has a transform length of 101250 = 270 x 375, a
# split without specialised kernels), products with 7 / 11 / 31, the L6 code
# length 2 x 10230
ANY_N = [7, 77, 1023, 4093, 16368, 20460, 50021]


@pytest.mark.parametrize("N", ANY_N)
def test_cpx_fft_any_length(sdr, orc, N):
    rng = np.random.default_rng(N)
    x = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    for d in (0, 1):
        ref = orc.fft_f64(x.astype(np.complex128), d)
        got = sdr.cpx_fft(x, d)
        err = rel(got, ref)
        # Bluestein on an FP32 FFT of length L <= 4N: three FFTs and three
        # chirp products, normwise 1.5e-6 log2 N + 1e-6
        assert err < 1.5e-6 * np.log2(N) + 1e-6, (N, d, err)


@pytest.mark.parametrize("N", [16368, 20460])
def test_corr_fft_any_length(sdr, orc, N):
    rng = np.random.default_rng(N + 1)
    IQ = rng.integers(-112, 113, (N, 2)).astype(np.int8)
    c = np.zeros(N, np.complex64)
    c[:N // 2] = rng.choice([-1, 1], N // 2)
    code_fft = np.conj(orc.fft_f32(c, 0))
    got = sdr.corr_fft(IQ, code_fft)
    ref = orc.corr_fft(IQ, code_fft)
    assert rel(got, ref) < 1e-5
    assert np.argmax(np.abs(got)) == np.argmax(np.abs(ref))


def test_search_code_any_length(sdr, orc):
    """A 1 ms L1 C/A search at fs = 8.184 MHz: M = 16368 = 2^4 3 11 31."""
    from pocketsdr_amd import workloads
    fs, T = 8.184e6, 1e-3
    N = int(round(fs * T))
    M = 2 * N
    rng = np.random.default_rng(5)
    buff = rng.integers(0, 256, 3 * N).astype(np.uint8)
    code = workloads.l1ca_code(3)
    code_fft = orc.gen_code_fft(code, T, 0.0, fs, N, N)
    fds = workloads.dop_bins(T, 0.0, 1500.0)
    P0 = rng.uniform(0, 1e-4, (len(fds), M)).astype(np.float32)
    got = sdr.search_code(code_fft, T, buff, 2 * N + 17, fs, 0.0, fds, P0.copy())   # ring wrap
    ref = orc.search_code(code_fft, T, buff, 2 * N + 17, fs, 0.0, fds, P0.copy())
    assert rel(got, ref) < 1e-5
    assert np.argmax(got[:, :N]) == np.argmax(ref[:, :N])
