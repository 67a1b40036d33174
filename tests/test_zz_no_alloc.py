"""sdr_search_code_noalloc: sdr_search_code on the library's fixed device
scratch array (no device allocation).  Checked through the C ABI against
sdr_search_code itself, bit for bit, on consecutive calls of decreasing and
increasing size (the array keeps the previous call's contents), including a
call whose bins are split into several passes, and the rejection of lengths
the array cannot hold."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def case(seed, fs, N, nbin, buffN, ix):
    rng = np.random.default_rng(seed)
    buff = rng.integers(0, 256, buffN).astype(np.uint8)
    code = rng.choice([-1.0, 1.0], N // 2)
    c = np.zeros(N, np.complex64)
    c[:N // 2] = code
    code_fft = np.conj(np.fft.fft(c)).astype(np.complex64)
    fds = (np.arange(nbin) - nbin // 2).astype(np.float32) * np.float32(500.0)
    P0 = rng.uniform(0.0, 1e-3, (nbin, N)).astype(np.float32)    # accumulated into
    return buff, code_fft, fds, P0


@pytest.mark.parametrize("sizes", [
    # (seed, fs, FFT length N, bins, ring length, start): large, smaller, chunked
    [(1, 12e6, 24000, 21, 36000, 30000), (2, 4e6, 8000, 5, 8000, 0), (3, 12e6, 96000, 81, 100000, 50000)],
])
def test_noalloc_matches_search_code(sdr, sizes):
    for seed, fs, N, nbin, buffN, ix in sizes:
        buff, code_fft, fds, P0 = case(seed, fs, N, nbin, buffN, ix)
        ref = sdr.search_code(code_fft, N / 2 / fs, buff, ix, fs, 0.0, fds, P0.copy())
        ok, got = sdr.search_code_noalloc(code_fft, N / 2 / fs, buff, ix, fs, 0.0, fds, P0.copy())
        assert ok == 1, (N, nbin)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (N, nbin)


def test_noalloc_rejects_oversize(sdr):
    # 2,000,000 = 2^7 5^6: one bin needs about 74 MB of the 64 MiB array
    N = 2_000_000
    buff = np.zeros(N, np.uint8)
    code_fft = np.zeros(N, np.complex64)
    fds = np.zeros(1, np.float32)
    P = np.full((1, N), 7.0, np.float32)
    ok, out = sdr.search_code_noalloc(code_fft, 1e-3, buff, 0, 2e9, 0.0, fds, P)
    assert ok == 0 and np.all(out == 7.0)
    # a length with a prime factor above 5
    ok, _ = sdr.search_code_noalloc(np.zeros(77, np.complex64), 1e-3, np.zeros(77, np.uint8), 0, 77e3,
                                    0.0, fds, np.zeros((1, 77), np.float32))
    assert ok == 0
    # a valid call after the rejections still matches
    buff, code_fft, fds, P0 = case(4, 4e6, 8000, 3, 8000, 100)
    ref = sdr.search_code(code_fft, 1e-3, buff, 100, 4e6, 0.0, fds, P0.copy())
    ok, got = sdr.search_code_noalloc(code_fft, 1e-3, buff, 100, 4e6, 0.0, fds, P0.copy())
    assert ok == 1 and np.array_equal(got.view(np.uint32), ref.view(np.uint32))
