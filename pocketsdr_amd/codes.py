"""Spreading-code replicas consumed by the hot path.

* GPS L1 C/A Gold codes (IS-GPS-200 G1/G2 shift registers and G2 delays; the
  reference generates them in src/sdr_code.c:776-812).
* Resampling of a code to the sample rate, ``sdr_res_code``
  (src/sdr_code.c:2556-2569), and the tracking code bank of SDR_N_CODES = 10
  fractional offsets built in src/sdr_ch.c:266-272.
* Code spectra for the FFT correlator, ``sdr_gen_code_fft``
  (src/sdr_code.c:2589-2608), computed with the library's GPU FFT.
"""
import numpy as np

from . import sdr_func

N_CODES = 10

# IS-GPS-200 G2 delays (chips) for PRN 1-32
_L1CA_G2_DELAY = (5, 6, 7, 8, 17, 18, 139, 140, 141, 251, 252, 254, 255, 256, 257, 258,
                  469, 470, 471, 472, 473, 474, 509, 512, 513, 514, 515, 516, 859, 860, 861, 862)


def _lfsr10(reg, tap):
    out = np.empty(1023, np.int8)
    for i in range(1023):
        out[i] = -1 if reg & 1 else 1
        fb = bin(reg & tap).count("1") & 1
        reg = (fb << 9) | (reg >> 1)
    return out


_G1 = None
_G2 = None


def gen_code_L1CA(prn):
    """GPS L1 C/A code of PRN 1-32 as int8 +/-1 (1023 chips)."""
    global _G1, _G2
    if not 1 <= prn <= 32:
        raise ValueError("L1CA PRN must be 1-32")
    if _G1 is None:
        _G1 = _lfsr10(0x3FF, 0x081)
        _G2 = _lfsr10(0x3FF, 0x197)
    d = _L1CA_G2_DELAY[prn - 1]
    idx = (np.arange(1023) + 1023 - d) % 1023
    return (_G1 * _G2[idx]).astype(np.int8)


def res_index(len_code, T, coff, fs, N):
    """chip index of each sample: (int)((coff fs + i) dx) % len, dx = len / T / fs"""
    dx = len_code / T / fs
    i = np.arange(N, dtype=np.float64)
    return ((coff * fs + i) * dx).astype(np.int64) % len_code


def res_code(code, T, coff, fs, N, Nz=0):
    """sdr_res_code for a real code: cpx16 array (N + Nz, 2) with I = Q."""
    out = np.zeros((N + Nz, 2), np.int8)
    c = np.asarray(code, np.int8)[res_index(len(code), T, coff, fs, N)]
    out[:N, 0] = c
    out[:N, 1] = c
    return out


def code_bank(code, T, fs, N):
    """Tracking bank: SDR_N_CODES resampled codes at offsets -k / fs / 10
    (src/sdr_ch.c:266-272), shape (10 N, 2) int8."""
    return np.concatenate([res_code(code, T, -k / fs / N_CODES, fs, N) for k in range(N_CODES)])


def gen_code_fft(code, T, coff, fs, N, Nz):
    """sdr_gen_code_fft for a real code: conj(FFT(resampled code, zero padded))."""
    x = np.zeros(N + Nz, np.complex64)
    x[:N] = np.asarray(code, np.float32)[res_index(len(code), T, coff, fs, N)]
    return np.conj(sdr_func.cpx_fft(x, 0)).astype(np.complex64)
