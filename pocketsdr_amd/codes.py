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


def _nh(bits):
    return np.array([1 if b == "0" else -1 for b in bits], np.int8)


def sec_code_NH10():
    """10-bit Neuman-Hoffman code 0000110101 (IS-GPS-705 L5-I), bit 0 -> +1,
    1 -> -1 (src/sdr_code.c:693-695)."""
    return _nh("0000110101")


def sec_code_NH20():
    """20-bit Neuman-Hoffman code 00000100110101001110 (IS-GPS-705 L5-Q), bit
    0 -> +1, 1 -> -1 (src/sdr_code.c:697-699, 1114-1118)."""
    return _nh("00000100110101001110")


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


# ---------------------------------------------------------------------------
# other signals of the batched search (ICD parameter tables in
# data/code_params.json, generation steps as src/sdr_code.c restates them)
# ---------------------------------------------------------------------------
import json as _json
import os as _os

_PARAMS = None


def _params():
    global _PARAMS
    if _PARAMS is None:
        with open(_os.path.join(_os.path.dirname(__file__), "data", "code_params.json")) as f:
            _PARAMS = _json.load(f)
    return _PARAMS


def lfsr(N, R, tap, n):
    """Fibonacci LFSR of src/sdr_code.c:772-781: chip = {+1, -1}[R & 1],
    R = (parity(R & tap) << (n - 1)) | (R >> 1)."""
    out = np.empty(N, np.int8)
    for i in range(N):
        out[i] = -1 if R & 1 else 1
        R = ((bin(R & tap).count("1") & 1) << (n - 1)) | (R >> 1)
    return out


def rev_reg(R, N):
    RR = 0
    for i in range(N):
        RR = (RR << 1) | ((R >> i) & 1)
    return RR


_L5 = {}


def gen_code_L5Q(prn):
    """GPS L5 Q5 (IS-GPS-705): XA (8190-chip short cycle, repeated) times XB
    advanced per PRN; 10230 chips (src/sdr_code.c:1007-1061)."""
    if not 1 <= prn <= 32:
        raise ValueError("L5Q PRN must be 1-32")
    if not _L5:
        xa = lfsr(8190, 0x1FFF, 0x001B, 13)
        _L5["XA"] = np.concatenate([xa, xa[:10230 - 8190]])
        _L5["XB"] = lfsr(10230, 0x1FFF, 0x16E3, 13)
    adv = _params()["L5Q_XB_adv"][prn - 1]
    return (_L5["XA"] * _L5["XB"][(np.arange(10230) + adv) % 10230]).astype(np.int8)


def gen_code_L6D(prn):
    """QZSS L6D (IS-QZSS-L6): 10-stage x 20-stage Kasami-like product, TDM
    with zero chips, 20460 entries (src/sdr_code.c:1166-1193)."""
    if not 193 <= prn <= 202:
        raise ValueError("L6D PRN must be 193-202")
    c1 = lfsr(10230, 0x3FF, 0x0F3, 10)
    c2 = lfsr(10230, rev_reg(_params()["L6D_R_init"][prn - 193], 20), 0x00053, 20)
    out = np.zeros(20460, np.int8)
    out[0::2] = c1 * c2
    return out


def gen_code_E1B(prn):
    """Galileo E1-B memory code (hex table) with a BOC(1,1) sub-carrier,
    8184 entries (src/sdr_code.c:1386-1400, 742-754)."""
    if not 1 <= prn <= 36:
        raise ValueError("E1B PRN must be 1-36")
    h = _params()["E1B_hex"][prn - 1]
    bits = np.array([(int(c, 16) >> (3 - j)) & 1 for c in h for j in range(4)], np.int8)[:4092]
    chips = np.where(bits == 1, -1, 1).astype(np.int8)
    return np.stack([chips, -chips], 1).reshape(-1).astype(np.int8)


_LEG = None


def gen_code_L1CP(prn):
    """GPS L1C pilot (IS-GPS-800): Weil code from the 10223 Legendre sequence
    with a 7-chip insertion, TMBOC(6,1,4/33), 122760 entries
    (src/sdr_code.c:838-913)."""
    global _LEG
    if not 1 <= prn <= 32:
        raise ValueError("L1CP PRN must be 1-32")
    if _LEG is None:
        L = np.ones(10223, np.int8)
        i = np.arange(1, 10223, dtype=np.int64)
        L[(i * i) % 10223] = -1
        _LEG = L
    w = _params()["L1CP_weil_idx"][prn - 1]
    p = _params()["L1CP_ins_idx"][prn - 1]
    t = np.arange(10230)
    src = np.where(t < p - 1, t, t - 7)
    code = (_LEG[src % 10223] * _LEG[(src + w) % 10223]).astype(np.int8)
    code[p - 1:p + 6] = [1, -1, -1, 1, -1, 1, 1]
    i = np.arange(10230) % 33
    b6 = (i == 0) | (i == 4) | (i == 6) | (i == 29)
    j = np.arange(12)
    sub_boc61 = np.where(j % 2 == 1, -1, 1)
    sub_boc11 = np.where(j >= 6, -1, 1)
    sub = np.where(b6[:, None], sub_boc61[None, :], sub_boc11[None, :])
    return (code[:, None] * sub).reshape(-1).astype(np.int8)
