"""Seeded synthetic IF data in the hot path's cpx8 format.

Signal model: complex baseband at IF ``fi``,
    x(t) = sum_p A_p c_p(t - tau_p) exp(j 2 pi (fi + fd_p(t)) t + j phi_p) + w(t)
with AWGN of sigma 1 per I/Q component (noise density N0 = 2 / fs), so the
carrier amplitude for a C/N0 (dB-Hz) is A = sqrt(2 10^(CN0/10) / fs).  The
code rate follows the Doppler (code Doppler = fd / f_carrier).  Samples are
quantised to 2 bits {-3, -1, 1, 3} (threshold 1 sigma) or 3 bits
{-7, ..., 7} odd levels (step 0.5 sigma) and packed as cpx8 (I low nibble,
Q high nibble, src/pocket_sdr.h:105-110).
"""
import numpy as np


def quantize(x, bits):
    if bits == 2:
        return np.where(np.abs(x) < 1.0, 1, 3).astype(np.int8) * np.where(x < 0, -1, 1).astype(np.int8)
    if bits == 3:
        q = 2 * np.floor(x / 0.5) + 1
        return np.clip(q, -7, 7).astype(np.int8)
    raise ValueError("bits must be 2 or 3")


def pack_cpx8(I, Q):
    return ((np.asarray(I).astype(np.int16) & 0xF) | ((np.asarray(Q).astype(np.int16) & 0xF) << 4)).astype(np.uint8)


def synth_if(fs, n_samp, sats, seed, bits=2, fi=0.0, chunk=1 << 22):
    """sats: list of dicts with code (int8 chips), T (code period s), tau (s),
    fd (Hz), drift (Hz/s, optional), cn0 (dB-Hz), fc (carrier Hz), phi (cyc),
    shift (optional: code delay in entries per code period, cycled), sec
    (optional: secondary code, one chip per code period, cycled).
    Returns the uint8 cpx8 buffer."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = np.empty(n_samp, np.uint8)
    for s0 in range(0, n_samp, chunk):
        n = min(chunk, n_samp - s0)
        t = (np.arange(n, dtype=np.float64) + s0) / fs
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        for sat in sats:
            code = np.asarray(sat["code"], np.float64)
            L = len(code)
            drift = sat.get("drift", 0.0)
            fd = sat["fd"] + drift * t
            # code phase (chips): rate L/T scaled by (1 + fd / fc)
            ph_code = (t - sat["tau"]) * (L / sat["T"]) + \
                (L / sat["T"]) * (sat["fd"] * t + 0.5 * drift * t * t) / sat["fc"]
            ic = np.floor(ph_code).astype(np.int64)
            if "shift" in sat:
                # code-shift keying: period k of the code is delayed by
                # shift[k mod len] entries
                sh = np.asarray(sat["shift"], np.int64)
                ic = ic - sh[np.floor_divide(ic, L) % len(sh)]
            chips = code[ic % L]
            if "sec" in sat:
                sec = np.asarray(sat["sec"], np.float64)
                chips = chips * sec[np.floor_divide(ic, L) % len(sec)]
            ph = 2 * np.pi * ((fi + sat["fd"]) * t + 0.5 * drift * t * t + sat.get("phi", 0.0))
            A = np.sqrt(2.0 * 10 ** (sat["cn0"] / 10.0) / fs)
            x += A * chips * np.exp(1j * ph)
            del fd
        out[s0:s0 + n] = pack_cpx8(quantize(x.real, bits), quantize(x.imag, bits))
    return out


def random_sats(rng, prns, codes, T, fc, cn0_range, dop_range, drift_range=(0.0, 0.0)):
    """Seeded satellite set: uniform code phase in [0, T), Doppler, C/N0."""
    sats = []
    for p in prns:
        sats.append(dict(prn=p, code=codes[p], T=T, fc=fc,
                         tau=float(rng.uniform(0, T)),
                         fd=float(rng.uniform(*dop_range)),
                         drift=float(rng.uniform(*drift_range)),
                         cn0=float(rng.uniform(*cn0_range)),
                         phi=float(rng.uniform(0, 1))))
    return sats
