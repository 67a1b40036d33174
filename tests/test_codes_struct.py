"""Structural pins of the GPS L5-Q code generator (pocketsdr_amd/codes.py,
after src/sdr_code.c:1005-1061), which has no capture among the reference's
samples to be checked against: the XA and XB sequences satisfy the
recurrences of the IS-GPS-705 polynomials (XA: 1 + x^9 + x^10 + x^12 + x^13,
short-cycled to 8190 chips; XB: 1 + x + x^3 + x^4 + x^6 + x^7 + x^8 + x^12 +
x^13, period 8191), both start from the all-ones register, the PRN code is
XA(i) XB((i + advance) mod 10230) as the reference forms it, and the codes
have the correlation properties of such a family (the L1 C/A, E1-B, L1C-P
and L6D generators are pinned by the real captures in test_codes_real.py).
"""
import numpy as np
import pytest


def bits(chips):
    """chip -1 <-> register bit 1"""
    return (np.asarray(chips) < 0).astype(np.int64)


def m_sequence(exps, n, length):
    """The sequence of the recurrence whose reversed characteristic
    polynomial has the given exponents (x^n and the constant included): a
    polynomial 1 + sum x^e + x^n gives s[k + n] = XOR of s[k + n - e] over
    its middle exponents and s[k]; all-ones start."""
    s = [1] * n
    mids = [e for e in exps if 0 < e < n]
    while len(s) < length:
        k = len(s) - n
        v = s[k]
        for e in mids:
            v ^= s[k + n - e]
        s.append(v)
    return np.array(s[:length], np.int64)


@pytest.fixture(scope="module")
def l5():
    from pocketsdr_amd import codes
    codes.gen_code_L5Q(1)
    return codes._L5


def test_xa_polynomial_and_short_cycle(l5):
    xa = bits(l5["XA"])
    ref = m_sequence([0, 9, 10, 12, 13], 13, 8190)
    assert np.array_equal(xa[:8190], ref)
    # short cycle: restarts from all ones after 8190 chips
    assert np.array_equal(xa[8190:], xa[:10230 - 8190])
    assert xa[:13].tolist() == [1] * 13


def test_xb_polynomial_and_period(l5):
    xb = bits(l5["XB"])
    ref = m_sequence([0, 1, 3, 4, 6, 7, 8, 12, 13], 13, 10230)
    assert np.array_equal(xb, ref)
    # a maximal-length sequence: period 8191, no shorter period
    assert np.array_equal(xb[8191:], xb[:10230 - 8191])
    assert all(not np.array_equal(xb[p:p + 500], xb[:500]) for p in range(1, 8191) if 8191 % p == 0)
    # balance of an m-sequence: 4096 ones in a period
    assert int(xb[:8191].sum()) == 4096


def test_code_composition(l5):
    from pocketsdr_amd import codes
    adv = codes._params()["L5Q_XB_adv"]
    for prn in (1, 7, 19, 32):
        c = codes.gen_code_L5Q(prn)
        i = np.arange(10230)
        want = l5["XA"] * l5["XB"][(i + adv[prn - 1]) % 10230]
        assert c.dtype == np.int8 and len(c) == 10230 and np.array_equal(c, want)


def test_correlation_properties():
    from pocketsdr_amd import codes
    C = np.stack([codes.gen_code_L5Q(p).astype(np.float64) for p in range(1, 33)])
    F = np.fft.fft(C, axis=1)
    auto = np.fft.ifft(F * np.conj(F), axis=1).real
    assert np.allclose(auto[:, 0], 10230)
    side = np.abs(auto[:, 1:]).max()
    assert side < 0.1 * 10230, side
    # distinct PRNs: every shift of every pair far below the peak
    worst = 0.0
    for p in range(0, 32, 3):
        for q in range(p + 1, 32, 5):
            x = np.fft.ifft(F[p] * np.conj(F[q])).real
            worst = max(worst, np.abs(x).max())
    assert worst < 0.1 * 10230, worst
    # no two PRNs share a code or a shift of one
    assert len({c.tobytes() for c in C.astype(np.int8)}) == 32
