"""Known answers for single updates of the oracle's closed tracking loop
(src/sdr_ch.c:488-624: FLL, PLL, DLL, C/N0 with the loss decision), so that
a wrong gain, band, filter constant, window or sign in the restatement fails
here and not only the pull-in physics of tests/test_trk_loop.py.  The GPU
loop is compared with this restatement epoch by epoch.  The secondary-code
sync (sync_sec_code) and bump-jump have their own cases at the end: planted
prompt histories of the NH20 code at known amplitude and phase, and VE / VL
powers on either side of BUMP_K.

Each case plants correlator values whose discriminator output is known in
closed form (a prompt rotated by a known angle, early and late amplitudes in
a known ratio, prompt and noise powers in a known ratio) and checks the
state change against the reference's loop constants:
B_FLL_W = 5 Hz for lock T < T_FPULLIN_W = 0.5 s, B_FLL_N = 2 Hz after, the
FLL gain B / 0.25; PLL W = B_PLL / 0.7845 with B_PLL = 5 Hz and the
coefficients 2.4 W, 1.1 W^2, W^3; DLL W = B_DLL / 0.53 with B_DLL = 0.25
Hz, T_DLL = 20 ms, the code error (E - L) / (E + L) / 2 chips, 1.414 W;
T_CN0 = 0.5 s, FILT_CN0 = 0.5, the loss after LOST_TH = 4 bad windows at
C/N0 < 30 dB-Hz or, with a Costas loop, PLI < 0.25.
"""
import math

import numpy as np
import pytest

T = 1e-3
TWO_PI = 2.0 * math.pi


def rot(a, amp=1.0):
    """a prompt of phase a as the loop receives it (FP32 correlators)"""
    return [float(np.float32(amp * math.cos(a))), float(np.float32(amp * math.sin(a)))]


def phase(c):
    """the planted phase of a rotated prompt, from its FP32 components"""
    return math.atan2(c[1], c[0])


@pytest.mark.parametrize("lock,band", [(2, 5.0), (499, 5.0), (500, 2.0), (900, 2.0)])
@pytest.mark.parametrize("dphi", [0.3, -0.05, 1.2])
def test_fll_phase_step(orc, lock, band, dphi):
    # the prompt advanced by dphi since the previous epoch: the frequency
    # error is dphi / (2 pi T) and the FLL moves fd by B / 0.25 dphi / (2 pi)
    c0, c1 = rot(0.4, 3.0), rot(0.4 + dphi, 3.0)
    st = orc.tl_state(T=T, lock=lock, fd=100.0, C0=c0)
    orc.tl_step("fll", st, c1)
    d = phase(c1) - phase(c0)
    assert st.fd == pytest.approx(100.0 + band / 0.25 * d / TWO_PI, rel=0, abs=1e-10)
    assert list(st.C0) == pytest.approx(rot(0.4 + dphi, 3.0), rel=1e-6)


def test_fll_costas_and_guards(orc):
    # a rotation past 90 degrees: atan2 keeps it, the Costas discriminator
    # (atan of the ratio) folds it by pi
    c = rot(2.5)
    for costas, expect in ((0, phase(c)), (1, phase(c) - math.pi)):
        st = orc.tl_state(T=T, costas=costas, lock=10, C0=rot(0.0))
        orc.tl_step("fll", st, c)
        assert st.fd == pytest.approx(5.0 / 0.25 * expect / TWO_PI, abs=1e-10)
    # first locked epoch: no update, the prompt is kept
    st = orc.tl_state(T=T, lock=1, fd=7.0, C0=rot(0.0))
    orc.tl_step("fll", st, rot(1.0))
    assert st.fd == 7.0 and list(st.C0) == pytest.approx(rot(1.0), rel=1e-6)
    # dot = 0 (a quarter turn): no update
    st = orc.tl_state(T=T, lock=5, fd=7.0, C0=[1.0, 0.0])
    orc.tl_step("fll", st, [0.0, 1.0])
    assert st.fd == 7.0


@pytest.mark.parametrize("phi", [0.2, -0.7])
def test_pll_two_steps(orc, phi):
    W = 5.0 / 0.7845
    c = rot(phi, 50.0)
    e = phase(c) / TWO_PI
    st = orc.tl_state(T=T, fd=-300.0)
    orc.tl_step("pll", st, c)
    acc1 = W ** 3 * e * T
    fd1 = -300.0 + 2.4 * W * e + 1.1 * W * W * e * T + acc1 * T
    assert st.phas_acc == pytest.approx(acc1, rel=1e-12)
    assert st.fd == pytest.approx(fd1, rel=0, abs=1e-9)
    assert st.err_phas == pytest.approx(e, rel=1e-12)
    # the same phase error again: no proportional change, the integrators grow
    orc.tl_step("pll", st, c)
    acc2 = 2 * acc1
    assert st.phas_acc == pytest.approx(acc2, rel=1e-12)
    assert st.fd == pytest.approx(fd1 + 1.1 * W * W * e * T + acc2 * T, rel=0, abs=1e-9)


def test_pll_costas_and_guard(orc):
    W = 5.0 / 0.7845
    c = rot(2.0)
    for costas, phi_eff in ((0, phase(c)), (1, phase(c) - math.pi)):
        st = orc.tl_state(T=T, costas=costas)
        orc.tl_step("pll", st, c)
        e = phi_eff / TWO_PI
        assert st.fd == pytest.approx(2.4 * W * e + 1.1 * W * W * e * T + W ** 3 * e * T * T, abs=1e-9)
    st = orc.tl_state(T=T, fd=5.0, err_phas=0.1)
    orc.tl_step("pll", st, [0.0, 1.0])          # IP = 0: no update
    assert st.fd == 5.0 and st.err_phas == 0.1


@pytest.mark.parametrize("a,b,code_int0", [(3.0, 1.0, 0.0), (1.0, 2.0, 2.5e-7), (2.0, 2.0, -1e-7)])
def test_dll_window(orc, a, b, code_int0):
    # early and late amplitudes a and b for a whole T_DLL window of 20 epochs:
    # E = sqrt(20) a, L = sqrt(20) b, code error (a - b) / (a + b) / 2 chips
    W = 0.25 / 0.53
    st = orc.tl_state(T=T, len_code=1023, coff=4e-4, code_int=code_int0)
    corr = [0.0, 0.0, a, 0.0, b, 0.0]
    for k in range(1, 20):
        st.lock = k
        orc.tl_step("dll", st, corr)
        assert st.coff == 4e-4 and st.code_int == code_int0      # inside the window
    assert st.sumE == pytest.approx(19 * a * a) and st.sumL == pytest.approx(19 * b * b)
    st.lock = 20
    orc.tl_step("dll", st, corr)
    err = (a - b) / (a + b) * 0.5 * T / 1023
    dt = 20 * T
    ci = code_int0 + W * W * err * dt
    assert st.code_int == pytest.approx(ci, rel=1e-12, abs=1e-20)
    assert st.coff == pytest.approx(4e-4 - (1.414 * W * err + ci) * dt, rel=1e-13)
    assert st.sumE == 0.0 and st.sumL == 0.0


def test_dll_window_length_follows_T(orc):
    # T = 4 ms: a window of 5 epochs; T = 50 ms: every epoch
    for Tk, nd in ((4e-3, 5), (50e-3, 1)):
        st = orc.tl_state(T=Tk, len_code=4092)
        for k in range(1, nd + 1):
            st.lock = k
            orc.tl_step("dll", st, [0, 0, 2.0, 0, 1.0, 0])
        err = (1.0 / 3.0) * 0.5 * Tk / 4092
        W = 0.25 / 0.53
        assert st.code_int == pytest.approx(W * W * err * nd * Tk, rel=1e-12)


def test_cn0_window_and_filter(orc):
    # prompt (p, q) and noise tap (n, 0) over one T_CN0 window of 500 epochs
    p, q, n = 0.5, 0.25, 0.125
    st = orc.tl_state(T=T, cn0=40.0)
    for k in range(1, 500):
        st.lock = k
        assert orc.tl_step("cn0", st, [0] * 6 + [n, 0.0], p, q) == 0
    st.lock = 500
    assert orc.tl_step("cn0", st, [0] * 6 + [n, 0.0], p, q) == 1
    est = 10.0 * math.log10((p * p + q * q) / (n * n) / T)
    assert st.cn0 == pytest.approx(40.0 + 0.5 * (est - 40.0), rel=1e-13)
    assert st.pli == pytest.approx((p * p - q * q) / (p * p + q * q), rel=1e-12)
    assert st.pli_valid == 1 and st.sumP == 0.0 and st.sumN == 0.0 and st.sumD == 0.0


def run_windows(orc, st, n_win, P, noise):
    flags = []
    for _ in range(n_win):
        for _k in range(500):
            st.lock += 1
            orc.tl_step("cn0", st, [0] * 6 + [noise, 0.0], P[0], P[1])
            if st.idle:
                break
        flags.append((round(st.cn0, 6), st.lost_cnt, st.idle))
        if st.idle:
            break
    return flags


def test_loss_after_four_bad_windows(orc):
    # estimate 20 dB-Hz from a filter state of 40: 30 (not below 30), 25,
    # 22.5, 21.25, 20.625 - the fourth bad window declares the loss
    n = 1.0
    p = math.sqrt(10 ** 2.0 * T) * n        # (p^2 / n^2) / T = 10^(20/10)
    st = orc.tl_state(T=T, cn0=40.0)
    got = run_windows(orc, st, 6, (p, 0.0), n)
    assert [g[0] for g in got] == pytest.approx([30.0, 25.0, 22.5, 21.25, 20.625], abs=1e-9)
    assert [g[1] for g in got] == [0, 1, 2, 3, 4]
    assert got[-1][2] == 1 and st.lock == 0
    # a good window resets the count
    st = orc.tl_state(T=T, cn0=31.0, lost_cnt=3)
    pg = math.sqrt(10 ** 4.5 * T)          # 45 dB-Hz
    run_windows(orc, st, 1, (pg, 0.0), 1.0)
    assert st.lost_cnt == 0 and st.idle == 0


def test_costas_pli_loss(orc):
    # prompt power split equally between I and Q: PLI = 0 < 0.25; a Costas
    # loop counts the window bad despite 45 dB-Hz, a plain PLL does not
    pg = math.sqrt(10 ** 4.5 * T / 2)
    for costas, expect in ((1, [1, 2, 3, 4]), (0, [0, 0, 0, 0])):
        st = orc.tl_state(T=T, costas=costas, cn0=45.0)
        got = run_windows(orc, st, 4, (pg, pg), 1.0)
        assert [g[1] for g in got] == expect
        assert st.pli == pytest.approx(0.0, abs=1e-6)


def test_update_switches_fll_to_pll(orc):
    # the update counts the epoch, runs the FLL while lock T <= 1 s and the
    # PLL after: at lock 1000 (1.0 s) the FLL, at 1001 the PLL
    W = 5.0 / 0.7845
    c = rot(0.1)
    st = orc.tl_state(T=T, lock=999, C0=rot(0.0))
    orc.tl_step("update", st, c)
    assert st.lock == 1000 and st.fd == pytest.approx(-2.0 / 0.25 * (-phase(c)) / TWO_PI, abs=1e-9)
    fd0 = st.fd
    orc.tl_step("update", st, c)
    e = phase(c) / TWO_PI
    assert st.lock == 1001
    assert st.fd == pytest.approx(fd0 + 2.4 * W * e + 1.1 * W * W * e * T + W ** 3 * e * T * T, abs=1e-9)


# --- secondary code (src/sdr_ch.c:453-486, 743-746) -----------------------
# THRES_SEC_RATIO = 0.8, THRES_SYNC = 0.02, THRES_LOST = 0.002, T_NPULLIN =
# 1.5 s; the sync compares the last len_sec in-phase prompts with the code
# from its first chip, so it fires on the epoch that ends a code period.

def nh20():
    from pocketsdr_amd import codes
    return codes.sec_code_NH20().astype(np.int64)


def shift_in(h, v):
    """the history's sdr_add_buff: drop the oldest, append v = (re, im)"""
    h[:-1] = h[1:]
    h[-1] = v


@pytest.mark.parametrize("amp", [0.05, -0.05])
def test_sec_sync_and_wipe(orc, amp):
    # prompts A c_0 .. A c_19: sync at this lock count with the sign of A;
    # the epoch's multiplier c_19 * pol, and every later epoch's the next chip
    code = nh20()
    st = orc.tl_state(T=T, lock=1600)
    h = orc.tl_hist(st)
    h[-20:, 0] = amp * code
    h[-20:, 1] = 0.01
    taps = [amp * code[-1], 0.01, 0.3, -0.2, 0.2, 0.1, 0.1, 0.0]
    m, c = orc.tl_sec(st, code, taps)
    pol = 1 if amp > 0 else -1
    assert st.sec_sync == 1600 and st.sec_pol == pol and m == code[-1] * pol
    assert np.array_equal(c, np.float32(taps) * m)
    assert h[-1, 0] == np.float32(abs(amp)) and h[-1, 1] == np.float32(0.01 * m)
    for k in range(45):                      # two periods and more
        shift_in(h, (amp * code[k % 20], 0.0))
        st.lock += 1
        m, c = orc.tl_sec(st, code, [1.0, 2.0])
        assert m == code[k % 20] * pol, k
        assert h[-1, 0] == np.float32(abs(amp)), k         # wiped: the constant sign
        assert list(c) == [m, 2.0 * m]


def test_sec_sync_thresholds(orc):
    code = nh20()

    def trial(x):
        st = orc.tl_state(T=T, lock=2000)
        h = orc.tl_hist(st)
        h[-20:, 0] = x
        m, _ = orc.tl_sec(st, code, [1.0, 0.0])
        return st.sec_sync, m

    A = 0.05
    flip1 = A * code.astype(np.float64)
    flip1[3] *= -1                            # |P| = 0.9 R
    assert trial(flip1)[0] == 2000
    flip3 = flip1.copy()
    flip3[[7, 11]] *= -1                      # |P| = 0.7 R < 0.8 R
    assert trial(flip3) == (0, 1)
    assert trial(0.021 * code)[0] == 2000     # R = 0.021 >= THRES_SYNC
    assert trial(0.019 * code) == (0, 1)      # R = 0.019 < THRES_SYNC
    # the code at any other phase: NH20's sidelobes stay below 0.8
    for k in range(1, 20):
        assert trial(A * np.roll(code, k)) == (0, 1), k
    # reversed in time: no sync either (the chips are read oldest first)
    assert trial(A * code[::-1])[0] == 0


def test_sec_loss_check_every_period(orc):
    code = nh20()

    def trial(lock, x):
        st = orc.tl_state(T=T, lock=lock, sec_sync=1000, sec_pol=1)
        h = orc.tl_hist(st)
        h[-20:, 0] = x
        m, _ = orc.tl_sec(st, code, [1.0, 0.0])
        return st.sec_sync, st.sec_pol, m

    alt = 0.05 * np.tile([1.0, -1.0], 10)            # mean 0
    assert trial(1020, alt) == (0, 0, 1)             # a period boundary: lost
    assert trial(1021, alt) == (1000, 1, code[0])    # between boundaries: kept
    bias = alt + 0.0025                              # mean 0.0025 >= THRES_LOST
    assert trial(1040, bias) == (1000, 1, code[19])


def test_sec_index_below_sync(orc):
    # adj_coff can take the lock count below the sync point: the chip index
    # (lock - sync - 1 + N) % N; more than N below, the reference's index
    # would go negative, and the oracle wraps it into the code
    code = nh20()
    for lock, idx in ((95, 14), (81, 0), (70, 9)):
        st = orc.tl_state(T=T, lock=lock, sec_sync=100, sec_pol=-1)
        orc.tl_hist(st)
        m, _ = orc.tl_sec(st, code, [1.0, 0.0])
        assert m == -code[idx], lock


@pytest.mark.parametrize("lock0,synced", [(1498, False), (1499, True)])
def test_update_wipes_from_npullin(orc, lock0, synced):
    # prompts -0.05 c_i with the newest pushed by this update: from lock
    # 1500 (T_NPULLIN) the sync fires with polarity -1; the code's last chip
    # is +1, so the multiplier is -1 and a plain PLL (atan2) sees the tap
    # planted at pi + 0.1 as 0.1; one epoch earlier it sees pi + 0.1
    code = nh20()
    assert code[-1] == 1
    st = orc.tl_state(T=T, lock=lock0, fd=10.0)
    h = orc.tl_hist(st)
    orc.tl_sec_code(st, code)
    h[-19:, 0] = -0.05 * code[:19]
    tap = [-math.cos(0.1), -math.sin(0.1)]
    tap = [float(np.float32(v)) for v in tap]
    m = orc.tl_step("update", st, tap + [0.0] * 6, -0.05, 0.0)
    W = 5.0 / 0.7845
    if synced:
        assert m == -1 and st.sec_sync == 1500 and st.sec_pol == -1
        assert h[-1, 0] == np.float32(0.05)
        e = math.atan2(-tap[1], -tap[0]) / TWO_PI
    else:
        assert m == 1 and st.sec_sync == 0 and h[-1, 0] == np.float32(-0.05)
        e = math.atan2(tap[1], tap[0]) / TWO_PI
    assert st.fd == pytest.approx(10.0 + 2.4 * W * e + 1.1 * W * W * e * T + W ** 3 * e * T * T, abs=1e-9)


# --- bump-jump (src/sdr_ch.c:551-572, 586-601): BUMP_K = 1.3 ----------------

@pytest.mark.parametrize("ve,vl,move", [(1.2, 0.5, -1), (0.5, 1.2, 1), (1.1, 0.5, 0), (1.2, 0.9, 0),
                                        (0.9, 1.2, 0)])
def test_bump_jump_window(orc, ve, vl, move):
    step = T / 8184
    st = orc.tl_state(T=T, lock=499, coff=2e-4, bump_step=step, cn0=40.0)
    corr = [0.0] * 6 + [0.1, 0.0, ve, 0.0, vl, 0.0]
    orc.tl_step("cn0", st, corr, 1.0, 0.0)          # not a window end: sums only
    assert st.coff == 2e-4
    assert st.sumVE == pytest.approx(ve * ve, rel=1e-6) and st.sumVL == pytest.approx(vl * vl, rel=1e-6)
    st.sumVE = st.sumVL = st.sumP = st.sumN = st.sumD = 0.0
    st.lock = 500
    assert orc.tl_step("cn0", st, corr, 1.0, 0.0) == 1
    assert st.coff == 2e-4 + move * step
    assert st.sumVE == 0.0 and st.sumVL == 0.0


def test_no_bump_without_step(orc):
    st = orc.tl_state(T=T, lock=500, coff=2e-4, cn0=40.0)
    orc.tl_step("cn0", st, [0.0] * 6 + [0.1, 0.0, 1.5, 0.0, 0.1, 0.0], 1.0, 0.0)
    assert st.coff == 2e-4 and st.sumVE == 0.0
