"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP library.  Loaded by tests/,
by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py, never by
the product package.  When oracle/_ref/libpocketfft_ref.so (the reference's
PocketFFT wrapper, lib/pocketfft/pocketfft_wrap.cpp, built from the reference
sources by ``make -C oracle ref``) is present, the FP32 FFTs go through it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_FFT = os.path.join(HERE, "_ref", "libpocketfft_ref.so")

if not os.path.exists(LIB):
    raise ImportError("oracle/liboracle.so not built (make -C oracle)")
lib = C.CDLL(LIB)

vp, fp, dp, ip = C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int)


def _d(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args


class orc_trk_ch_t(C.Structure):
    _fields_ = [("fs", C.c_double), ("fi", C.c_double), ("fc_sig", C.c_double),
                ("T", C.c_double), ("N", C.c_int), ("fd0", C.c_double), ("drift", C.c_double),
                ("phi0", C.c_double), ("coff0", C.c_double)]


_d("orc_func_init", None, [])
_d("orc_mix_tbl", C.POINTER(C.c_int8), [])
_d("orc_load_ref_fft", C.c_int, [C.c_char_p])
_d("orc_ref_fft_loaded", C.c_int, [])
_d("orc_mix_carr", None, [vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, vp])
_d("orc_corr_std", None, [vp, vp, C.c_int, C.c_double, dp, C.c_int, vp, vp])
_d("orc_corr_std_cpx_code", None, [vp, vp, C.c_int, C.c_double, dp, C.c_int, vp, vp])
_d("orc_cpx_mul", None, [vp, vp, C.c_int, C.c_float, vp])
_d("orc_fft_f32", C.c_int, [vp, C.c_int, C.c_int, vp])
_d("orc_fft_f64", None, [vp, C.c_int, C.c_int, vp])
_d("orc_corr_fft", None, [vp, vp, C.c_int, vp])
_d("orc_search_code", None, [vp, C.c_double, vp, C.c_int, C.c_int, C.c_int, C.c_double,
                             C.c_double, fp, C.c_int, fp])
_d("orc_search_code_f64", None, [vp, C.c_double, vp, C.c_int, C.c_int, C.c_int, C.c_double,
                                 C.c_double, fp, C.c_int, dp])
_d("orc_corr_max", C.c_float, [fp, C.c_int, C.c_int, C.c_int, C.c_double, ip])
_d("orc_fine_dop", C.c_double, [fp, C.c_int, fp, C.c_int, ip])
_d("orc_dop_bins", C.c_int, [C.c_double, C.c_float, C.c_float, fp])
_d("orc_res_code", None, [vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, vp])
_d("orc_gen_code_fft", None, [vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, vp])
_d("orc_gen_code_fft_f64", None, [vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                  C.c_int, vp])
_d("orc_gen_code_L1CA", C.c_int, [C.c_int, vp])
_d("orc_trk_run", None, [C.POINTER(orc_trk_ch_t), vp, vp, C.c_int, C.c_int, dp, C.c_int, vp, vp, dp])
_d("orc_trk_stitch", None, [vp, C.c_int, C.c_int, vp])
_d("orc_l6l_run", None, [C.POINTER(orc_trk_ch_t), C.c_int, C.c_int, C.c_double, vp, vp, C.c_int, C.c_int,
                         dp, C.c_int, dp, vp, vp])
_d("orc_trkl_run", None, [C.POINTER(orc_trk_ch_t), C.c_int, C.c_int, C.c_double, vp, vp, C.c_int,
                          C.c_int, dp, C.c_int, dp, vp, vp, C.c_int, vp, C.c_double])
class orc_tl_t(C.Structure):
    _fields_ = [("T", C.c_double), ("len_code", C.c_int), ("costas", C.c_int),
                ("fd", C.c_double), ("coff", C.c_double), ("cn0", C.c_double), ("pli", C.c_double),
                ("err_phas", C.c_double), ("phas_acc", C.c_double), ("code_int", C.c_double),
                ("sumE", C.c_double), ("sumL", C.c_double), ("sumP", C.c_double), ("sumN", C.c_double),
                ("sumD", C.c_double), ("C0", C.c_float * 2),
                ("lock", C.c_int), ("lost_cnt", C.c_int), ("pli_valid", C.c_int), ("idle", C.c_int),
                ("len_sec", C.c_int), ("sec_code", vp), ("sec_sync", C.c_int), ("sec_pol", C.c_int),
                ("bump_step", C.c_double), ("sumVE", C.c_double), ("sumVL", C.c_double),
                ("hist", fp)]


_d("orc_tl_fll", None, [C.POINTER(orc_tl_t), fp])
_d("orc_tl_pll", None, [C.POINTER(orc_tl_t), fp])
_d("orc_tl_dll", None, [C.POINTER(orc_tl_t), fp])
_d("orc_tl_cn0", C.c_int, [C.POINTER(orc_tl_t), fp, C.c_float, C.c_float])
_d("orc_tl_update", C.c_int, [C.POINTER(orc_tl_t), fp, C.c_float, C.c_float])
_d("orc_tl_push", None, [C.POINTER(orc_tl_t), C.c_float, C.c_float])
_d("orc_tl_sec", C.c_int, [C.POINTER(orc_tl_t), fp, C.c_int])
_d("orc_csk", C.c_int, [vp, C.c_int, C.c_int, dp, C.c_int, vp])
_d("orc_trk_l6_run", None, [C.POINTER(orc_trk_ch_t), C.c_int, vp, vp, C.c_int, C.c_int, dp, C.c_int,
                            vp, vp, vp, dp])
_d("orc_gen_lutc", None, [C.c_int, C.c_int, C.c_int, C.c_int, vp])
_d("orc_gen_lutc_rtoc", None, [C.c_int, vp])
_d("orc_gen_lut8", None, [C.c_int, C.c_double, C.c_double, vp])
_d("orc_write_buff", C.c_int, [C.c_int, vp, C.c_int, C.c_longlong, ip, C.POINTER(vp), C.POINTER(vp)])
_d("orc_arch_combine", None, [C.c_int, C.POINTER(vp), C.c_int, vp, vp])

lib.orc_func_init()
REF_LOADED = bool(os.path.exists(REF_FFT) and lib.orc_load_ref_fft(REF_FFT.encode()))


def _p(a):
    return a.ctypes.data_as(vp)


def mix_tbl():
    return np.ctypeslib.as_array(lib.orc_mix_tbl(), shape=(65536 * 2,)).reshape(65536, 2).copy()


def mix_carr(buff, ix, N, fs, fc, phi):
    buff = np.ascontiguousarray(buff, np.uint8)
    IQ = np.zeros((N, 2), np.int8)
    lib.orc_mix_carr(_p(buff), len(buff), ix, N, fs, fc, phi, _p(IQ))
    return IQ


def corr_std(IQ, code, N, coff, pos, cpx_code=False):
    IQ = np.ascontiguousarray(IQ, np.int8)
    code = np.ascontiguousarray(code, np.int8)
    pos = np.ascontiguousarray(pos, np.float64)
    corr = np.zeros(len(pos), np.complex64)
    Cc = np.zeros(2, np.complex64)
    f = lib.orc_corr_std_cpx_code if cpx_code else lib.orc_corr_std
    f(_p(IQ), _p(code), N, coff, pos.ctypes.data_as(dp), len(pos), _p(corr), _p(Cc))
    return corr, Cc


def cpx_mul(a, b, s):
    a = np.ascontiguousarray(a, np.complex64)
    b = np.ascontiguousarray(b, np.complex64)
    c = np.zeros_like(a)
    lib.orc_cpx_mul(_p(a), _p(b), len(a), s, _p(c))
    return c


def fft_f32(x, dir=0):
    x = np.ascontiguousarray(x, np.complex64)
    y = np.zeros_like(x)
    lib.orc_fft_f32(_p(x), len(x), dir, _p(y))
    return y


def fft_f64(x, dir=0):
    x = np.ascontiguousarray(x, np.complex128)
    y = np.zeros_like(x)
    lib.orc_fft_f64(_p(x), len(x), dir, _p(y))
    return y


def corr_fft(IQ, code_fft):
    IQ = np.ascontiguousarray(IQ, np.int8)
    code_fft = np.ascontiguousarray(code_fft, np.complex64)
    out = np.zeros(len(code_fft), np.complex64)
    lib.orc_corr_fft(_p(IQ), _p(code_fft), len(code_fft), _p(out))
    return out


def search_code(code_fft, T, buff, ix, fs, fi, fds, P=None):
    code_fft = np.ascontiguousarray(code_fft, np.complex64)
    buff = np.ascontiguousarray(buff, np.uint8)
    fds = np.ascontiguousarray(fds, np.float32)
    N = len(code_fft)
    if P is None:
        P = np.zeros((len(fds), N), np.float32)
    lib.orc_search_code(_p(code_fft), T, _p(buff), len(buff), ix, N, fs, fi,
                        fds.ctypes.data_as(fp), len(fds), P.ctypes.data_as(fp))
    return P


def search_code_f64(code_fft, T, buff, ix, fs, fi, fds, P=None):
    code_fft = np.ascontiguousarray(code_fft, np.complex128)
    buff = np.ascontiguousarray(buff, np.uint8)
    fds = np.ascontiguousarray(fds, np.float32)
    N = len(code_fft)
    if P is None:
        P = np.zeros((len(fds), N), np.float64)
    lib.orc_search_code_f64(_p(code_fft), T, _p(buff), len(buff), ix, N, fs, fi,
                            fds.ctypes.data_as(fp), len(fds), P.ctypes.data_as(dp))
    return P


def corr_max(P, N, Nmax, M, T, ix=None):
    P = np.ascontiguousarray(P, np.float32)
    ix = np.zeros(2, np.int32) if ix is None else np.ascontiguousarray(ix, np.int32)
    cn0 = lib.orc_corr_max(P.ctypes.data_as(fp), N, Nmax, M, T, ix.ctypes.data_as(ip))
    return cn0, ix


def fine_dop(P, N, fds, ix):
    P = np.ascontiguousarray(P, np.float32)
    fds = np.ascontiguousarray(fds, np.float32)
    ix = np.ascontiguousarray(ix, np.int32)
    return lib.orc_fine_dop(P.ctypes.data_as(fp), N, fds.ctypes.data_as(fp), len(fds),
                            ix.ctypes.data_as(ip))


def dop_bins(T, dop, max_dop):
    n = lib.orc_dop_bins(T, dop, max_dop, None)
    out = np.zeros(n, np.float32)
    lib.orc_dop_bins(T, dop, max_dop, out.ctypes.data_as(fp))
    return out


def res_code(code, T, coff, fs, N, Nz=0):
    code = np.ascontiguousarray(code, np.int8)
    out = np.zeros((N + Nz, 2), np.int8)
    lib.orc_res_code(_p(code), len(code), T, coff, fs, N, Nz, _p(out))
    return out


def gen_code_fft(code, T, coff, fs, N, Nz):
    code = np.ascontiguousarray(code, np.int8)
    out = np.zeros(N + Nz, np.complex64)
    lib.orc_gen_code_fft(_p(code), len(code), T, coff, fs, N, Nz, _p(out))
    return out


def gen_code_fft_f64(code, T, coff, fs, N, Nz):
    code = np.ascontiguousarray(code, np.int8)
    out = np.zeros(N + Nz, np.complex128)
    lib.orc_gen_code_fft_f64(_p(code), len(code), T, coff, fs, N, Nz, _p(out))
    return out


def gen_code_L1CA(prn):
    out = np.zeros(1023, np.int8)
    if not lib.orc_gen_code_L1CA(prn, _p(out)):
        raise ValueError("bad PRN")
    return out


def code_bank(code, T, fs, N):
    return np.concatenate([res_code(code, T, -k / fs / 10, fs, N) for k in range(10)])


def trk_run(ch, code_bank_arr, buff, n_epoch, pos):
    """ch: dict fs, fi, fc_sig, T, N, fd0, drift, phi0, coff0"""
    c = orc_trk_ch_t(ch["fs"], ch["fi"], ch["fc_sig"], ch["T"], ch["N"], ch["fd0"],
                     ch["drift"], ch["phi0"], ch["coff0"])
    code_bank_arr = np.ascontiguousarray(code_bank_arr, np.int8)
    buff = np.ascontiguousarray(buff, np.uint8)
    pos = np.ascontiguousarray(pos, np.float64)
    corr = np.zeros((n_epoch, len(pos)), np.complex64)
    Cc = np.zeros((n_epoch, 2), np.complex64)
    state = np.zeros((n_epoch, 3), np.float64)
    lib.orc_trk_run(C.byref(c), _p(code_bank_arr), _p(buff), len(buff), n_epoch,
                    pos.ctypes.data_as(dp), len(pos), _p(corr), _p(Cc), state.ctypes.data_as(dp))
    return corr, Cc, state


def trkl_run(ch, len_code, costas, cn0, code_bank_arr, buff, n_epoch, pos, sec_code=None, bump_step=0.0):
    """Closed-loop tracking of one channel (src/sdr_ch.c:488-624, 694-763):
    state (n_epoch, 6) = fd, coff, phi, cn0, lock, idle; corr (n_epoch,
    n_pos); P (n_epoch,).  sec_code: the secondary code (int8 +-1) to
    synchronise and wipe off; bump_step > 0: taps 4 and 5 are VE and VL and
    bump-jump moves the code by that step (s)."""
    sec = None if sec_code is None else np.ascontiguousarray(sec_code, np.int8)
    c = orc_trk_ch_t(ch["fs"], ch["fi"], ch["fc_sig"], ch["T"], ch["N"], ch["fd0"],
                     ch.get("drift", 0.0), ch["phi0"], ch["coff0"])
    code_bank_arr = np.ascontiguousarray(code_bank_arr, np.int8)
    buff = np.ascontiguousarray(buff, np.uint8)
    pos = np.ascontiguousarray(pos, np.float64)
    state = np.zeros((n_epoch, 6), np.float64)
    corr = np.zeros((n_epoch, len(pos)), np.complex64)
    P = np.zeros(n_epoch, np.complex64)
    lib.orc_trkl_run(C.byref(c), len_code, costas, cn0, _p(code_bank_arr), _p(buff), len(buff), n_epoch,
                     pos.ctypes.data_as(dp), len(pos), state.ctypes.data_as(dp), _p(corr), _p(P),
                     0 if sec is None else len(sec), None if sec is None else _p(sec), float(bump_step))
    return state, corr, P


def trk_stitch(Cc, N):
    """Prompt history (src/sdr_ch.c:736-740) from trk_run's C (n_epoch, 2)."""
    Cc = np.ascontiguousarray(Cc, np.complex64)
    P = np.zeros(len(Cc), np.complex64)
    lib.orc_trk_stitch(_p(Cc), len(Cc), N, _p(P))
    return P


def csk(corr, len_code, pos):
    """L6 CSK (src/sdr_ch.c:627-664) of one correlation: (ix, sym, taps)."""
    corr = np.ascontiguousarray(corr, np.complex64)
    pos = np.ascontiguousarray(pos, np.float64)
    taps = np.zeros(len(pos), np.complex64)
    ix = lib.orc_csk(_p(corr), len(corr), len_code, pos.ctypes.data_as(dp), len(pos), _p(taps))
    return ix, (255 - ix) & 0xFF, taps


def trk_l6_run(ch, len_code, code_fft_banks, buff, n_epoch, pos, keep_corr=False):
    """L6D / L6E FFT tracking of one channel (src/sdr_ch.c:694-720).
    code_fft_banks: (10, N) complex64.  Returns syms, taps, corr (or None),
    state (n_epoch x {phi, coff, i, j})."""
    c = orc_trk_ch_t(ch["fs"], ch["fi"], ch["fc_sig"], ch["T"], ch["N"], ch["fd0"],
                     ch["drift"], ch["phi0"], ch["coff0"])
    cf = np.ascontiguousarray(code_fft_banks, np.complex64)
    buff = np.ascontiguousarray(buff, np.uint8)
    pos = np.ascontiguousarray(pos, np.float64)
    N = ch["N"]
    syms = np.zeros(n_epoch, np.uint8)
    taps = np.zeros((n_epoch, len(pos)), np.complex64)
    corr = np.zeros((n_epoch, N), np.complex64) if keep_corr else None
    state = np.zeros((n_epoch, 4), np.float64)
    lib.orc_trk_l6_run(C.byref(c), len_code, _p(cf), _p(buff), len(buff), n_epoch,
                       pos.ctypes.data_as(dp), len(pos), _p(syms), _p(taps),
                       _p(corr) if keep_corr else None, state.ctypes.data_as(dp))
    return syms, taps, corr, state

def l6l_run(ch, len_code, cn0, code_fft_banks, buff, n_epoch, pos, costas=0):
    """L6D / L6E closed-loop tracking of one channel (src/sdr_ch.c:694-720,
    741-763): state (n_epoch, 6) = fd, coff, phi, cn0, lock, idle; syms
    (n_epoch,); taps (n_epoch, n_pos)."""
    c = orc_trk_ch_t(ch["fs"], ch["fi"], ch["fc_sig"], ch["T"], ch["N"], ch["fd0"],
                     ch.get("drift", 0.0), ch["phi0"], ch["coff0"])
    cf = np.ascontiguousarray(code_fft_banks, np.complex64)
    buff = np.ascontiguousarray(buff, np.uint8)
    pos = np.ascontiguousarray(pos, np.float64)
    state = np.zeros((n_epoch, 6), np.float64)
    syms = np.zeros(n_epoch, np.uint8)
    taps = np.zeros((n_epoch, len(pos)), np.complex64)
    lib.orc_l6l_run(C.byref(c), len_code, costas, cn0, _p(cf), _p(buff), len(buff), n_epoch,
                    pos.ctypes.data_as(dp), len(pos), state.ctypes.data_as(dp), _p(syms), _p(taps))
    return state, syms, taps


# front-end raw data to cpx8 (src/sdr_rcv.c:776-936)
FMT_INT8, FMT_INT8X2, FMT_RAW8, FMT_RAW16, FMT_RAW16I, FMT_RAW32, FMT_CS8, FMT_CS16 = range(1, 9)
RAW_NCH = {FMT_INT8: 1, FMT_INT8X2: 1, FMT_CS8: 1, FMT_CS16: 1, FMT_RAW8: 2, FMT_RAW16: 4,
           FMT_RAW16I: 8, FMT_RAW32: 8}
RAW_BPS = {FMT_INT8: 1, FMT_INT8X2: 2, FMT_CS8: 2, FMT_CS16: 4, FMT_RAW8: 1, FMT_RAW16: 2,
           FMT_RAW16I: 2, FMT_RAW32: 4}


def gen_lutc(fmt, rfch, IQ, bits):
    out = np.zeros(256, np.uint8)
    lib.orc_gen_lutc(fmt, rfch, IQ, bits, _p(out))
    return out


def gen_lutc_rtoc(rfch):
    out = np.zeros(1024, np.uint8)
    lib.orc_gen_lutc_rtoc(rfch, _p(out))
    return out


def gen_lut8(fmt, off, scale):
    out = np.zeros(65536 if fmt == FMT_CS16 else 256, np.int8)
    lib.orc_gen_lut8(fmt, off, scale, _p(out))
    return out


def write_buff(fmt, raw, N, ix, rtoc, luts):
    """Per-channel cpx8 arrays (nch x N) of N samples of block ix.  luts: per
    channel cpx8 LUT (raw formats) or [LUT8, LUTQ or None] (int8 formats)."""
    raw = np.ascontiguousarray(raw, np.uint8)
    nch = RAW_NCH[fmt]
    out = np.zeros((8, N), np.uint8)
    luts = [None if l is None else np.ascontiguousarray(l) for l in luts] + [None] * 8
    lp = (vp * 8)(*[None if l is None else l.ctypes.data for l in luts[:8]])
    op = (vp * 8)(*[out[k].ctypes.data for k in range(8)])
    rt = (C.c_int * 8)(*(list(rtoc) + [0] * 8)[:8])
    n = lib.orc_write_buff(fmt, _p(raw), N, ix, rt, lp, op)
    assert n == nch
    return out[:nch]


def arch_combine(ins, w):
    """ins: (nrfch, N) uint8 cpx8; w: (nrfch, 2) int16 Q8 weights -> (N,) cpx8."""
    ins = np.ascontiguousarray(ins, np.uint8)
    w = np.ascontiguousarray(w, np.int16)
    nrfch, N = ins.shape
    out = np.zeros(N, np.uint8)
    ip_ = (vp * nrfch)(*[ins[j].ctypes.data for j in range(nrfch)])
    lib.orc_arch_combine(nrfch, ip_, N, _p(w), _p(out))
    return out



def tl_state(T=1e-3, len_code=1023, costas=0, **kw):
    """A closed-loop state (orc_tl_t) with the given fields."""
    st = orc_tl_t()
    st.T, st.len_code, st.costas, st.pli = T, len_code, costas, 1.0
    for k, v in kw.items():
        if k == "C0":
            st.C0[0], st.C0[1] = v
        else:
            setattr(st, k, v)
    return st



def tl_step(step, st, corr, Px=0.0, Py=0.0):
    """One loop step of the oracle on state st: step "fll", "pll", "dll",
    "cn0" or "update"; corr = the epoch's correlators as (re, im) pairs
    P, E, L, N, VE, VL (12 floats, missing ones zero).  Returns orc_tl_cn0's
    window flag for "cn0", orc_tl_update's secondary-code multiplier for
    "update", else None."""
    c = np.zeros(12, np.float32)
    c[:len(corr)] = corr
    ref = C.byref(st)
    cp = c.ctypes.data_as(fp)
    if step == "cn0":
        return lib.orc_tl_cn0(ref, cp, Px, Py)
    if step == "update":
        return lib.orc_tl_update(ref, cp, Px, Py)
    getattr(lib, "orc_tl_" + step)(ref, cp)
    return None


def tl_hist(st):
    """Attach a zeroed prompt history (SDR_N_HIST entries) to state st;
    returns the (n, 2) float32 array behind it."""
    h = np.zeros((5000, 2), np.float32)
    st._hist = h
    st.hist = h.ctypes.data_as(fp)
    return h


def tl_sec_code(st, sec_code):
    """Attach a secondary code (int8 +-1) to state st."""
    code = np.ascontiguousarray(sec_code, np.int8)
    st._sec = code
    st.len_sec = len(code)
    st.sec_code = code.ctypes.data_as(vp)


def tl_sec(st, sec_code, taps):
    """orc_tl_sec on state st (its history attached) with the secondary code
    and the epoch's taps ((re, im) pairs, wiped in place); returns the
    multiplier and the wiped taps."""
    tl_sec_code(st, sec_code)
    c = np.ascontiguousarray(taps, np.float32).copy()
    m = lib.orc_tl_sec(C.byref(st), c.ctypes.data_as(fp), len(c) // 2)
    return m, c
