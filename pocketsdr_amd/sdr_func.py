"""numpy front end of libpocketsdr_amd (the search / correlate hot path).

Function names and argument meaning follow the reference C API
(src/pocket_sdr.h:391-465); arrays follow its layouts:

* IF buffer ``buff``: uint8 array of cpx8 samples (I low nibble, Q high nibble)
* mixed samples (cpx16): int8 array of shape (N, 2)
* complex arrays (sdr_cpx_t): complex64 arrays
"""
import ctypes as C

import numpy as np

from . import lib

c_float_p = C.POINTER(C.c_float)
c_double_p = C.POINTER(C.c_double)
c_int_p = C.POINTER(C.c_int)
c_void_p = C.c_void_p


class sdr_buff_t(C.Structure):
    """src/pocket_sdr.h:242-245"""
    _fields_ = [("data", C.POINTER(C.c_uint8)), ("IQ", C.c_int), ("N", C.c_int)]


class sdr_search_sig_t(C.Structure):
    _fields_ = [("buff", C.c_int), ("M", C.c_int), ("N", C.c_int),
                ("T", C.c_double), ("fs", C.c_double), ("fi", C.c_double),
                ("n_prn", C.c_int), ("code_fft", c_void_p),
                ("len_fds", C.c_int), ("fds", c_float_p),
                ("n_blk", C.c_int), ("ix0", C.c_int), ("blk_step", C.c_int)]


class sdr_raw_stream_t(C.Structure):
    _fields_ = [("raw", c_void_p), ("N", C.c_int), ("fmt", C.c_int), ("rfch", C.c_int),
                ("rtoc", C.c_int), ("lut", c_void_p), ("ph0", C.c_int)]


class sdr_trk_ch_t(C.Structure):
    _fields_ = [("fs", C.c_double), ("fi", C.c_double), ("fc_sig", C.c_double),
                ("T", C.c_double), ("N", C.c_int), ("fd0", C.c_double),
                ("drift", C.c_double), ("phi0", C.c_double), ("coff0", C.c_double),
                ("code", c_void_p)]


class sdr_trk_loop_t(C.Structure):
    _fields_ = [("len_code", C.c_int), ("costas", C.c_int), ("cn0", C.c_double),
                ("len_sec_code", C.c_int), ("sec_code", c_void_p), ("bump_step", C.c_double)]


class sdr_l6_ch_t(C.Structure):
    _fields_ = [("fs", C.c_double), ("fi", C.c_double), ("fc_sig", C.c_double),
                ("T", C.c_double), ("N", C.c_int), ("len_code", C.c_int), ("fd0", C.c_double),
                ("drift", C.c_double), ("phi0", C.c_double), ("coff0", C.c_double),
                ("code_fft", c_void_p)]


def _decl(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_decl("sdr_func_init", None, [C.c_char_p])
_decl("sdr_cpx_mul", None, [c_void_p, c_void_p, C.c_int, C.c_float, c_void_p])
_decl("sdr_cpx_fft", C.c_int, [c_void_p, C.c_int, C.c_int, c_void_p])
_decl("sdr_search_code", None, [c_void_p, C.c_double, C.POINTER(sdr_buff_t), C.c_int, C.c_int,
                                C.c_double, C.c_double, c_float_p, C.c_int, c_float_p])
_decl("sdr_search_code_noalloc", C.c_int, [c_void_p, C.c_double, C.POINTER(sdr_buff_t), C.c_int, C.c_int,
                                        C.c_double, C.c_double, c_float_p, C.c_int, c_float_p])
_decl("sdr_corr_max", C.c_float, [c_float_p, C.c_int, C.c_int, C.c_int, C.c_double, c_int_p])
_decl("sdr_fine_dop", C.c_double, [c_float_p, C.c_int, c_float_p, C.c_int, c_int_p])
_decl("sdr_dop_bins", c_float_p, [C.c_double, C.c_float, C.c_float, c_int_p])
_decl("sdr_corr_std", None, [c_void_p, c_void_p, C.c_int, C.c_double, c_double_p, C.c_int,
                             c_void_p, c_void_p])
_decl("sdr_corr_std_cpx_code", None, [c_void_p, c_void_p, C.c_int, C.c_double, c_double_p,
                                      C.c_int, c_void_p, c_void_p])
_decl("sdr_corr_std_cpx", None, [c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                 C.c_double, c_float_p, c_double_p, C.c_int, c_void_p])
_decl("sdr_corr_fft", None, [c_void_p, c_void_p, C.c_int, c_void_p])
_decl("sdr_corr_fft_cpx", None, [c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                 C.c_double, c_void_p, c_void_p])
_decl("sdr_mix_carr", None, [C.POINTER(sdr_buff_t), C.c_int, C.c_int, C.c_double, C.c_double,
                             C.c_double, c_void_p])
_decl("sdr_search_plan_new", c_void_p, [C.c_int, C.POINTER(C.POINTER(sdr_buff_t)), C.c_int,
                                        C.POINTER(sdr_search_sig_t), C.c_int])
_decl("sdr_search_plan_new_dev", c_void_p, [C.c_int, C.POINTER(c_void_p), c_int_p, C.c_int,
                                             C.POINTER(sdr_search_sig_t), C.c_int])
_decl("sdr_search_plan_new_raw", c_void_p, [C.c_int, C.POINTER(sdr_raw_stream_t), C.c_int,
                                             C.POINTER(sdr_search_sig_t), C.c_int])
_decl("sdr_search_plan_run", C.c_int, [c_void_p])
_decl("sdr_search_plan_result", C.c_int, [c_void_p, C.c_int, c_float_p, c_int_p, c_double_p])
_decl("sdr_search_plan_power", C.c_int, [c_void_p, C.c_int, C.c_int, c_float_p])
_decl("sdr_search_plan_form", C.c_int, [c_void_p, C.c_int])
_decl("sdr_search_plan_tiles", C.c_int, [c_void_p, C.c_int, c_int_p])
_decl("sdr_search_pass_shape", C.c_int, [C.c_int, C.c_int, C.c_int, c_int_p])
_decl("sdr_fft_pass_groups", C.c_int, [C.c_int])
_decl("sdr_search_plan_timing", None, [c_void_p, C.c_int])
_decl("sdr_search_plan_kernel_ms", C.c_int, [c_void_p, C.c_char_p, c_double_p,
                                             C.POINTER(C.c_longlong), c_double_p])
_decl("sdr_search_plan_kernel_bytes", C.c_int, [c_void_p, C.c_char_p, c_double_p])
_decl("sdr_search_plan_kernel_grid", C.c_int, [c_void_p, C.c_char_p, c_double_p])
_decl("sdr_search_plan_free", None, [c_void_p])
_decl("sdr_search_plan_bytes", C.c_longlong, [])
_decl("sdr_trk_plan_new", c_void_p, [C.POINTER(sdr_buff_t), C.c_int, C.POINTER(sdr_trk_ch_t),
                                     C.c_int, c_double_p, C.c_int])
_decl("sdr_trk_plan_run", C.c_int, [c_void_p])
_decl("sdr_trk_plan_result", C.c_int, [c_void_p, c_void_p, c_void_p])
_decl("sdr_trk_plan_prompt", C.c_int, [c_void_p, c_void_p])
_decl("sdr_trk_plan_new_loop", c_void_p, [C.POINTER(sdr_buff_t), C.c_int, C.POINTER(sdr_trk_ch_t),
                                          C.POINTER(sdr_trk_loop_t), C.c_int, c_double_p, C.c_int])
_decl("sdr_trk_plan_state", C.c_int, [c_void_p, c_void_p])
_decl("sdr_trk_plan_timing", None, [c_void_p, C.c_int])
_decl("sdr_trk_plan_kernel_ms", C.c_int, [c_void_p, C.c_char_p, c_double_p,
                                          C.POINTER(C.c_longlong), c_double_p])
_decl("sdr_trk_plan_free", None, [c_void_p])
_decl("sdr_l6_plan_new", c_void_p, [C.POINTER(sdr_buff_t), C.c_int, C.POINTER(sdr_l6_ch_t),
                                    C.c_int, c_double_p, C.c_int, C.c_int])
_decl("sdr_l6_plan_new_dev", c_void_p, [c_void_p, C.c_int, C.c_int, C.POINTER(sdr_l6_ch_t),
                                        C.c_int, c_double_p, C.c_int, C.c_int])
_decl("sdr_l6_plan_new_loop", c_void_p, [C.POINTER(sdr_buff_t), C.c_int, C.POINTER(sdr_l6_ch_t),
                                         C.POINTER(sdr_trk_loop_t), C.c_int, c_double_p, C.c_int, C.c_int])
_decl("sdr_l6_plan_state", C.c_int, [c_void_p, c_void_p])
_decl("sdr_l6_plan_run", C.c_int, [c_void_p])
_decl("sdr_l6_plan_result", C.c_int, [c_void_p, c_void_p, c_void_p])
_decl("sdr_l6_plan_corr", C.c_int, [c_void_p, C.c_int, C.c_int, c_void_p])
_decl("sdr_l6_plan_timing", None, [c_void_p, C.c_int])
_decl("sdr_l6_plan_kernel_ms", C.c_int, [c_void_p, C.c_char_p, c_double_p,
                                         C.POINTER(C.c_longlong), c_double_p])
_decl("sdr_l6_plan_free", None, [c_void_p])
_decl("sdr_csk", C.c_int, [c_void_p, C.c_int, C.c_int, C.c_int, c_double_p, C.c_int, c_void_p,
                           c_void_p])
_decl("sdr_raw_gen_lut", None, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, c_void_p])
_decl("sdr_raw_gen_lut8", None, [C.c_int, C.c_double, C.c_double, c_void_p])
_decl("sdr_raw_plan_new", c_void_p, [C.c_int, c_int_p, C.POINTER(c_void_p)])
_decl("sdr_raw_plan_nch", C.c_int, [c_void_p])
_decl("sdr_raw_unpack", C.c_int, [c_void_p, c_void_p, C.c_int, C.c_int64, C.POINTER(c_void_p)])
_decl("sdr_raw_unpack_dev", C.c_int, [c_void_p, c_void_p, C.c_longlong, C.c_int64, C.POINTER(c_void_p)])
_decl("sdr_raw_plan_kernel_ms", C.c_double, [c_void_p])
_decl("sdr_raw_plan_free", None, [c_void_p])
_decl("sdr_arch_combine_raw", C.c_int, [C.c_int, C.POINTER(c_void_p), C.c_int, C.c_int, c_void_p,
                                     C.POINTER(c_void_p)])
_decl("sdr_arch_combine_raw_dev", C.c_double, [C.c_int, C.POINTER(c_void_p), C.c_longlong, C.c_int,
                                            c_void_p, C.POINTER(c_void_p)])
_decl("sdr_amd_version", C.c_char_p, [])


def _ptr(a):
    return a.ctypes.data_as(c_void_p)


def _fp(a):
    return a.ctypes.data_as(c_float_p)


def _dp(a):
    return a.ctypes.data_as(c_double_p)


def _buff(data):
    data = np.ascontiguousarray(data, dtype=np.uint8)
    b = sdr_buff_t(data.ctypes.data_as(C.POINTER(C.c_uint8)), 2, len(data))
    b._keep = data
    return b


def func_init():
    lib.sdr_func_init(b"")


def mix_carr(buff, ix, N, fs, fc, phi):
    """sdr_mix_carr (src/sdr_func.c:824-836) -> int8 (N, 2)"""
    b = _buff(buff)
    IQ = np.zeros((N, 2), dtype=np.int8)
    lib.sdr_mix_carr(C.byref(b), ix, N, fs, fc, phi, _ptr(IQ))
    return IQ


def corr_std(IQ, code, N, coff, pos, cpx_code=False):
    """sdr_corr_std / sdr_corr_std_cpx_code -> (corr[n] complex64, C[2])"""
    IQ = np.ascontiguousarray(IQ, dtype=np.int8)
    code = np.ascontiguousarray(code, dtype=np.int8)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n = len(pos)
    corr = np.zeros(n, dtype=np.complex64)
    Cc = np.zeros(2, dtype=np.complex64)
    f = lib.sdr_corr_std_cpx_code if cpx_code else lib.sdr_corr_std
    f(_ptr(IQ), _ptr(code), N, coff, _dp(pos), n, _ptr(corr), _ptr(Cc))
    return corr, Cc


def corr_std_cpx(buff, ix, N, fs, fc, phi, code, pos):
    """sdr_corr_std_cpx (src/sdr_func.c:1060-1072), buff complex64"""
    buff = np.ascontiguousarray(buff, dtype=np.complex64)
    code = np.ascontiguousarray(code, dtype=np.float32)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    corr = np.zeros(len(pos), dtype=np.complex64)
    lib.sdr_corr_std_cpx(_ptr(buff), len(buff), ix, N, fs, fc, phi, _fp(code), _dp(pos),
                         len(pos), _ptr(corr))
    return corr


def corr_fft(IQ, code_fft):
    """sdr_corr_fft (src/sdr_func.c:1088-1102)"""
    IQ = np.ascontiguousarray(IQ, dtype=np.int8)
    code_fft = np.ascontiguousarray(code_fft, dtype=np.complex64)
    N = len(code_fft)
    corr = np.zeros(N, dtype=np.complex64)
    lib.sdr_corr_fft(_ptr(IQ), _ptr(code_fft), N, _ptr(corr))
    return corr


def corr_fft_cpx(buff, ix, N, fs, fc, phi, code_fft):
    buff = np.ascontiguousarray(buff, dtype=np.complex64)
    code_fft = np.ascontiguousarray(code_fft, dtype=np.complex64)
    corr = np.zeros(N, dtype=np.complex64)
    lib.sdr_corr_fft_cpx(_ptr(buff), len(buff), ix, N, fs, fc, phi, _ptr(code_fft), _ptr(corr))
    return corr


def cpx_fft(x, dir=0):
    """sdr_cpx_fft (src/sdr_func.c:236-264); raises on unsupported length"""
    x = np.ascontiguousarray(x, dtype=np.complex64)
    y = np.zeros_like(x)
    if not lib.sdr_cpx_fft(_ptr(x), len(x), dir, _ptr(y)):
        raise ValueError("sdr_cpx_fft failed (N=%d dir=%d)" % (len(x), dir))
    return y


def cpx_mul(a, b, s):
    a = np.ascontiguousarray(a, dtype=np.complex64)
    b = np.ascontiguousarray(b, dtype=np.complex64)
    c = np.zeros_like(a)
    lib.sdr_cpx_mul(_ptr(a), _ptr(b), len(a), s, _ptr(c))
    return c


def search_code(code_fft, T, buff, ix, fs, fi, fds, P=None):
    """sdr_search_code (src/sdr_func.c:664-687): returns P (len_fds, N)"""
    code_fft = np.ascontiguousarray(code_fft, dtype=np.complex64)
    fds = np.ascontiguousarray(fds, dtype=np.float32)
    N = len(code_fft)
    if P is None:
        P = np.zeros((len(fds), N), dtype=np.float32)
    b = _buff(buff)
    lib.sdr_search_code(_ptr(code_fft), T, C.byref(b), ix, N, fs, fi, _fp(fds), len(fds), _fp(P))
    return P


def search_code_noalloc(code_fft, T, buff, ix, fs, fi, fds, P=None):
    """sdr_search_code_noalloc: sdr_search_code on the library's fixed device
    scratch (no device allocation).  Returns (ok, P); P is untouched when the
    call is rejected (ok == 0)."""
    code_fft = np.ascontiguousarray(code_fft, dtype=np.complex64)
    fds = np.ascontiguousarray(fds, dtype=np.float32)
    N = len(code_fft)
    if P is None:
        P = np.zeros((len(fds), N), dtype=np.float32)
    b = _buff(buff)
    ok = lib.sdr_search_code_noalloc(_ptr(code_fft), T, C.byref(b), ix, N, fs, fi, _fp(fds), len(fds), _fp(P))
    return ok, P


def corr_max(P, N, Nmax, M, T):
    P = np.ascontiguousarray(P, dtype=np.float32)
    ix = np.zeros(2, dtype=np.int32)
    cn0 = lib.sdr_corr_max(_fp(P), N, Nmax, M, T, ix.ctypes.data_as(c_int_p))
    return cn0, ix


def fine_dop(P, N, fds, ix):
    P = np.ascontiguousarray(P, dtype=np.float32)
    fds = np.ascontiguousarray(fds, dtype=np.float32)
    ix = np.ascontiguousarray(ix, dtype=np.int32)
    return lib.sdr_fine_dop(_fp(P), N, _fp(fds), len(fds), ix.ctypes.data_as(c_int_p))


def dop_bins(T, dop, max_dop):
    n = C.c_int(0)
    p = lib.sdr_dop_bins(T, dop, max_dop, C.byref(n))
    out = np.ctypeslib.as_array(p, shape=(n.value,)).copy()
    C.CDLL(None).free(C.cast(p, c_void_p))
    return out


def pass_shape(col, L, var):
    """(R0, R1, R2, G, threads) of workgroup shape `var` of the precision
    form's column (col True) or row passes of length L, else None."""
    t = (C.c_int * 5)()
    if not lib.sdr_search_pass_shape(1 if col else 0, L, var, C.cast(t, c_int_p)):
        return None
    return tuple(t)


def fft_pass_groups(L):
    """Sub-transforms per workgroup of the FP32 passes' compiled kernel of
    length L (0: the runtime-length kernel)."""
    return lib.sdr_fft_pass_groups(L)


def search_plan_bytes():
    """sdr_search_plan_bytes: device bytes held by the live search plans"""
    return lib.sdr_search_plan_bytes()


SEARCH_BATCHED = 0
SEARCH_PLAIN = 1
SEARCH_GRAPH = 2
SEARCH_FP16 = 4
SEARCH_BF16 = 8
SEARCH_TUNE = 16
FORM_FP32, FORM_PLAIN, FORM_FP16, FORM_BF16 = 0, 1, 2, 3


class SearchPlan:
    """Batched HBM-resident acquisition over signal groups (sdr_search_plan_*).

    buffs: list of uint8 cpx8 IF streams, or with dev_buffs a list of
    (device pointer, length) pairs read in place (sdr_search_plan_new_dev).
    sigs: list of dicts with keys buff, M, N, T, fs, fi, code_fft (n_prn, M)
    complex64, fds float32, n_blk, ix0, blk_step.
    """

    def __init__(self, buffs, sigs, flags=SEARCH_BATCHED, dev_buffs=False, raw_streams=None):
        """raw_streams: instead of buffs, packed raw rings in device memory:
        dicts raw (device pointer), N, fmt, rfch, rtoc, lut (uint8 table),
        ph0 (sdr_search_plan_new_raw)."""
        if raw_streams is not None:
            buffs = []
            dev_buffs = True
        if not dev_buffs:
            self._bufs = [_buff(b) for b in buffs]
            arr = (C.POINTER(sdr_buff_t) * len(buffs))(*[C.pointer(b) for b in self._bufs])
        self._keep = []
        sa = (sdr_search_sig_t * len(sigs))()
        for i, s in enumerate(sigs):
            cf = np.ascontiguousarray(s["code_fft"], dtype=np.complex64)
            fds = np.ascontiguousarray(s["fds"], dtype=np.float32)
            self._keep += [cf, fds]
            sa[i] = sdr_search_sig_t(s["buff"], s["M"], s["N"], s["T"], s["fs"], s["fi"],
                                     cf.shape[0], _ptr(cf), len(fds), _fp(fds), s["n_blk"],
                                     s["ix0"], s["blk_step"])
        self.sigs = sigs
        if raw_streams is not None:
            luts = [np.ascontiguousarray(r["lut"], np.uint8) for r in raw_streams]
            self._keep += luts
            ra = (sdr_raw_stream_t * len(raw_streams))()
            for i, r in enumerate(raw_streams):
                ra[i] = sdr_raw_stream_t(r["raw"], r["N"], r["fmt"], r["rfch"], int(r.get("rtoc", 0)),
                                         _ptr(luts[i]), int(r.get("ph0", 0)))
            self.p = lib.sdr_search_plan_new_raw(len(raw_streams), ra, len(sigs), sa, flags)
        elif dev_buffs:
            dp_ = (c_void_p * len(buffs))(*[b[0] for b in buffs])
            nb = (C.c_int * len(buffs))(*[b[1] for b in buffs])
            self.p = lib.sdr_search_plan_new_dev(len(buffs), dp_, nb, len(sigs), sa, flags)
        else:
            self.p = lib.sdr_search_plan_new(len(buffs), arr, len(sigs), sa, flags)
        if not self.p:
            raise ValueError("sdr_search_plan_new failed")

    def run(self):
        if not lib.sdr_search_plan_run(self.p):
            raise RuntimeError("sdr_search_plan_run failed")

    def result(self, sig):
        n = self.sigs[sig]["code_fft"].shape[0]
        cn0 = np.zeros(n, np.float32)
        ix = np.zeros((n, 2), np.int32)
        dop = np.zeros(n, np.float64)
        lib.sdr_search_plan_result(self.p, sig, _fp(cn0), ix.ctypes.data_as(c_int_p), _dp(dop))
        return cn0, ix, dop

    def power(self, sig, prn_idx):
        s = self.sigs[sig]
        P = np.zeros((len(s["fds"]), s["N"]), np.float32)
        lib.sdr_search_plan_power(self.p, sig, prn_idx, _fp(P))
        return P

    def form(self, sig):
        """FORM_FP32 / FORM_PLAIN / FORM_FP16 / FORM_BF16 of signal group sig."""
        return lib.sdr_search_plan_form(self.p, sig)

    def tiles(self, sig):
        """(M1, M2, column shape, row shape) of a precision-form group, else None."""
        t = (C.c_int * 4)()
        if not lib.sdr_search_plan_tiles(self.p, sig, C.cast(t, c_int_p)):
            return None
        return tuple(t)

    def timing(self, on=True):
        lib.sdr_search_plan_timing(self.p, 1 if on else 0)

    def kernel_ms(self, name):
        ms = C.c_double(0)
        cnt = C.c_longlong(0)
        el = C.c_double(0)
        if not lib.sdr_search_plan_kernel_ms(self.p, name.encode(), C.byref(ms), C.byref(cnt),
                                             C.byref(el)):
            return None
        by = C.c_double(0)
        lib.sdr_search_plan_kernel_bytes(self.p, name.encode(), C.byref(by))
        return ms.value, cnt.value, el.value, by.value

    def kernel_grid(self, name):
        """Average launch size in threads (the profiler's grid size)."""
        g = C.c_double(0)
        if not lib.sdr_search_plan_kernel_grid(self.p, name.encode(), C.byref(g)):
            return None
        return g.value

    def close(self):
        if self.p:
            lib.sdr_search_plan_free(self.p)
            self.p = None

    def __del__(self):
        self.close()


class TrkPlan:
    """Batched tracking correlators (sdr_trk_plan_*): open loop, or with
    loops (dicts len_code, costas, cn0 per channel) the closed-loop plan
    (sdr_trk_plan_new_loop)."""

    def __init__(self, buff, chans, pos, n_epoch, loops=None):
        self._buf = _buff(buff)
        self._keep = []
        ca = (sdr_trk_ch_t * len(chans))()
        for i, c in enumerate(chans):
            code = np.ascontiguousarray(c["code"], dtype=np.int8)
            self._keep.append(code)
            ca[i] = sdr_trk_ch_t(c["fs"], c["fi"], c["fc_sig"], c["T"], c["N"], c["fd0"],
                                 c["drift"], c["phi0"], c["coff0"], _ptr(code))
        self.pos = np.ascontiguousarray(pos, dtype=np.float64)
        self.n_ch, self.n_epoch, self.n_pos = len(chans), n_epoch, len(self.pos)
        if loops is not None:
            la = (sdr_trk_loop_t * len(chans))()
            for i, lp in enumerate(loops):
                sec = lp.get("sec_code")
                if sec is not None:
                    sec = np.ascontiguousarray(sec, np.int8)
                    self._keep.append(sec)
                la[i] = sdr_trk_loop_t(lp["len_code"], 1 if lp.get("costas", 1) else 0, lp["cn0"],
                                       0 if sec is None else len(sec), None if sec is None else _ptr(sec),
                                       float(lp.get("bump_step", 0.0)))
            self.p = lib.sdr_trk_plan_new_loop(C.byref(self._buf), len(chans), ca, la, self.n_pos,
                                               _dp(self.pos), n_epoch)
        else:
            self.p = lib.sdr_trk_plan_new(C.byref(self._buf), len(chans), ca, self.n_pos,
                                          _dp(self.pos), n_epoch)
        if not self.p:
            raise ValueError("sdr_trk_plan_new failed")

    def run(self):
        if not lib.sdr_trk_plan_run(self.p):
            raise RuntimeError("sdr_trk_plan_run failed")

    def result(self):
        corr = np.zeros((self.n_ch, self.n_epoch, self.n_pos), np.complex64)
        Cc = np.zeros((self.n_ch, self.n_epoch, 2), np.complex64)
        lib.sdr_trk_plan_result(self.p, _ptr(corr), _ptr(Cc))
        return corr, Cc

    def prompt(self):
        """Full-period prompt per (channel, epoch), the trk->P history."""
        P = np.zeros((self.n_ch, self.n_epoch), np.complex64)
        lib.sdr_trk_plan_prompt(self.p, _ptr(P))
        return P

    def state(self):
        """Closed-loop plans: (n_ch, n_epoch, 6) fd, coff, phi, cn0, lock, idle."""
        st = np.zeros((self.n_ch, self.n_epoch, 6), np.float64)
        if not lib.sdr_trk_plan_state(self.p, _ptr(st)):
            raise ValueError("sdr_trk_plan_state: not a closed-loop plan")
        return st

    def timing(self, on=True):
        lib.sdr_trk_plan_timing(self.p, 1 if on else 0)

    def kernel_ms(self, name):
        ms = C.c_double(0)
        cnt = C.c_longlong(0)
        el = C.c_double(0)
        if not lib.sdr_trk_plan_kernel_ms(self.p, name.encode(), C.byref(ms), C.byref(cnt),
                                          C.byref(el)):
            return None
        return ms.value, cnt.value, el.value

    def close(self):
        if self.p:
            lib.sdr_trk_plan_free(self.p)
            self.p = None

    def __del__(self):
        self.close()


# --- L6D / L6E FFT tracking and CSK (src/sdr_ch.c:627-664, 694-720) ----------
SDR_L6_KEEP_CORR = 1


def csk(corr, len_code, pos):
    """CSK of a batch of correlations (n_item, N): (syms uint8, taps complex64)."""
    corr = np.ascontiguousarray(np.atleast_2d(corr), np.complex64)
    pos = np.ascontiguousarray(pos, np.float64)
    n_item, N = corr.shape
    syms = np.zeros(n_item, np.uint8)
    taps = np.zeros((n_item, len(pos)), np.complex64)
    if not lib.sdr_csk(_ptr(corr), n_item, N, len_code, _dp(pos), len(pos), _ptr(syms), _ptr(taps)):
        raise ValueError("sdr_csk failed")
    return syms, taps


class L6Plan:
    """Batched L6D / L6E FFT tracking with CSK decoding (sdr_l6_plan_*).
    chans: dicts fs, fi, fc_sig, T, N, len_code, fd0, drift, phi0, coff0 and
    code_fft, the (SDR_N_CODES, N) conjugated code spectra.  buff: uint8 cpx8
    IF ring, or with dev_buff a (device pointer, length) pair read in place
    (sdr_l6_plan_new_dev)."""

    def __init__(self, buff, chans, pos, n_epoch, keep_corr=False, dev_buff=False, loops=None):
        """loops: per channel dicts len_code, costas, cn0 (bump_step) for the
        closed-loop plan (sdr_l6_plan_new_loop)."""
        self._buf = None if dev_buff else _buff(buff)
        self._keep = []
        ca = (sdr_l6_ch_t * len(chans))()
        for i, c in enumerate(chans):
            cf = np.ascontiguousarray(c["code_fft"], dtype=np.complex64)
            self._keep.append(cf)
            ca[i] = sdr_l6_ch_t(c["fs"], c["fi"], c["fc_sig"], c["T"], c["N"], c["len_code"],
                                c["fd0"], c["drift"], c["phi0"], c["coff0"], _ptr(cf))
        self.pos = np.ascontiguousarray(pos, dtype=np.float64)
        self.n_ch, self.n_epoch, self.n_pos = len(chans), n_epoch, len(self.pos)
        self.N = chans[0]["N"]
        flags = SDR_L6_KEEP_CORR if keep_corr else 0
        self.loop = loops is not None
        if loops is not None:
            if dev_buff:
                raise ValueError("closed-loop L6 plans take a host IF ring")
            la = (sdr_trk_loop_t * len(chans))()
            for i, lp in enumerate(loops):
                sec = lp.get("sec_code")
                if sec is not None:
                    sec = np.ascontiguousarray(sec, np.int8)
                    self._keep.append(sec)
                la[i] = sdr_trk_loop_t(lp["len_code"], 1 if lp.get("costas", 0) else 0, lp["cn0"],
                                       0 if sec is None else len(sec), None if sec is None else _ptr(sec),
                                       float(lp.get("bump_step", 0.0)))
            self.p = lib.sdr_l6_plan_new_loop(C.byref(self._buf), len(chans), ca, la, self.n_pos,
                                              _dp(self.pos), n_epoch, flags)
        elif dev_buff:
            self.p = lib.sdr_l6_plan_new_dev(buff[0], buff[1], len(chans), ca, self.n_pos,
                                             _dp(self.pos), n_epoch, flags)
        else:
            self.p = lib.sdr_l6_plan_new(C.byref(self._buf), len(chans), ca, self.n_pos,
                                         _dp(self.pos), n_epoch, flags)
        if not self.p:
            raise ValueError("sdr_l6_plan_new failed")

    def run(self):
        if not lib.sdr_l6_plan_run(self.p):
            raise RuntimeError("sdr_l6_plan_run failed")

    def result(self):
        syms = np.zeros((self.n_ch, self.n_epoch), np.uint8)
        taps = np.zeros((self.n_ch, self.n_epoch, self.n_pos), np.complex64)
        lib.sdr_l6_plan_result(self.p, _ptr(syms), _ptr(taps))
        return syms, taps

    def state(self):
        """Closed-loop plans: (n_ch, n_epoch, 6) fd, coff, phi, cn0, lock, idle."""
        st = np.zeros((self.n_ch, self.n_epoch, 6), np.float64)
        if not lib.sdr_l6_plan_state(self.p, _ptr(st)):
            raise ValueError("sdr_l6_plan_state: not a closed-loop plan")
        return st

    def corr(self, ch, epoch):
        out = np.zeros(self.N, np.complex64)
        if not lib.sdr_l6_plan_corr(self.p, ch, epoch, _ptr(out)):
            raise ValueError("sdr_l6_plan_corr failed (plan without keep_corr?)")
        return out

    def timing(self, on=True):
        lib.sdr_l6_plan_timing(self.p, 1 if on else 0)

    def kernel_ms(self, name):
        ms = C.c_double(0)
        cnt = C.c_longlong(0)
        el = C.c_double(0)
        if not lib.sdr_l6_plan_kernel_ms(self.p, name.encode(), C.byref(ms), C.byref(cnt),
                                         C.byref(el)):
            return None
        return ms.value, cnt.value, el.value

    def close(self):
        if self.p:
            lib.sdr_l6_plan_free(self.p)
            self.p = None

    def __del__(self):
        self.close()

# --- front end: raw data to cpx8 (write_buff, src/sdr_rcv.c:776-936) ----------
FMT_INT8, FMT_INT8X2, FMT_RAW8, FMT_RAW16, FMT_RAW16I, FMT_RAW32, FMT_CS8, FMT_CS16 = range(1, 9)
RAW_BPS = {FMT_INT8: 1, FMT_INT8X2: 2, FMT_CS8: 2, FMT_CS16: 4, FMT_RAW8: 1, FMT_RAW16: 2,
           FMT_RAW16I: 2, FMT_RAW32: 4}


def raw_gen_lut(fmt, rfch, IQ, bits, rtoc=False):
    """gen_LUTC / gen_LUTC_rtoc: uint8 cpx8 table (1024 entries with rtoc)."""
    out = np.zeros(1024 if rtoc else 256, np.uint8)
    lib.sdr_raw_gen_lut(fmt, rfch, IQ, bits, 1 if rtoc else 0, _ptr(out))
    return out


def raw_gen_lut8(fmt, off, scale):
    """gen_LUT8: int8 table (65536 entries for CS16)."""
    out = np.zeros(65536 if fmt == FMT_CS16 else 256, np.int8)
    lib.sdr_raw_gen_lut8(fmt, off, scale, _ptr(out))
    return out


class RawPlan:
    """sdr_raw_plan_*: write_buff for one format and its channel tables."""

    def __init__(self, fmt, luts, rtoc=None):
        self._luts = [None if l is None else np.ascontiguousarray(l) for l in luts]
        lp = (c_void_p * 8)(*[None if l is None else l.ctypes.data for l in self._luts[:8]])
        rt = (C.c_int * 8)(*((list(rtoc) if rtoc else []) + [0] * 8)[:8])
        self.p = lib.sdr_raw_plan_new(fmt, rt, lp)
        if not self.p:
            raise ValueError("sdr_raw_plan_new failed")
        self.fmt = fmt
        self.nch = lib.sdr_raw_plan_nch(self.p)

    def unpack(self, raw, N, ix=0):
        """Host arrays: (nch, N) uint8 cpx8 samples of block ix."""
        raw = np.ascontiguousarray(raw, np.uint8)
        assert raw.size >= N * RAW_BPS[self.fmt]
        out = np.zeros((self.nch, N), np.uint8)
        op = (c_void_p * 8)(*[out[k].ctypes.data for k in range(self.nch)])
        if not lib.sdr_raw_unpack(self.p, _ptr(raw), N, ix, op):
            raise RuntimeError("sdr_raw_unpack failed")
        return out

    def unpack_dev(self, raw_ptr, N, ix, out_ptrs):
        """Device pointers (16-byte aligned): raw bytes and one output per channel."""
        op = (c_void_p * 8)(*list(out_ptrs)[:self.nch])
        if not lib.sdr_raw_unpack_dev(self.p, c_void_p(raw_ptr), N, ix, op):
            raise RuntimeError("sdr_raw_unpack_dev failed")

    def kernel_ms(self):
        return lib.sdr_raw_plan_kernel_ms(self.p)

    def close(self):
        if self.p:
            lib.sdr_raw_plan_free(self.p)
            self.p = None

    def __del__(self):
        self.close()


def arch_combine(ins, w):
    """sdr_arch_combine for several beams: ins (nrfch, N) uint8 cpx8, w
    (narch, nrfch, 2) int16 Q8 weights -> (narch, N) uint8 cpx8."""
    ins = np.ascontiguousarray(ins, np.uint8)
    w = np.ascontiguousarray(w, np.int16)
    nrfch, N = ins.shape
    narch = w.shape[0]
    out = np.zeros((narch, N), np.uint8)
    ip_ = (c_void_p * 8)(*[ins[j].ctypes.data for j in range(nrfch)])
    op = (c_void_p * 8)(*[out[m].ctypes.data for m in range(narch)])
    if not lib.sdr_arch_combine_raw(nrfch, ip_, N, narch, _ptr(w), op):
        raise RuntimeError("sdr_arch_combine_raw failed")
    return out


def arch_combine_dev(in_ptrs, N, w, out_ptrs):
    """Device pointers (16-byte aligned); returns the kernel time (ms)."""
    w = np.ascontiguousarray(w, np.int16)
    ip_ = (c_void_p * 8)(*list(in_ptrs))
    op = (c_void_p * 8)(*list(out_ptrs))
    ms = lib.sdr_arch_combine_raw_dev(len(in_ptrs), ip_, N, len(out_ptrs), _ptr(w), op)
    if ms < 0:
        raise RuntimeError("sdr_arch_combine_raw_dev failed")
    return ms

