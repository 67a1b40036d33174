"""The C-ABI library loads and exports every function include/pocket_sdr_amd.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "pocket_sdr_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = set(re.findall(r"\b(sdr_\w+)\s*\(", src))
    return sorted(names)


def test_header_declares_reference_hot_path():
    names = declared_functions()
    for n in ["sdr_func_init", "sdr_mix_carr", "sdr_corr_std", "sdr_corr_std_cpx_code",
              "sdr_corr_std_cpx", "sdr_corr_fft", "sdr_corr_fft_cpx", "sdr_search_code",
              "sdr_cpx_fft", "sdr_cpx_mul", "sdr_cpx_malloc", "sdr_cpx_free", "sdr_corr_max",
              "sdr_fine_dop", "sdr_dop_bins", "sdr_buff_new", "sdr_buff_free"]:
        assert n in names


def test_library_exports_all_declared():
    import pocketsdr_amd
    lib = ctypes.CDLL(pocketsdr_amd.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_host_helpers_without_gpu():
    """Host-only helpers (no device work) behave as the reference's known
    answers (test/src/test_sdr_func.c:255-280)."""
    import numpy as np
    from pocketsdr_amd import sdr_func as s
    P = np.array([1, 1, 1, 1, 1, 1, 4, 1, 1, 1, 1, 1], np.float32)
    cn0, ix = s.corr_max(P, 4, 4, 3, 0.001)
    assert cn0 > 0 and ix.tolist() == [1, 2]
    fds = np.array([-500, 0, 500], np.float32)
    assert abs(s.fine_dop(P, 4, fds, ix)) < 1e-6
    assert abs(s.fine_dop(P, 4, fds, np.array([0, 0])) + 500) < 1e-6
    assert s.dop_bins(0.001, 0.0, 1000.0).tolist() == [-1000, -500, 0, 500, 1000]
