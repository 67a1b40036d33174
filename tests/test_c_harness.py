"""The C caller (tests/c/harness.c): the reference applications' loops over
the drop-in entry points, compiled against include/pocket_sdr_amd.h, checked
against the oracle, including eight host threads calling at once."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "tests", "c", "harness")


def test_c_harness():
    assert os.path.exists(HARNESS), "tests/c/harness is not built (make -C tests/c)"
    env = dict(os.environ)
    ref = os.path.join(ROOT, "oracle", "_ref", "libpocketfft_ref.so")
    if os.path.exists(ref):
        env["ORACLE_REF_FFT"] = ref
    r = subprocess.run([HARNESS, "8", "50"], capture_output=True, text=True, timeout=240, env=env)
    print(r.stdout)
    assert r.returncode == 0 and "harness ok" in r.stdout, r.stdout + r.stderr
