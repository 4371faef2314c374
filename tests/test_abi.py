"""The C ABI: the library loads, exports every entry point include/bcmhip.h
declares, and its device entry points fail loudly (no CPU fallback) when no
GPU is visible."""

import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bcmhip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(bcm_\w+)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if not n.startswith("bcm_cb_")))


def test_header_declarations_exported(B):
    lib = B.lib()
    names = declared_functions()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python mirror binds exactly these
    assert set(names) == set(B.SIGNATURES)


def test_dynamic_symbols_match_header(B):
    out = subprocess.run(["nm", "-D", "--defined-only", B.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (bcm_\w+)", out))
    assert set(declared_functions()) <= exported


def test_params_defaults_mirror_cslvparam(B):
    p = B.default_params()
    # CSlvParam::CSlvParam (slvparam.h:121-214)
    assert (p.tau1, p.tau2, p.theta, p.sclmin, p.pivmin) == (1e-2, 1e-4, 1e-1, 1e-8, 1e-8)
    assert (p.ordtype, p.ittype, p.sttype, p.niter, p.eps) == (1, 2, 2, 1000, 1e-9)


def test_errors_are_status_codes(B):
    with pytest.raises(B.BcmError) as e:
        B.Matrix.from_arrays([0, 2, 4], [0, 1, 2], [2, 0], np.ones(8), symmetric=False)  # column out of range
    assert e.value.code in (B.E_ARG,)
    assert B.lib().bcm_matrix_destroy(None) == B.OK


@pytest.mark.skipif(os.environ.get("HIP_VISIBLE_DEVICES", None) is None and os.path.exists("/dev/kfd"),
                    reason="a GPU may be visible here")
def test_device_calls_fail_without_gpu(B):
    assert B.device_count() == 0
    with pytest.raises(B.BcmError) as e:
        B.Context()
    assert e.value.code == B.E_NODEVICE
    with pytest.raises(B.BcmError) as e:
        B.DeviceArray(16)
    assert e.value.code == B.E_NODEVICE
