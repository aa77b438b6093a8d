"""The float step kernels compute 1/rho, u/a and sqrt(gamma*R*T) with short instruction
sequences (rcp_fast, div_fast, sqrt_fast in shll_device.h) that are meant to equal the
IEEE operators bit for bit.  These checks run the kernels' own functions on the GPU:
every normal reciprocal, every square root over 2^-100..2^101, and 2^29 random quotients.
The F64F kernels' faithful FP64 sequences (hardware seed and two Newton steps): rcp_nr2
equal to IEEE 1.0/b and rsq_nr2 within 2 ulp of 1.0/sqrt(x) on 2^28 random doubles each
(exponents -1000..1000)."""
import ctypes
import os

import pytest

pytestmark = pytest.mark.gpu

PROBE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "probes", "librounding_probe.so")


@pytest.fixture(scope="module")
def probe():
    if not os.path.exists(PROBE):
        pytest.fail("tests/probes/librounding_probe.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(PROBE)
    for name in ("probe_rcp_mismatches", "probe_sqrt_mismatches", "probe_div_mismatches", "probe_rcp64_mismatches",
                 "probe_rsq64_max_ulp"):
        getattr(lib, name).restype = ctypes.c_longlong
    for name in ("probe_div_mismatches", "probe_rcp64_mismatches", "probe_rsq64_max_ulp"):
        getattr(lib, name).argtypes = [ctypes.c_int, ctypes.c_ulonglong]
    return lib


def test_rcp_fast_is_correctly_rounded_for_every_normal_float(probe):
    print("[gpu-case] rcp_fast exhaustive", flush=True)
    assert probe.probe_rcp_mismatches() == 0


def test_sqrt_fast_is_correctly_rounded_over_its_range(probe):
    print("[gpu-case] sqrt_fast exhaustive", flush=True)
    assert probe.probe_sqrt_mismatches() == 0


def test_div_fast_equals_ieee_division_on_random_pairs(probe):
    print("[gpu-case] div_fast random pairs", flush=True)
    assert probe.probe_div_mismatches(64, 20261203) == 0


def test_fp64_faithful_reciprocal_and_reciprocal_square_root(probe):
    print("[gpu-case] rcp_nr2 / rsq_nr2 random doubles", flush=True)
    assert probe.probe_rcp64_mismatches(32, 20261204) == 0
    worst = probe.probe_rsq64_max_ulp(32, 20261205)
    print("rsq_nr2: at most %d ulp from 1.0/sqrt(x)" % worst)
    assert 0 <= worst <= 2
