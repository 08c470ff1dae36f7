"""The C++ host layer (sgn-hyperbolic-gpu_amd/cpp/sgnx/sgn2d.hpp) and the
constant-divisor quotient, through the test programs in tests/cpp/ and the
reference drop-in program oracle/_ref/dropin_ref."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
DROPIN_REF = os.path.join(ROOT, "oracle", "_ref", "dropin_ref")


def _binary(name):
    path = os.path.join(CPP, name)
    if not os.path.exists(path):
        subprocess.run(["make", name], cwd=CPP, check=True, capture_output=True)
    return path


def _no_gpu_env():
    return dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")


def test_constant_divisor_quotient_is_ieee_division():
    """div_const == a / b bit for bit on 5.4M inputs per divisor (random bit
    patterns, lift-sized values, neighbours of exact multiples and of halfway
    quotients, zeros, infinities, NaN, subnormals, range limits), for the
    Jacobians of every workload here and a few others."""
    divisors = ["0.25", "0.046875", "0.625", "0.6666666666666666", "0.65", "0.35", "0.0703125",
                "0.5", "3.0", "1.7e-5", "47.3", "0.1", "0.05859375", "0.4166666666666667"]
    out = subprocess.run([_binary("fastdiv_check"), "1000000"] + divisors, capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert " 0 mismatches" in out.stdout


@pytest.mark.parametrize("prog", ["dropin_sgnx", "adjoint_sgnx", "dropin_ref"])
def test_cpp_layer_without_device_reports_error(prog):
    path = DROPIN_REF if prog == "dropin_ref" else _binary(prog)
    if not os.path.exists(path):
        pytest.skip("reference drop-in program not built (no reference checkout at build time)")
    out = subprocess.run([path], env=_no_gpu_env(), capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "NO_DEVICE" in out.stdout


@pytest.mark.gpu
def test_reference_class_drops_into_sgnx():
    """The reference's sgn::SGN2D handed to sgnx::SGN2DHip as to sgn::SGN2DCuda:
    RHS bitwise against the reference's own rhs(), adaptive run with the same
    step counts within 1e-12."""
    if not os.path.exists(DROPIN_REF):
        pytest.skip("reference drop-in program not built (no reference checkout at build time)")
    out = subprocess.run([DROPIN_REF], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0 and "DROPIN_OK" in out.stdout, out.stdout + out.stderr


@pytest.mark.gpu
def test_sgnx_cpp_class_matches_oracle():
    out = subprocess.run([_binary("dropin_sgnx")], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0 and "DROPIN_OK" in out.stdout, out.stdout + out.stderr


@pytest.mark.gpu
def test_sgnx_cpp_adjoint_matches_differences():
    """sgnx::SGN2DHip's gauges, tape and reverse sweep: the directional
    derivative of a taped run against fourth-order differences of its replay."""
    out = subprocess.run([_binary("adjoint_sgnx")], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0 and "ADJOINT_OK" in out.stdout, out.stdout + out.stderr
