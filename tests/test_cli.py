"""The drop-in CLI rejects bad input before touching any GPU (tests/test_solver.py:25-125
of the reference): every case runs with all devices hidden and must fail with the
reference's message, print no HIP error and write no output."""
import os
import struct
import subprocess
from pathlib import Path

import pytest

import prc_mlem_amd

EXE = Path(os.environ.get("PRC_EXE", prc_mlem_amd.PRC_EXE)).resolve()
HIDDEN = {"HIP_VISIBLE_DEVICES": "", "ROCR_VISIBLE_DEVICES": "", "CUDA_VISIBLE_DEVICES": ""}

pytestmark = pytest.mark.skipif(not EXE.exists(), reason="prc executable not built")


def write_raw(path, values):
    path.write_bytes(struct.pack(f"<{len(values)}f", *values))


@pytest.fixture
def files(tmp_path):
    d, i, o = tmp_path / "density.raw", tmp_path / "input.raw", tmp_path / "output.raw"
    write_raw(d, [1.0] * 8)
    write_raw(i, [0.0] * 8)
    return d, i, o


def run_prc(files, *options, geometry=None):
    d, i, o = files
    args = geometry or ["68Ga", "2", "2", "2", "1", "1", "1"]
    return subprocess.run([str(EXE), *args, "--density", str(d), "--input", str(i), "--output", str(o), *options],
                          env={**os.environ, **HIDDEN}, text=True, capture_output=True, timeout=30)


def assert_rejected(files, result, message):
    assert result.returncode != 0, result.stdout
    assert message in result.stderr
    assert "HIP error" not in result.stderr and "prc:" not in result.stderr
    assert not files[2].exists()


@pytest.mark.parametrize("option", ["--help", "-h"])
def test_help_succeeds_without_gpu(option):
    r = subprocess.run([str(EXE), option], env={**os.environ, **HIDDEN}, text=True, capture_output=True, timeout=30)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Usage:" in r.stdout
    assert r.stderr == ""


@pytest.mark.parametrize("option,value", [("--beta", "garbage"), ("--beta", "nan"), ("--gamma", "inf"),
                                          ("--rays", "128x"), ("--rays", "2147483648"), ("--iterations", "2.5"),
                                          ("--beta", "1e999")])
def test_invalid_numeric_options(files, option, value):
    assert_rejected(files, run_prc(files, option, value), "Invalid")


@pytest.mark.parametrize("option,value", [("--support-rho", "-1"), ("--air-cap-mm", "0"), ("--save-every", "-1"),
                                          ("--iterations", "2147483647"), ("--rays", "2147483647"),
                                          ("--rays", "1"), ("--rays", "3"), ("--rays", "129"),
                                          ("--precision", "fp8"), ("--mode", "blur")])
def test_invalid_bounds(files, option, value):
    assert_rejected(files, run_prc(files, option, value), "Invalid")


@pytest.mark.parametrize("geometry", [["68Ga", "2x", "2", "2", "1", "1", "1"], ["68Ga", "2", "2", "2", "nan", "1", "1"],
                                      ["68Ga", "0", "2", "2", "1", "1", "1"]])
def test_invalid_geometry(files, geometry):
    assert_rejected(files, run_prc(files, geometry=geometry), "Invalid")


def test_index_overflow(files):
    assert_rejected(files, run_prc(files, geometry=["68Ga", "65536", "65536", "1", "1", "1", "1"]), "INT_MAX")


def test_missing_value_and_unknown_option(files):
    assert_rejected(files, run_prc(files, "--beta"), "Missing value for --beta")
    assert_rejected(files, run_prc(files, "--frobnicate"), "Unknown option --frobnicate")


def test_unknown_isotope_prints_usage(files):
    r = run_prc(files, geometry=["99Xx", "2", "2", "2", "1", "1", "1"])
    assert r.returncode != 0 and "Usage:" in r.stderr and not files[2].exists()


def test_too_few_arguments():
    r = subprocess.run([str(EXE), "68Ga", "2"], env={**os.environ, **HIDDEN}, text=True, capture_output=True,
                       timeout=30)
    assert r.returncode != 0 and "Usage:" in r.stderr


@pytest.mark.parametrize("size", [7, 9])
def test_raw_size_must_match_exactly(files, size):
    write_raw(files[1], [0.0] * size)
    assert_rejected(files, run_prc(files), "Expected exactly")


def test_raw_size_in_bytes(files):
    files[1].write_bytes(b"\x00" * 33)
    assert_rejected(files, run_prc(files), "Expected exactly")


@pytest.mark.parametrize("which", [0, 1])
@pytest.mark.parametrize("value", [float("nan"), float("inf"), float("-inf")])
def test_nonfinite_raw_values(files, which, value):
    write_raw(files[which], [value] + [1.0] * 7)
    assert_rejected(files, run_prc(files), "Non-finite")


@pytest.mark.parametrize("mode", ["correct", "forward", "adjoint"])
def test_zero_activity_is_rejected_without_gpu(files, mode):
    assert_rejected(files, run_prc(files, "--mode", mode), "No activity")


def test_empty_support_is_rejected_without_gpu(files):
    write_raw(files[0], [0.0] * 8)
    write_raw(files[1], [3.0] * 8)
    assert_rejected(files, run_prc(files), "No activity")


def test_missing_input_file(files, tmp_path):
    r = subprocess.run([str(EXE), "68Ga", "2", "2", "2", "1", "1", "1", "--density", str(files[0]), "--input",
                        str(tmp_path / "absent.raw"), "--output", str(files[2])],
                       env={**os.environ, **HIDDEN}, text=True, capture_output=True, timeout=30)
    assert_rejected(files, r, "Expected exactly")
