"""Numerical behaviour of the drop-in `prc` executable on the GPU: the invariant suite of the
reference's GPU tests (tests/test_solver.py:141-361), run through the CLI and raw files on
the same seeded 16x14x12 anisotropic phantom, with the reference's tolerances (and 1e-6 for
the identity and conservation where that is tighter)."""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

import prc_mlem_amd
from conftest import announce
from prc_mlem_amd import phantoms
from test_oracle import water_kernel

pytestmark = pytest.mark.gpu
EXE = Path(os.environ.get("PRC_EXE", prc_mlem_amd.PRC_EXE)).resolve()
SHAPE = (12, 14, 16)


@pytest.fixture(scope="module")
def ws(tmp_path_factory):
    d = tmp_path_factory.mktemp("prc_cli")
    rho, x, y, _ = phantoms.test_phantom()
    rho.tofile(d / "density.raw")
    return {"dir": d, "rho": rho, "x": x, "y": y, "support": rho >= 0.05, "n": [0]}


def run_prc(ws, image, mode="correct", *options, isotope="68Ga"):
    ws["n"][0] += 1
    stem = ws["dir"] / str(ws["n"][0])
    src, out = f"{stem}_in.raw", f"{stem}_out.raw"
    image.astype(np.float32).tofile(src)
    cmd = [str(EXE), isotope, "16", "14", "12", "1.1", "1.6", "2.3", "--input", src, "--density",
           str(ws["dir"] / "density.raw"), "--output", out, "--mode", mode, "--rays", "32", *options]
    announce(" ".join(cmd[1:2] + [mode] + list(options)))
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=90)
    assert r.returncode == 0, r.stdout + r.stderr
    res = np.fromfile(out, dtype=np.float32)
    assert res.size == image.size and np.isfinite(res).all()
    return res.reshape(SHAPE)


def test_projector_adjoint_and_normalization(ws):
    x, y, sup = ws["x"], ws["y"], ws["support"]
    ax, aty = run_prc(ws, x, "forward"), run_prc(ws, y, "adjoint")
    assert abs(np.sum(ax.astype(np.float64) * y) / np.sum(x.astype(np.float64) * aty) - 1) <= 1e-6
    assert abs(ax.sum(dtype=np.float64) / x.sum(dtype=np.float64) - 1) <= 1e-6
    np.testing.assert_array_equal(ax[~sup], 0)
    one = run_prc(ws, np.ones(SHAPE, np.float32), "adjoint")
    np.testing.assert_allclose(one[sup], 1, rtol=3e-6, atol=3e-6)
    np.testing.assert_array_equal(one[~sup], 0)


def test_cached_recomputed_and_exact_corrections_agree(ws):
    opts = ("--iterations", "3", "--beta", "0.1")
    cached = run_prc(ws, ws["x"], "correct", *opts)
    uncached = run_prc(ws, ws["x"], "correct", *opts, "--no-cache-near")
    np.testing.assert_allclose(cached, uncached, rtol=3e-5, atol=3e-6)
    exact = run_prc(ws, ws["x"], "correct", *opts, "--no-cache-near", "--exact-range")
    np.testing.assert_allclose(cached, exact, rtol=3e-5, atol=3e-6)


@pytest.mark.parametrize("size", [4.0, 2.0])
def test_water_kernel_matches_exact_voxel_average(ws, oracle_mod, size):
    # tests/test_solver.py:216-243: within 1% of the blur; mirrored rays leave no shift.
    half = int(np.ceil(10 / size)) + 1
    n, c = 2 * half + 5, half + 2
    d = ws["dir"]
    delta = np.zeros((n, n, n), np.float32)
    delta[c, c, c] = 1
    delta.tofile(d / f"w{size}_in.raw")
    np.ones((n, n, n), np.float32).tofile(d / f"w{size}_rho.raw")
    r = subprocess.run([str(EXE), "68Ga", str(n), str(n), str(n), str(size), str(size), str(size), "--mode",
                        "forward", "--input", str(d / f"w{size}_in.raw"), "--density", str(d / f"w{size}_rho.raw"),
                        "--output", str(d / f"w{size}_out.raw")], capture_output=True, text=True, timeout=90)
    assert r.returncode == 0, r.stdout + r.stderr
    P = np.fromfile(d / f"w{size}_out.raw", np.float32).reshape(n, n, n)
    P = P[c - half:c + half + 1, c - half:c + half + 1, c - half:c + half + 1].astype(np.float64)
    exact = water_kernel(oracle_mod.model("68Ga").astype(np.float64), (size,) * 3, half)
    blur = exact.copy()
    blur[half, half, half] -= 1
    assert np.linalg.norm(P - exact) / np.linalg.norm(blur) < 0.01
    for axis in np.indices(P.shape) - half:
        assert abs((P * axis).sum()) < 1e-6


PRECISIONS = [(), ("--precision", "mixed")]


@pytest.mark.parametrize("precision", PRECISIONS)
def test_sparse_observations_do_not_lock_supported_zeros(ws, precision):
    sparse = ws["x"].copy()
    sparse[np.random.default_rng(21).random(SHAPE) < 0.5] = 0
    out = run_prc(ws, sparse, "correct", "--iterations", "5", "--beta", "0.1", *precision)
    if precision:  # the mixed form against FP32 on the same data
        ref = run_prc(ws, sparse, "correct", "--iterations", "5", "--beta", "0.1")
        assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-3
    sup = ws["support"]
    assert np.mean(out[(sparse == 0) & sup] > 0) > 0.95
    assert out.min() >= 0
    np.testing.assert_array_equal(out[~sup], 0)
    assert abs(out.sum(dtype=np.float64) / sparse.sum(dtype=np.float64) - 1) <= 3e-6


def test_zero_iterations_preserves_supported_input_exactly(ws):
    src = ws["x"].copy()
    src[5, 4, 5] = 0
    np.testing.assert_array_equal(run_prc(ws, src, "correct", "--iterations", "0"), src)


@pytest.mark.skipif(not Path("/dev/full").exists(), reason="/dev/full is unavailable")
def test_buffered_output_write_failure_is_reported(ws):
    d = ws["dir"]
    np.ones(8, np.float32).tofile(d / "full_in.raw")
    np.ones(8, np.float32).tofile(d / "full_rho.raw")
    r = subprocess.run([str(EXE), "68Ga", "2", "2", "2", "1", "1", "1", "--mode", "forward", "--input",
                        str(d / "full_in.raw"), "--density", str(d / "full_rho.raw"), "--output", "/dev/full",
                        "--rays", "32"], capture_output=True, text=True, timeout=90)
    assert r.returncode != 0 and "Cannot write /dev/full" in r.stderr


def test_overflowing_update_fails_without_writing_output(ws):
    d = ws["dir"]
    np.full(8, 1e20, np.float32).tofile(d / "x_in.raw")
    np.ones(8, np.float32).tofile(d / "x_rho.raw")
    out = d / "x_out.raw"
    r = subprocess.run([str(EXE), "68Ga", "2", "2", "2", "1", "1", "1", "--input", str(d / "x_in.raw"), "--density",
                        str(d / "x_rho.raw"), "--output", str(out), "--iterations", "1", "--no-accel", "--rays", "32"],
                       capture_output=True, text=True, timeout=90)
    assert r.returncode != 0 and "Non-finite output" in r.stderr
    assert not out.exists()


def test_activity_unit_scaling(ws):
    a = run_prc(ws, ws["x"], "correct", "--iterations", "4", "--beta", "0.1")
    b = run_prc(ws, ws["x"] * 1000, "correct", "--iterations", "4", "--beta", "0.1") / 1000
    np.testing.assert_allclose(b, a, rtol=5e-5, atol=5e-6)


@pytest.mark.parametrize("iterations", ["1", "2"])
def test_acceleration_warmup_matches_first_two_plain_steps(ws, iterations):
    opts = ("--iterations", iterations, "--beta", "0.2", "--gamma", "20")
    np.testing.assert_allclose(run_prc(ws, ws["x"], "correct", *opts),
                               run_prc(ws, ws["x"], "correct", "--no-accel", *opts), rtol=3e-5, atol=3e-6)


@pytest.mark.parametrize("accel", [(), ("--no-accel",)])
def test_unregularized_rl_conserves_counts(ws, accel):
    out = run_prc(ws, ws["x"], "correct", "--iterations", "25", "--beta", "0", *accel)
    assert out.min() >= 0
    np.testing.assert_array_equal(out[~ws["support"]], 0)
    assert abs(out.sum(dtype=np.float64) / ws["x"].sum(dtype=np.float64) - 1) <= 1e-6


@pytest.mark.parametrize("precision", PRECISIONS)
def test_long_sparse_cold_reconstruction_stays_finite_and_preserves_counts(ws, precision):
    sup = ws["support"]
    truth = sup.astype(np.float32)
    cold = np.zeros(SHAPE, bool)
    cold[4:8, 3:7, 3:7] = True
    cold &= sup
    truth[cold] = 0
    blurred = run_prc(ws, truth, "forward")
    sparse = (np.random.default_rng(81).poisson(2 * blurred) / 2).astype(np.float32)
    missed = (sparse == 0) & sup & ~cold
    assert missed.sum() > 20
    for accel in ((), ("--no-accel",)):
        out = run_prc(ws, sparse, "correct", "--iterations", "30", "--beta", "0.2", "--gamma", "20", *accel,
                      *precision)
        if precision:
            ref = run_prc(ws, sparse, "correct", "--iterations", "30", "--beta", "0.2", "--gamma", "20", *accel)
            assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-3
        assert out.min() >= 0
        np.testing.assert_array_equal(out[~sup], 0)
        assert np.mean(out[missed] > 0) > 0.95
        assert out[cold].mean() < out[sup & ~cold].mean()
        assert abs(out.sum(dtype=np.float64) / sparse.sum(dtype=np.float64) - 1) <= 5e-6


def test_acceleration_approaches_a_long_plain_reference_sooner(ws):
    blurred = run_prc(ws, ws["x"], "forward")
    opts = ("--beta", "0.2", "--gamma", "20")
    plain = run_prc(ws, blurred, "correct", "--iterations", "400", "--no-accel", *opts)
    accel = run_prc(ws, blurred, "correct", "--iterations", "100", *opts)
    short = run_prc(ws, blurred, "correct", "--iterations", "100", "--no-accel", *opts)

    def dist(img):
        return np.linalg.norm(img - plain) / np.linalg.norm(plain)

    assert dist(accel) < 1e-3
    assert dist(accel) < 0.5 * dist(short)
    assert abs(accel.sum(dtype=np.float64) / blurred.sum(dtype=np.float64) - 1) <= 3e-6


def test_checkpoint_and_reprojection_names_in_dotted_directory(ws):
    d = ws["dir"] / "dir.with.dots"
    d.mkdir()
    src = ws["dir"] / "dotted_in.raw"
    ws["x"].tofile(src)
    r = subprocess.run([str(EXE), "68Ga", "16", "14", "12", "1.1", "1.6", "2.3", "--input", str(src), "--density",
                        str(ws["dir"] / "density.raw"), "--output", str(d / "output"), "--rays", "32",
                        "--iterations", "3", "--save-every", "1", "--reproject"], capture_output=True, text=True,
                       timeout=90)
    assert r.returncode == 0, r.stdout + r.stderr
    for name in ["output", "output_it001", "output_it002", "output_REPROJ"]:
        assert (d / name).stat().st_size == ws["x"].nbytes


def test_mixed_precision_cli_close_to_fp32(ws):
    a = run_prc(ws, ws["x"], "correct", "--iterations", "6", "--beta", "0.03")
    b = run_prc(ws, ws["x"], "correct", "--iterations", "6", "--beta", "0.03", "--precision", "mixed")
    assert np.linalg.norm(b - a) / np.linalg.norm(a) < 1e-3
    assert abs(b.sum(dtype=np.float64) / a.sum(dtype=np.float64) - 1) < 1e-6
