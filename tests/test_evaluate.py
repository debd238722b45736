"""The image-quality evaluator (prc_mlem_amd/evaluate.py), CPU only. The metric contract follows
the reference's metric tests (tests/test_metrics.py:24-173); the 3D SSIM, which the reference takes
from scikit-image (absent here), is pinned by a closed form for constant volumes, by the identity,
and by an independent brute-force evaluation of its definition (explicit 11-tap Gaussian weights,
mirrored edges, direct separable sums)."""
import json
import subprocess
import sys

import numpy as np
import pytest
from scipy import ndimage as ndi

from prc_mlem_amd import evaluate as ev


def phantom():
    truth = np.zeros((35, 35, 35), dtype=np.float32)
    truth[4:31, 4:31, 4:31] = 1
    truth[13:22, 13:22, 13:22] = 4
    return truth


def brute_ssim_map(a, b, data_range, sigma=1.5, radius=5, k1=0.01, k2=0.03):
    x = np.arange(-radius, radius + 1, dtype=np.float64)
    w = np.exp(-0.5 * (x / sigma) ** 2)
    w /= w.sum()

    def blur(v):
        for axis in range(3):
            pad = [(0, 0)] * 3
            pad[axis] = (radius, radius)
            p = np.pad(v, pad, mode="symmetric")  # d c b a | a b c d | d c b a
            acc = np.zeros_like(v)
            for i, wi in enumerate(w):
                sl = [slice(None)] * 3
                sl[axis] = slice(i, i + v.shape[axis])
                acc += wi * p[tuple(sl)]
            v = acc
        return v

    a, b = a.astype(np.float64), b.astype(np.float64)
    ma, mb = blur(a), blur(b)
    va, vb, cab = blur(a * a) - ma ** 2, blur(b * b) - mb ** 2, blur(a * b) - ma * mb
    c1, c2 = (k1 * data_range) ** 2, (k2 * data_range) ** 2
    return ((2 * ma * mb + c1) * (2 * cab + c2)) / ((ma ** 2 + mb ** 2 + c1) * (va + vb + c2))


def test_ssim_matches_brute_force_definition():
    rng = np.random.default_rng(11)
    a = rng.uniform(0, 3, (13, 14, 15))
    b = a + rng.normal(0, 0.3, a.shape)
    np.testing.assert_allclose(ev.ssim_map(a, b, 3.0), brute_ssim_map(a, b, 3.0), rtol=0, atol=1e-12)


def test_ssim_of_constant_volumes_is_the_luminance_term():
    c1v, c2v, r = 2.0, 1.5, 4.0
    s = ev.ssim_map(np.full((12, 12, 12), c1v), np.full((12, 12, 12), c2v), r)
    k1 = (0.01 * r) ** 2
    np.testing.assert_allclose(s, (2 * c1v * c2v + k1) / (c1v ** 2 + c2v ** 2 + k1), rtol=1e-13)


def test_identity_and_strict_json():
    truth = phantom()
    m = ev.Evaluator(truth)(truth)
    for key in ("nrmse", "nrmse_tissue", "mse", "mae_tissue", "edge_nrmse", "ring", "overshoot", "undershoot",
                "bg_bias", "bg_cov", "halo", "count_bias"):
        assert m[key] == 0, key
    for key in ("ssim", "ssim_tissue", "ssim_edge", "rc_hot", "rc_min"):
        assert m[key] == pytest.approx(1, abs=1e-12), key
    assert m["psnr"] is None
    json.dumps(m, allow_nan=False)


def test_global_metrics_follow_their_definitions():
    truth = phantom()
    cand = ndi.gaussian_filter(truth.astype(np.float64), sigma=0.7)
    e = ev.Evaluator(truth)
    m = e(cand)
    t = truth.astype(np.float64)
    assert m["nrmse"] == pytest.approx(np.linalg.norm(cand - t) / np.linalg.norm(t), abs=1e-12)
    assert m["count_bias"] == pytest.approx(cand.sum() / t.sum() - 1, abs=1e-12)
    smap = brute_ssim_map(t[e.roi], cand[e.roi], 4.0)
    inner = (slice(5, -5),) * 3
    assert m["ssim"] == pytest.approx(smap[inner].mean(), abs=1e-12)
    assert m["ssim_tissue"] == pytest.approx(smap[inner][e.body[e.roi][inner]].mean(), abs=1e-12)


def test_blur_and_local_range_violations_are_distinct():
    truth = phantom()
    e = ev.Evaluator(truth)
    blur = e(ndi.uniform_filter(truth, size=3))
    assert blur["edge_nrmse"] > 0.05 and blur["ring"] < 1e-7
    ring = truth.copy()
    ring[13:15, 13:22, 13:22] += 0.5
    ring[11:13, 13:22, 13:22] = -0.2
    m = e(ring)
    assert m["overshoot"] > 0 and m["undershoot"] > 0
    assert m["ring"] ** 2 == pytest.approx(m["overshoot"] ** 2 + m["undershoot"] ** 2)


def test_scale_invariant_scores():
    truth = phantom()
    cand = ndi.gaussian_filter(truth, 0.8)
    a = ev.Evaluator(truth)(cand)
    for scale in (1e-4, 20, 1e4):
        b = ev.Evaluator(truth.astype(np.float64) * scale)(cand.astype(np.float64) * scale)
        for key in ("nrmse", "ssim", "ssim_tissue", "ssim_edge", "ring", "halo", "edge_nrmse", "bg_bias", "bg_cov",
                    "rc_hot", "rc_min"):
            assert a[key] == pytest.approx(b[key], abs=1e-9), key


def test_missing_metrics_are_explicit():
    const = np.ones((15, 15, 15), dtype=np.float32)
    m = ev.Evaluator(const)(const)
    for key in ("rc_hot", "rc_min", "ring", "halo", "edge_nrmse", "ssim_edge"):
        assert m[key] is None, key
    assert m["ssim"] == 1
    assert "—" in ev.format_table([("constant", m)])
    small = np.ones((5, 15, 15), dtype=np.float32)
    m = ev.Evaluator(small)(small)
    assert m["ssim"] is None and m["nrmse"] == 0


def test_masks_cold_regions_and_hot_lesions():
    truth = phantom()
    body = truth > 0
    truth[8:12, 8:12, 8:12] = 0
    cand = truth.copy()
    cand[8:12, 8:12, 8:12] = 1
    assert ev.Evaluator(truth)(cand)["nrmse_tissue"] == 0
    assert ev.Evaluator(truth, mask=body)(cand)["nrmse_tissue"] > 0
    truth = phantom()
    lesions = truth == 4
    truth[5:9, 5:9, 5:9] = 2
    cand = truth.copy()
    cand[5:9, 5:9, 5:9] = 1
    threshold = ev.Evaluator(truth, background=1)
    explicit = ev.Evaluator(truth, background=1, hot_mask=lesions)
    assert threshold(cand)["rc_hot"] < 1
    assert explicit(cand)["rc_hot"] == 1 and explicit(cand)["rc_min"] == 1
    assert explicit.metric_settings["hot_region_count"] == 1 and threshold.metric_settings["hot_region_count"] == 2
    bad = lesions.copy()
    bad[0, 0, 0] = True
    with pytest.raises(ValueError, match="inside tissue"):
        ev.Evaluator(truth, hot_mask=bad)
    with pytest.raises(ValueError, match="truth shape"):
        ev.Evaluator(truth, hot_mask=lesions[:, :, :-1])


def test_rejects_invalid_inputs_and_sizes(tmp_path):
    truth = phantom()
    e = ev.Evaluator(truth)
    bad = truth.copy()
    bad[0, 0, 0] = np.nan
    for value in (bad, -truth, np.zeros_like(truth), np.ones((3, 3))):
        with pytest.raises(ValueError):
            ev.Evaluator(value)
    for kwargs in ({"data_range": 0}, {"data_range": np.nan}, {"background": -1}, {"mask": np.zeros_like(truth)},
                   {"mask": np.ones((3, 3, 3))}):
        with pytest.raises(ValueError):
            ev.Evaluator(truth, **kwargs)
    for value in (bad, truth[:, :, :-1]):
        with pytest.raises(ValueError):
            e(value)
    path = tmp_path / "bad.raw"
    np.zeros(7, dtype=np.float32).tofile(path)
    with pytest.raises(ValueError, match="expected 32 bytes, got 28"):
        ev.load(path, (2, 2, 2))


def test_cli_writes_flat_json_and_names_the_hot_definition(tmp_path):
    truth = phantom()
    truth.tofile(tmp_path / "truth.raw")
    (truth > 1.5).astype(np.float32).tofile(tmp_path / "hot.raw")
    r = subprocess.run([sys.executable, "-m", "prc_mlem_amd.evaluate", "--truth", str(tmp_path / "truth.raw"),
                        "--shape", "35", "35", "35", "--hot-mask", str(tmp_path / "hot.raw"), "--background", "1",
                        "--json", str(tmp_path / "m.json"), f"identity={tmp_path / 'truth.raw'}"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    report = json.loads((tmp_path / "m.json").read_text())
    assert set(report) == {"identity"} and report["identity"]["rc_hot"] == 1
    assert "explicit hot mask" in r.stdout
