#!/usr/bin/env python3
"""Reference-based image metrics for corrected volumes: NRMSE, 3D SSIM, counts, recovery, background
noise, local-range (ring) violations and halo.

Drop-in for the reference's evaluator (tools/evaluate.py:44-150, CLI tools/evaluate.py:169-192):
same metric names and definitions, same command line. scikit-image is not needed: the 3D SSIM is
restated here from its published definition with scikit-image's conventions (Gaussian window
sigma 1.5 truncated at 3.5 sigma = 11 voxels, mirrored edges, population covariances,
K1 = 0.01, K2 = 0.03, mean over the map cropped by the window radius).

Usage: python -m prc_mlem_amd.evaluate --truth TRUTH.raw --shape NZ NY NX name=image.raw [...]
"""
import argparse
import json
from pathlib import Path

import numpy as np
from scipy import ndimage as ndi

SSIM_SIGMA, SSIM_TRUNCATE, SSIM_K1, SSIM_K2 = 1.5, 3.5, 0.01, 0.03
SSIM_RADIUS = int(SSIM_TRUNCATE * SSIM_SIGMA + 0.5)  # 5: an 11-voxel window
COLUMNS = ["nrmse", "ssim", "ssim_tissue", "ssim_edge", "edge_nrmse", "rc_hot", "rc_min", "bg_cov", "ring", "halo"]


def load(path, shape):
    """Headerless float32 volume of the given (NZ, NY, NX) shape; the file size must match."""
    if len(shape) != 3 or min(shape) <= 0:
        raise ValueError("shape must contain three positive dimensions (NZ NY NX)")
    want = int(np.prod(shape)) * 4
    got = Path(path).stat().st_size
    if got != want:
        raise ValueError(f"{path}: expected {want} bytes, got {got}")
    return np.fromfile(path, dtype=np.float32).reshape(shape)


def ssim_map(a, b, data_range):
    """Local SSIM of two 3D volumes (float64): Gaussian-weighted means, population variances."""
    blur = lambda v: ndi.gaussian_filter(v, sigma=SSIM_SIGMA, truncate=SSIM_TRUNCATE, mode="reflect")
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    mu_a, mu_b = blur(a), blur(b)
    var_a = blur(a * a) - mu_a * mu_a
    var_b = blur(b * b) - mu_b * mu_b
    cov = blur(a * b) - mu_a * mu_b
    c1, c2 = (SSIM_K1 * data_range) ** 2, (SSIM_K2 * data_range) ** 2
    return ((2 * mu_a * mu_b + c1) * (2 * cov + c2)) / ((mu_a ** 2 + mu_b ** 2 + c1) * (var_a + var_b + c2))


def _rel(diff, ref):
    den = float(np.sum(ref ** 2))
    return float(np.sqrt(np.sum(diff ** 2) / den)) if den > 0 else None


def _mean(values):
    return float(values.mean()) if values.size else None


def _check_like(name, array, shape):
    array = np.asarray(array)
    if array.shape != shape or not np.all(np.isfinite(array)):
        raise ValueError(f"{name} must be finite and have the truth shape")
    return array


class Evaluator:
    """Metrics of reconstructions against a known, nonnegative truth volume.

    mask: tissue mask (positive = tissue; may include cold regions), default truth > 0.
    data_range: SSIM/PSNR intensity range, default truth.max().
    background: background level, default the most frequent positive tissue value.
    hot_mask: explicit lesion mask, default truth > 1.5 * background inside tissue.
    """

    def __init__(self, truth, mask=None, data_range=None, background=None, hot_mask=None):
        truth = np.asarray(truth)
        if truth.ndim != 3 or not np.all(np.isfinite(truth)) or np.any(truth < 0):
            raise ValueError("truth must be a finite, nonnegative 3D activity volume")
        if not np.any(truth > 0):
            raise ValueError("truth must contain positive activity")
        self.t = truth.astype(np.float64)
        if mask is None:
            self.body = truth > 0
        else:
            self.body = _check_like("mask", mask, truth.shape) > 0
            if not self.body.any():
                raise ValueError("mask must contain tissue voxels")
        self.data_range = float(truth.max() if data_range is None else data_range)
        if not (np.isfinite(self.data_range) and self.data_range > 0):
            raise ValueError("data_range must be finite and positive")
        if background is None:
            levels, freq = np.unique(truth[self.body & (truth > 0)], return_counts=True)
            if not levels.size:
                raise ValueError("mask must include positive truth activity, or supply background")
            background = levels[np.argmax(freq)]
        self.bg = float(background)
        if not (np.isfinite(self.bg) and self.bg > 0):
            raise ValueError("background must be finite and positive")
        if hot_mask is None:
            self.hot = (truth > 1.5 * self.bg) & self.body
            hot_text = "truth > 1.5 * background (high-uptake regions; not necessarily lesions)"
        else:
            self.hot = _check_like("hot_mask", hot_mask, truth.shape) > 0
            if np.any(self.hot & (~self.body | (truth <= 0))):
                raise ValueError("hot_mask must lie inside tissue with positive truth activity")
            hot_text = "explicit hot mask; connected positive components define lesions"
        self.hot_labels, self.n_hot = ndi.label(self.hot)
        self.metric_settings = {
            "data_range": self.data_range, "background": self.bg,
            "tissue_mask": "truth > 0" if mask is None else "explicit mask",
            "hot_regions": hot_text, "hot_region_count": int(self.n_hot),
            "hot_region_voxels": int(self.hot.sum()),
            "ssim_window": "3D Gaussian sigma 1.5, 11 voxels",
        }
        # Region of interest: the tissue's bounding box grown by four voxels.
        self.roi = ndi.find_objects(ndi.binary_dilation(self.body, iterations=4).astype(np.uint8))[0]
        # Background interior (relative tolerance keeps membership under unit changes).
        at_bg = (np.abs(self.t - self.bg) < 1e-6 * self.bg) & self.body
        self.bg_interior = ndi.binary_erosion(at_bg, structure=np.ones((3, 3, 3)), iterations=4)
        # Local range of the truth over 5x5x5 neighbourhoods; edges are where it is not flat.
        self.lo = ndi.minimum_filter(truth, size=5, mode="nearest").astype(np.float64)
        self.hi = ndi.maximum_filter(truth, size=5, mode="nearest").astype(np.float64)
        self.edge = self.hi > self.lo
        # Halo shell: background voxels two to three voxels outside the hot regions, off the skin.
        shell = ndi.binary_dilation(self.hot, iterations=3) & ~ndi.binary_dilation(self.hot, iterations=1)
        self.halo_shell = shell & ndi.binary_erosion(self.body, iterations=4) & at_bg

    def __call__(self, f):
        f = _check_like("reconstruction", f, self.t.shape).astype(np.float64)
        t, body = self.t, self.body
        d = f - t
        m = {
            "nrmse": _rel(d, t),
            "nrmse_tissue": _rel(d[body], t[body]),
            "mse": float(np.mean(d ** 2)),
            "mae_tissue": _mean(np.abs(d[body])),
            "count_bias": float(f.sum() / t.sum() - 1.0),
        }
        fr, tr = f[self.roi], t[self.roi]
        rmse = float(np.sqrt(np.mean((fr - tr) ** 2)))
        m["psnr"] = float(20.0 * np.log10(self.data_range / rmse)) if rmse > 0 else None
        m["ssim"] = m["ssim_tissue"] = m["ssim_edge"] = None
        if min(tr.shape) >= 2 * SSIM_RADIUS + 1:
            smap = ssim_map(tr, fr, self.data_range)
            inner = (slice(SSIM_RADIUS, -SSIM_RADIUS),) * 3
            m["ssim"] = float(smap[inner].mean())
            m["ssim_tissue"] = _mean(smap[inner][body[self.roi][inner]])
            m["ssim_edge"] = _mean(smap[inner][self.edge[self.roi][inner]])
        m["rc_hot"] = m["rc_min"] = None
        if self.n_hot:
            m["rc_hot"] = float(f[self.hot].mean() / t[self.hot].mean())
            ids = np.arange(1, self.n_hot + 1)
            m["rc_min"] = float(np.min(ndi.mean(f, self.hot_labels, ids) / ndi.mean(t, self.hot_labels, ids)))
        bgv = f[self.bg_interior]
        m["bg_bias"] = float(bgv.mean() / self.bg - 1.0) if bgv.size else None
        m["bg_cov"] = float(bgv.std() / bgv.mean()) if bgv.size and bgv.mean() > 0 else None
        over = np.maximum(f - self.hi, 0.0)[self.edge]
        under = np.maximum(self.lo - f, 0.0)[self.edge]
        has_edge = over.size > 0
        m["ring"] = float(np.sqrt(np.mean((over + under) ** 2)) / self.bg) if has_edge else None
        m["overshoot"] = float(np.sqrt(np.mean(over ** 2)) / self.bg) if has_edge else None
        m["undershoot"] = float(np.sqrt(np.mean(under ** 2)) / self.bg) if has_edge else None
        m["edge_nrmse"] = _rel(d[self.edge], t[self.edge])
        halo = _mean(f[self.halo_shell])
        m["halo"] = 1.0 - halo / self.bg if halo is not None else None
        return m


def format_table(rows):
    """Markdown table of the main metrics; missing values print as an em dash."""
    out = ["| image | " + " | ".join(COLUMNS) + " |", "|---|" + "---:|" * len(COLUMNS)]
    for name, m in rows:
        out.append(f"| {name} | " + " | ".join("—" if m.get(c) is None else f"{m[c]:.4f}" for c in COLUMNS) + " |")
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--truth", required=True)
    ap.add_argument("--shape", type=int, nargs=3, required=True, metavar=("NZ", "NY", "NX"))
    ap.add_argument("--json", help="write metrics to this JSON file")
    ap.add_argument("--mask", help="float32 tissue mask (positive = tissue), including cold regions")
    ap.add_argument("--hot-mask", help="float32 true hot-lesion mask (positive = lesion)")
    ap.add_argument("--data-range", type=float, help="fixed reference intensity range for SSIM/PSNR")
    ap.add_argument("--background", type=float, help="known background level for phantom metrics")
    ap.add_argument("images", nargs="+", help="name=path.raw")
    a = ap.parse_args(argv)
    shape = tuple(a.shape)
    ev = Evaluator(load(a.truth, shape), mask=load(a.mask, shape) if a.mask else None, data_range=a.data_range,
                   background=a.background, hot_mask=load(a.hot_mask, shape) if a.hot_mask else None)
    rows = []
    for item in a.images:
        name, path = item.rsplit("=", 1)
        rows.append((name, ev(load(path, shape))))
    print(format_table(rows))
    print("Hot-region definition: " + ev.metric_settings["hot_regions"])
    if a.json:
        Path(a.json).write_text(json.dumps(dict(rows), indent=2, allow_nan=False))


if __name__ == "__main__":
    main()
