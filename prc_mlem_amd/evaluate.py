#!/usr/bin/env python3
"""Image-quality scores of corrected volumes against a known truth: NRMSE, 3D SSIM, counts,
recovery, background noise, local-range (ring) violations and halo.

Interface of the reference's evaluator (tools/evaluate.py:44-197): the metric names, the
Evaluator(truth, mask, data_range, background, hot_mask) signature, the table columns, the
command line and the printed hot-region line are the same, so its users and tests can switch.
The computation is written here from the definitions below, and the 3D SSIM is restated from its
published definition with scikit-image's conventions (scikit-image is not needed): Gaussian
window sigma 1.5 truncated at 3.5 sigma (11 voxels), mirrored edges, population covariances,
K1 = 0.01, K2 = 0.03, mean over the map cropped by the window radius.

Definitions (t = truth, f = reconstruction, both float64):
  tissue      explicit mask > 0, else t > 0
  background  given, else the most frequent positive truth value in tissue
  hot         explicit lesion mask, else tissue with t > 1.5 * background
  ROI         bounding box of the tissue grown by 4 voxels per side (clipped to the volume)
  flat        tissue voxels at the background level (relative tolerance 1e-6)
  bg interior flat voxels whose whole 9x9x9 neighbourhood is flat and inside the volume
  lo, hi      min and max of t over 5x5x5 neighbourhoods (edges replicated); edge band: hi > lo
  halo shell  flat voxels at city-block distance 2-3 from the hot set and more than 4 from the
              outside of the tissue (the volume's outside counts as outside)

Usage: python -m prc_mlem_amd.evaluate --truth TRUTH.raw --shape NZ NY NX name=image.raw [...]
"""
import argparse
import json
from dataclasses import dataclass
from pathlib import Path

import numpy as np
from scipy import ndimage as ndi

SSIM_SIGMA, SSIM_TRUNCATE, SSIM_K1, SSIM_K2 = 1.5, 3.5, 0.01, 0.03
SSIM_RADIUS = int(SSIM_TRUNCATE * SSIM_SIGMA + 0.5)  # 5: an 11-voxel window
COLUMNS = ["nrmse", "ssim", "ssim_tissue", "ssim_edge", "edge_nrmse", "rc_hot", "rc_min", "bg_cov", "ring", "halo"]
HOT_BY_THRESHOLD = "truth > 1.5 * background (high-uptake regions; not necessarily lesions)"
HOT_BY_MASK = "explicit hot mask; connected positive components define lesions"


def load(path, shape):
    """Headerless float32 volume of the given (NZ, NY, NX) shape; the file size must match."""
    shape = tuple(int(s) for s in shape)
    if len(shape) != 3 or min(shape) <= 0:
        raise ValueError("shape must contain three positive dimensions (NZ NY NX)")
    nbytes = 4 * int(np.prod(shape))
    size = Path(path).stat().st_size
    if size != nbytes:
        raise ValueError(f"{path}: expected {nbytes} bytes, got {size}")
    return np.fromfile(path, dtype=np.float32).reshape(shape)


def ssim_map(a, b, data_range):
    """Local SSIM of two 3D volumes (float64): Gaussian-weighted means, population variances."""
    def window_mean(v):
        return ndi.gaussian_filter(v, sigma=SSIM_SIGMA, truncate=SSIM_TRUNCATE, mode="reflect")

    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    ma, mb = window_mean(a), window_mean(b)
    saa = window_mean(a * a) - ma * ma
    sbb = window_mean(b * b) - mb * mb
    sab = window_mean(a * b) - ma * mb
    c1, c2 = (SSIM_K1 * data_range) ** 2, (SSIM_K2 * data_range) ** 2
    return ((2 * ma * mb + c1) * (2 * sab + c2)) / ((ma * ma + mb * mb + c1) * (saa + sbb + c2))


def _finite_volume(what, v, shape):
    v = np.asarray(v)
    if v.shape != shape or not np.isfinite(v).all():
        raise ValueError(f"{what} must be finite and have the truth shape")
    return v


def _modal_level(values):
    """Most frequent value (the smallest among equally frequent ones)."""
    levels, counts = np.unique(values, return_counts=True)
    return levels[np.argmax(counts)]


def _grown_box(mask, grow):
    """Slices of the bounding box of `mask` grown by `grow` voxels per side, clipped."""
    idx = np.nonzero(mask)
    return tuple(slice(max(int(i.min()) - grow, 0), min(int(i.max()) + grow + 1, n))
                 for i, n in zip(idx, mask.shape))


def _taxicab_distance(mask):
    """City-block distance of every voxel to the nearest voxel of `mask`, in voxels."""
    return ndi.distance_transform_cdt(~mask, metric="taxicab")


def _taxicab_depth(mask):
    """City-block distance of every voxel of `mask` to the nearest voxel outside it, the
    volume's outside included."""
    padded = np.pad(mask, 1, constant_values=False)
    return ndi.distance_transform_cdt(padded, metric="taxicab")[1:-1, 1:-1, 1:-1]


@dataclass(frozen=True)
class Regions:
    tissue: np.ndarray
    hot: np.ndarray
    hot_labels: np.ndarray
    n_hot: int
    roi: tuple
    bg_interior: np.ndarray
    lo: np.ndarray
    hi: np.ndarray
    edge: np.ndarray
    halo_shell: np.ndarray


def _regions(t, tissue, background, hot):
    flat = tissue & (np.abs(t - background) < 1e-6 * background)
    # 9x9x9 all-flat neighbourhood, voxels near the volume's edge excluded.
    interior = ndi.minimum_filter(flat.astype(np.uint8), size=9, mode="constant", cval=0) > 0
    lo = ndi.minimum_filter(t, size=5, mode="nearest")
    hi = ndi.maximum_filter(t, size=5, mode="nearest")
    if hot.any():
        d = _taxicab_distance(hot)
        shell = (d >= 2) & (d <= 3)
    else:
        shell = np.zeros_like(hot)
    labels, n = ndi.label(hot)
    return Regions(tissue=tissue, hot=hot, hot_labels=labels, n_hot=int(n), roi=_grown_box(tissue, 4),
                   bg_interior=interior, lo=lo, hi=hi, edge=hi > lo,
                   halo_shell=shell & flat & (_taxicab_depth(tissue) > 4))


def _relative_l2(err, ref):
    ref2 = float(np.sum(ref * ref))
    return None if ref2 <= 0 else float(np.sqrt(np.sum(err * err) / ref2))


def _mean(v):
    return None if v.size == 0 else float(np.mean(v))


# Metric groups: each maps (reconstruction, truth, regions, evaluator) to named scores; None marks
# a score the volume cannot define (no hot region, no edge, volume too small for the window).
def _global_scores(f, t, r, ev):
    err = f - t
    return {"nrmse": _relative_l2(err, t), "nrmse_tissue": _relative_l2(err[r.tissue], t[r.tissue]),
            "mse": float(np.mean(err * err)), "mae_tissue": _mean(np.abs(err[r.tissue])),
            "count_bias": float(np.sum(f) / np.sum(t) - 1.0)}


def _roi_scores(f, t, r, ev):
    fr, tr = f[r.roi], t[r.roi]
    rmse = float(np.sqrt(np.mean((fr - tr) ** 2)))
    out = {"psnr": None if rmse == 0 else float(20.0 * np.log10(ev.data_range / rmse)),
           "ssim": None, "ssim_tissue": None, "ssim_edge": None}
    if min(tr.shape) > 2 * SSIM_RADIUS:
        core = (slice(SSIM_RADIUS, -SSIM_RADIUS),) * 3
        s = ssim_map(tr, fr, ev.data_range)[core]
        out["ssim"] = float(np.mean(s))
        out["ssim_tissue"] = _mean(s[r.tissue[r.roi][core]])
        out["ssim_edge"] = _mean(s[r.edge[r.roi][core]])
    return out


def _recovery_scores(f, t, r, ev):
    if r.n_hot == 0:
        return {"rc_hot": None, "rc_min": None}
    ids = np.arange(1, r.n_hot + 1)
    per_region = ndi.mean(f, r.hot_labels, ids) / ndi.mean(t, r.hot_labels, ids)
    return {"rc_hot": float(np.mean(f[r.hot]) / np.mean(t[r.hot])), "rc_min": float(np.min(per_region))}


def _background_scores(f, t, r, ev):
    v = f[r.bg_interior]
    if v.size == 0:
        return {"bg_bias": None, "bg_cov": None}
    m = float(np.mean(v))
    return {"bg_bias": m / ev.bg - 1.0, "bg_cov": float(np.std(v)) / m if m > 0 else None}


def _local_range_scores(f, t, r, ev):
    above = np.maximum(f - r.hi, 0.0)[r.edge]
    below = np.maximum(r.lo - f, 0.0)[r.edge]

    def rms(v):
        return None if v.size == 0 else float(np.sqrt(np.mean(v * v)) / ev.bg)

    return {"ring": rms(above + below), "overshoot": rms(above), "undershoot": rms(below),
            "edge_nrmse": _relative_l2((f - t)[r.edge], t[r.edge])}


def _halo_scores(f, t, r, ev):
    m = _mean(f[r.halo_shell])
    return {"halo": None if m is None else 1.0 - m / ev.bg}


METRIC_GROUPS = (_global_scores, _roi_scores, _recovery_scores, _background_scores, _local_range_scores,
                 _halo_scores)


class Evaluator:
    """Scores of reconstructions against a known, nonnegative truth volume (definitions in the
    module docstring). mask: tissue mask; data_range: SSIM/PSNR intensity range (default
    truth.max()); background: background level; hot_mask: explicit lesion mask."""

    def __init__(self, truth, mask=None, data_range=None, background=None, hot_mask=None):
        t = np.asarray(truth)
        if t.ndim != 3 or not np.isfinite(t).all() or (t < 0).any():
            raise ValueError("truth must be a finite, nonnegative 3D activity volume")
        if not (t > 0).any():
            raise ValueError("truth must contain positive activity")
        self.t = t.astype(np.float64)
        if mask is None:
            tissue = self.t > 0
        else:
            tissue = _finite_volume("mask", mask, t.shape) > 0
            if not tissue.any():
                raise ValueError("mask must contain tissue voxels")
        self.data_range = float(self.t.max() if data_range is None else data_range)
        if not (np.isfinite(self.data_range) and self.data_range > 0):
            raise ValueError("data_range must be finite and positive")
        if background is None:
            positive = self.t[tissue & (self.t > 0)]
            if positive.size == 0:
                raise ValueError("the tissue holds no positive truth activity: give the background level")
            background = _modal_level(positive)
        self.bg = float(background)
        if not (np.isfinite(self.bg) and self.bg > 0):
            raise ValueError("background must be finite and positive")
        if hot_mask is None:
            hot = tissue & (self.t > 1.5 * self.bg)
            hot_text = HOT_BY_THRESHOLD
        else:
            hot = _finite_volume("hot_mask", hot_mask, t.shape) > 0
            if (hot & ~(tissue & (self.t > 0))).any():
                raise ValueError("hot_mask must lie inside tissue with positive truth activity")
            hot_text = HOT_BY_MASK
        self.regions = _regions(self.t, tissue, self.bg, hot)
        self.metric_settings = {
            "data_range": self.data_range, "background": self.bg,
            "tissue_mask": "explicit mask" if mask is not None else "truth > 0",
            "hot_regions": hot_text, "hot_region_count": self.regions.n_hot,
            "hot_region_voxels": int(hot.sum()), "ssim_window": "3D Gaussian sigma 1.5, 11 voxels",
        }

    # Names the reference evaluator's users read.
    body = property(lambda self: self.regions.tissue)
    roi = property(lambda self: self.regions.roi)
    hot = property(lambda self: self.regions.hot)

    def __call__(self, f):
        f = _finite_volume("reconstruction", f, self.t.shape).astype(np.float64)
        scores = {}
        for group in METRIC_GROUPS:
            scores.update(group(f, self.t, self.regions, self))
        return scores


def format_table(rows):
    """Markdown table of the main scores; a score that is not defined prints as an em dash."""
    def cell(v):
        return "—" if v is None else f"{v:.4f}"

    lines = ["| image | " + " | ".join(COLUMNS) + " |", "|---|" + "---:|" * len(COLUMNS)]
    lines += [f"| {name} | " + " | ".join(cell(m.get(c)) for c in COLUMNS) + " |" for name, m in rows]
    return "\n".join(lines)


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
    args = ap.parse_args(argv)
    shape = tuple(args.shape)

    def optional(path):
        return None if path is None else load(path, shape)

    ev = Evaluator(load(args.truth, shape), mask=optional(args.mask), data_range=args.data_range,
                   background=args.background, hot_mask=optional(args.hot_mask))
    rows = [(name, ev(load(path, shape))) for name, path in (item.rsplit("=", 1) for item in args.images)]
    print(format_table(rows))
    print("Hot-region definition: " + ev.metric_settings["hot_regions"])
    if args.json:
        Path(args.json).write_text(json.dumps(dict(rows), indent=2, allow_nan=False))


if __name__ == "__main__":
    main()
