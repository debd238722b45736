#!/usr/bin/env python3
"""Corrected volumes/s on the 128x128x159 68Ga thorax workload: mixed + batched vs plain FP32.

One step = the full correction of one batch: the normalisation pass (every system-matrix
weight marched and stored) followed by 6 accelerated MAP-EM iterations (beta 0.03, gamma
20, 128 far rays), for each volume of the batch. Inputs (density and measured counts) are
resident in HBM before the timed region; nothing is copied over PCIe inside it.

Prints one JSON line. `value` is the mixed + batched throughput; the plain FP32 path (one
volume per launch, no graph), the per-kernel timings (HIP events, a profiled step outside
the timed region), the roofline of the dominant kernel and a CPU baseline (the oracle on a
bounded slab of the same workload, extrapolated by its voxel-ray-step rate) sit beside it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import prc_mlem_amd as prc  # noqa: E402
from prc_mlem_amd import phantoms  # noqa: E402

SHAPE = (159, 128, 128)
SPACING = (0.776, 0.776, 0.795)
ISOTOPE, RAYS, BETA, GAMMA, ITERS, COUNTS = "68Ga", 128, 0.03, 20.0, 6, 5e7
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload(seed):
    rho, truth = phantoms.thorax_phantom(SHAPE, SPACING)
    return rho, truth


def timed(fn, steps, warmup, sync):
    for _ in range(warmup):
        fn()
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    sync()
    return time.perf_counter() - t0


def measured_traffic(kernel, form, sched=None):
    """HBM bytes per launch of `kernel` on this workload from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the same counter's reading for a
    plain copy of known size), or None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_thorax_128x128x159.md")
    try:
        with open(path) as fh:
            text = fh.read()
        table = json.loads(text.split("```json", 1)[1].split("```", 1)[0])
    except (OSError, ValueError, IndexError):
        return None
    key = f"{kernel}/{form}" + (f"/{sched}" if sched else "")
    entry = table.get("kernels", {}).get(key)
    return None if entry is None else entry.get("hbm_bytes_per_launch")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_correction(rho, data, spacing, iso, rays, iters, beta, gamma, accel):
    """The FP32 oracle (reference semantics) on one volume: (normalisation s, iterations s, VRS)."""
    from oracle import oracle
    r = np.ascontiguousarray(np.maximum(rho, np.float32(1e-3)))
    g = np.where(r >= np.float32(0.05), np.maximum(data, 0), 0).astype(np.float32)
    t0 = time.perf_counter()
    P = oracle.Projector(r, spacing, iso, rays)
    t1 = time.perf_counter()
    P.solve(g, iters, beta, gamma, accel)
    t2 = time.perf_counter()
    near, far = P.vrs
    return t1 - t0, t2 - t1, near + far


def cpu_baselines(rho, data, slices_all, slices_one, n_gpu_vrs):
    """CPU baselines on the host, the oracle in FP32 (the reference's arithmetic; it has no CPU
    path of its own): the 64^3 18F cylinder workload in full (Richardson-Lucy, 64 rays, 10
    iterations), and Z-slabs of the thorax workload (the full volume would take minutes), each on
    every host thread and on one thread. The thorax figure in corrected vols/s is the slab's
    voxel-ray-step rate applied to the full volume's step count."""
    from oracle import oracle
    threads = oracle.threads()
    out = {"cpu_model": cpu_model(), "host_threads": threads, "nproc": os.cpu_count()}
    rho1, truth1 = phantoms.cylinder_phantom()
    A1 = prc.Projector(rho1, (2.0, 2.0, 2.0), "18F", 64)
    g1 = phantoms.poisson_counts(A1.forward(truth1), 1e6, seed=3001)
    A1.close()
    for label, nthreads, slices in (("all", threads, slices_all), ("one", 1, slices_one)):
        oracle.set_threads(nthreads)
        tb, ti, vrs = oracle_correction(rho1, g1, (2.0, 2.0, 2.0), "18F", 64, 10, 0.0, 20.0, False)
        out[f"cylinder_64_18F_{label}_threads"] = {"threads": nthreads, "normalisation_s": round(tb, 3),
                                                   "s_per_iteration": round(ti / 10, 4),
                                                   "vols_per_s": round(1.0 / (tb + ti), 5)}
        z0 = SHAPE[0] // 2 - slices // 2
        tb, ti, vrs = oracle_correction(rho[z0:z0 + slices], data[z0:z0 + slices], SPACING, ISOTOPE, RAYS, ITERS,
                                        BETA, GAMMA, True)
        rate = vrs / (tb + ti)
        out[f"thorax_slab_{label}_threads"] = {
            "threads": nthreads, "slices": slices, "normalisation_s": round(tb, 3),
            "s_per_iteration": round(ti / ITERS, 4), "vrs_per_s": round(rate, 1),
            "full_volume_vols_per_s": round(rate / n_gpu_vrs, 6)}
    oracle.set_threads(threads)
    return out


def accuracy_vs_fp64(rho, data, crop):
    """NRMSE, 3D SSIM and count error of the GPU corrections (plain FP32 and mixed) against the FP64
    oracle on a central crop of the workload (the FP64 CPU oracle evaluates the range model exactly
    at every step; the full volume would take it many minutes)."""
    from oracle import oracle
    from prc_mlem_amd.evaluate import Evaluator
    sl = tuple(slice(n // 2 - c // 2, n // 2 - c // 2 + c) for n, c in zip(SHAPE, crop))
    r = np.ascontiguousarray(np.maximum(rho[sl], np.float32(1e-3)))
    g = np.ascontiguousarray(np.where(r >= np.float32(0.05), np.maximum(data[sl], 0), 0).astype(np.float32))
    t0 = time.perf_counter()
    ref = oracle.Projector(r, SPACING, ISOTOPE, RAYS, prec=64).solve(g, ITERS, BETA, GAMMA, True).reshape(r.shape)
    ev = Evaluator(np.maximum(ref, 0))
    out = {"region": "central crop %dx%dx%d (NZ NY NX) of the workload" % tuple(crop), "reference": "FP64 oracle",
           "oracle_seconds": round(time.perf_counter() - t0, 1)}
    for form in ("fp32", "mixed"):
        A = prc.Projector(r, SPACING, ISOTOPE, RAYS, precision=form)
        est = np.asarray(prc.correct(A, g, ITERS, BETA, GAMMA, True)).reshape(r.shape)
        A.close()
        m = ev(est)
        out[form] = {"nrmse": m["nrmse"], "ssim": m["ssim"], "count_error": m["count_bias"]}
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=4, help="volumes per batched step")
    ap.add_argument("--no-pipeline", dest="pipeline", action="store_false",
                    help="mixed form: one batch at a time (no overlap of normalisation and iterations)")
    ap.add_argument("--only", choices=["both", "mixed", "fp32"], default="both")
    ap.add_argument("--cpu-slices", type=int, default=48, help="Z slices of the CPU-baseline slab (0 = skip)")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--accuracy-crop", type=int, nargs=3, default=[16, 64, 64], metavar=("NZ", "NY", "NX"),
                    help="central crop on which both corrections are compared with the FP64 oracle (0 0 0 = skip)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES bytes three times (the known-size kernel that "
                         "scales the HBM byte counters of a profiler run) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = args.copy_probe // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 4 * n, "reads_per_launch": 4 * n, "writes_per_launch": 4 * n}))
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} (WORLD_SIZE {world}) is not supported; the "
              "batched volumes would be sharded across GPUs by whole volumes with no communication.",
              file=sys.stderr)
        sys.exit(2)

    prc.lib()
    log("phantom + blurred data")
    rho, truth = workload(0)
    A_blur = prc.Projector(rho, SPACING, ISOTOPE, RAYS)
    blurred = A_blur.forward(truth)
    A_blur.close()
    datas = [phantoms.poisson_counts(blurred, COUNTS, seed=2000 + v) for v in range(max(args.batch, 1))]
    n = rho.size
    result = {}

    # Plain FP32: one volume per launch, direct launches.
    fp32 = None
    if args.only in ("both", "fp32"):
        log("plain fp32: build")
        A = prc.Projector(rho, SPACING, ISOTOPE, RAYS, precision="fp32")
        S = prc.Solver(A, datas[0], BETA, GAMMA, True)

        def step_fp32():
            A.normalise()
            S.reset()
            S.iterate(ITERS)

        t = timed(step_fp32, args.steps, args.warmup, S.sync)
        st = A.stats
        fp32 = dict(vols_per_s=args.steps / t, ms_per_volume=1e3 * t / args.steps, weight_bytes=st["weight_bytes"],
                    weight_entries=st["weight_entries"], vrs_near=st["vrs_near"], vrs_far=st["vrs_far"],
                    n_active=st["n_active"], deposits_far=st["deposits"], n_rows=st["n_rows"],
                    nonzero_entries=st["nonzero_entries"],
                    target_entries=st["target_entries"], near_uniform=st["near_uniform"],
                    weight_planes=st["weight_planes"], target_planes=st["target_planes"])
        est32 = S.estimate()
        log(f"plain fp32: {fp32['vols_per_s']:.3f} vols/s, weights {st['weight_bytes'] / 2**30:.2f} GiB")
        if not args.no_profile:
            prc.profile_enable(True)
            step_fp32()
            S.sync()
            fp32["kernels"] = prc.profile_report()
            prc.profile_enable(False)
        S.close()
        A.close()

    # Mixed + batched: FP16 weights, every stage one launch over the batch, graph-captured loop.
    # Sequential: normalise the batch's projectors, then iterate. Pipelined (the headline form):
    # two batches in flight, each building and iterating on its own stream, so that the
    # normalisation of one batch (VALU-bound march) overlaps the iterations of the other
    # (HBM-bound passes). Every step still normalises and corrects `batch` volumes.
    mixed = None
    if args.only in ("both", "mixed"):
        sets = 2 if args.pipeline else 1
        log(f"mixed+batched x{args.batch}: build ({sets} batch{'es' if sets > 1 else ''})")
        As = [[prc.Projector(rho, SPACING, ISOTOPE, RAYS, precision="mixed") for _ in range(args.batch)]
              for _ in range(sets)]
        Bs = [prc.Batch(a, datas[:args.batch], BETA, GAMMA, True) for a in As]
        turn = [0]

        def step_set(k):
            Bs[k].normalise()  # the batch's projectors build on the batch's stream, then it iterates
            Bs[k].run(ITERS)

        def step_seq():
            step_set(0)
            Bs[0].sync()

        def step_pipe():
            k = turn[0] % sets
            turn[0] += 1
            step_set(k)

        def sync_all():
            for b in Bs:
                b.sync()

        t_seq = timed(step_seq, args.steps, args.warmup, sync_all)
        mixed = dict(vols_per_s_sequential=args.batch * args.steps / t_seq)
        if sets > 1:
            t = timed(step_pipe, args.steps, max(args.warmup, sets), sync_all)
            mixed.update(vols_per_s=args.batch * args.steps / t, ms_per_step=1e3 * t / args.steps,
                         batches_in_flight=sets)
        else:
            mixed.update(vols_per_s=mixed["vols_per_s_sequential"], ms_per_step=1e3 * t_seq / args.steps,
                         batches_in_flight=1)
        st = As[0][0].stats
        mixed.update(weight_bytes_per_volume=st["weight_bytes"], weight_entries=st["weight_entries"],
                     nonzero_entries=st["nonzero_entries"],
                     target_entries=st["target_entries"], vrs_near=st["vrs_near"], vrs_far=st["vrs_far"],
                     weight_planes=st["weight_planes"], target_planes=st["target_planes"])
        Bs[0].run(ITERS)
        est_mixed = Bs[0].estimate(0)
        log(f"mixed+batched: {mixed['vols_per_s']:.3f} vols/s ({mixed['vols_per_s_sequential']:.3f} sequential)")
        if not args.no_profile:
            # Profiled step: the batch's projectors normalise one after another here, so that each
            # phase's event time is one volume's, uncontended; then the batch iterates.
            prc.profile_enable(True)
            for a in As[0]:
                a.normalise()
            Bs[0].run(ITERS)
            Bs[0].sync()
            mixed["kernels"] = prc.profile_report()
            prc.profile_enable(False)
        if fp32 is not None:
            d = est_mixed.astype(np.float64) - est32
            result["nrmse_mixed_vs_fp32"] = float(np.linalg.norm(d) / np.linalg.norm(est32))
            result["count_error_mixed_vs_fp32"] = float(est_mixed.sum(dtype=np.float64) /
                                                        est32.sum(dtype=np.float64) - 1)
        for b in Bs:
            b.close()
        for a in As:
            for p in a:
                p.close()

    head = mixed if mixed is not None else fp32
    value = head["vols_per_s"]
    k = head.get("kernels", {})

    # Roofline of the dominant kernel of the headline step (by total event time). The stored-
    # weight projector passes are HBM streams: algorithmic bytes per launch = stored weights of
    # every volume in the launch + the voxel vectors they read and write.
    roofline = None
    if k:
        name = max(k, key=lambda kk: k[kk]["total_ms"])
        vols = args.batch if mixed is not None else 1
        st0 = mixed if mixed is not None else fp32
        b_w = 2 if mixed is not None else 4
        # Algorithmic bytes: the nonzero stored weights (each weight of A once, in either copy),
        # the planes' shift lists and the vectors. The zero entries that pad a segment's planes past
        # the end of its shorter rays are streamed too but are not algorithmic; the stored-byte
        # figures sit beside as `streamed_*`.
        wbytes = st0["nonzero_entries"] * b_w * vols
        wbytes_t = st0["nonzero_entries"] * b_w * vols
        streamed = {"forward": st0["target_entries"] * b_w * vols, "adjoint": st0["weight_entries"] * b_w * vols}
        n_act = (fp32 or {}).get("n_active", int(np.sum(np.maximum(rho, 1e-3) >= 0.05)))
        b_v = 2 if mixed is not None else 4  # the mixed solver gathers FP16 copies of its iterates
        # + the planes' shift lists (4 B per plane), read once per launch like the weights
        lists_s, lists_t = 4 * st0["weight_planes"] * vols, 4 * st0["target_planes"] * vols
        per_launch = {"forward": wbytes_t + lists_t + vols * (b_v * n + 4 * n),
                      "adjoint": wbytes + lists_s + vols * (b_v * n + 8 * n_act + n)}
        if name in per_launch:
            mean_s = k[name]["mean_ms"] / 1e3
            achieved = per_launch[name] / mean_s / 1e9
            streamed_bytes = per_launch[name] - (wbytes if name == "adjoint" else wbytes_t) + streamed[name]
            roofline = {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                        "bytes_per_launch": per_launch[name], "mean_ms": k[name]["mean_ms"],
                        "streamed_bytes_per_launch": streamed_bytes,
                        "streamed_GBps": round(streamed_bytes / mean_s / 1e9, 1)}
        else:
            roofline = {"kernel": name, "bound": "valu", "achieved": None, "peak": None, "unit": None,
                        "frac": None, "traffic": None, "mean_ms": k[name]["mean_ms"]}

    # Projector throughput in the effective accounting of the on-the-fly design: one call = the far voxel-ray
    # steps plus 125 near taps per active source, at 8 B (FP32) or 6 B (mixed) per step. These
    # are effective bytes, not HBM traffic (the stored-weight stream is the roofline above).
    projector = {}
    for form, d, vols, b_step in (("fp32", fp32, 1, 8), ("mixed", mixed, args.batch, 6)):
        if d is None or "kernels" not in d:
            continue
        n_act = (fp32 or {}).get("n_active", int(np.sum(np.maximum(rho, 1e-3) >= 0.05)))
        vrs_call = d["vrs_far"] + 125 * n_act
        entry = {"vrs_per_call": vrs_call, "bytes_per_step": b_step}
        for op in ("forward", "adjoint"):
            if op in d["kernels"]:
                t = d["kernels"][op]["mean_ms"] / 1e3 / vols  # per volume
                entry[op] = {"ms_per_volume": round(t * 1e3, 3), "vrs_per_s": round(vrs_call / t, 1),
                             "effective_GBps": round(vrs_call * b_step / t / 1e9, 1),
                             "effective_frac_of_hbm_peak": round(vrs_call * b_step / t / 1e9 / HBM_PEAK_GBS, 3)}
        projector[form] = entry
    tuning = prc.tuning_report()
    if roofline is not None:
        form = "mixed" if mixed is not None else "fp32"
        sched = None
        if roofline["kernel"] in ("forward", "adjoint"):
            mine = [t for t in tuning if t["weight_bytes"] == (2 if mixed is not None else 4)]
            sched = mine[-1][roofline["kernel"]] if mine else None
            roofline["schedule"] = sched
        roofline["traffic"] = measured_traffic(roofline["kernel"], form, sched)

    cpu = None
    cpu_detail = None
    if args.cpu_slices > 0 and fp32 is not None:
        log(f"cpu baselines: oracle, all threads and one thread")
        # The oracle marches every near block: the full volume's step count without the memo.
        full = prc.Projector(rho, SPACING, ISOTOPE, RAYS, near_memo=False)
        vrs_full = full.stats["vrs_near"] + (1 + 2 * ITERS) * fp32["vrs_far"]
        full.close()
        cpu_detail = cpu_baselines(rho, datas[0], args.cpu_slices, max(2, args.cpu_slices // 8), vrs_full)
        slab = cpu_detail["thorax_slab_all_threads"]
        cpu = {"value": slab["full_volume_vols_per_s"], "unit": "corrected vols/s", "cores": slab["threads"],
               "kind": "port",
               "sample": f"FP32 oracle, normalisation + {ITERS} iterations on a {slab['slices']}-slice Z slab of the "
                         f"thorax workload ({slab['normalisation_s']} s + {slab['s_per_iteration']} s per iteration, "
                         f"{slab['vrs_per_s']:.3g} voxel-ray steps/s on {slab['threads']} threads of "
                         f"{cpu_detail['cpu_model']}); vols/s = that rate over the full volume's {vrs_full:.3g} steps"}

    accuracy = None
    if min(args.accuracy_crop) > 0 and args.cpu_slices > 0:
        log(f"accuracy: GPU corrections vs the FP64 oracle on a {args.accuracy_crop} crop")
        accuracy = accuracy_vs_fp64(rho, datas[0], args.accuracy_crop)

    out = {
        "metric": "corrected vols/s + projector GB/s, 128x128x159 68Ga: mixed+batched vs FP32",
        "value": round(value, 4), "unit": "corrected vols/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(head.get("ms_per_step", head.get("ms_per_volume")), 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f16 weights, f32 arithmetic, f64 reductions" if mixed is not None else "f32",
        "data": "synthetic thorax phantom (seeded), forward-blurred and Poisson-noised to 5e7 counts",
        "config": {"workload": "68Ga thorax 128x128x159 @ 0.776x0.776x0.795 mm, 128 far rays, beta 0.03, "
                               "gamma 20, Biggs-Andrews on, 6 iterations incl. normalisation",
                   "batch": args.batch if mixed is not None else 1, "form": "mixed+batched" if mixed else "fp32",
                   "batches_in_flight": mixed["batches_in_flight"] if mixed else 1},
        "fp32_vols_per_s": round(fp32["vols_per_s"], 4) if fp32 else None,
        "speedup_mixed_batched_vs_fp32": round(mixed["vols_per_s"] / fp32["vols_per_s"], 3) if (mixed and fp32) else None,
        "roofline": roofline, "cpu_baseline": cpu, "projector_effective": projector,
    }
    out.update(result)
    out["cpu_baselines"] = cpu_detail
    out["accuracy_vs_fp64"] = accuracy
    out["projector_tuning"] = tuning
    if fp32 is not None:
        out["fp32"] = fp32
    if mixed is not None:
        out["mixed"] = mixed
    print(json.dumps(out))


if __name__ == "__main__":
    main()
