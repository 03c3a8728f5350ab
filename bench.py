#!/usr/bin/env python3
"""Benchmark of the FDK back end on one MI355X.

Workload (dental512): 512^3 FP32 volume at 0.15 mm from 600 views of 768 x 768
uint16 at 0.2 mm, SOD 350 mm, SDD 600 mm, Shepp-Logan window, cubic beam
hardening, seeded 40-ellipsoid + tooth phantom, per-pixel gain and dark.

One step = the whole reconstruction on the device with the raw counts already
resident in HBM: preprocess -> row filter -> stage -> backprojection ->
normalised volume.  Two forms are timed:
  * fp32: FP32 Toeplitz filter on MFMA, FP32 projection storage;
  * bf16: hi/mid/lo split-BF16 Toeplitz filter on MFMA (6 products, FP32
    accumulation), BF16 projection storage (headline);
  * f16: the same filter, FP16 projection storage.
value = voxel-updates per second of the headline form over the whole step
(X*Y*Z*N / step time), in GUPS.

--workload batch16 times the batched form instead: 16 scans of uneven sizes
(volumes 128^3..400^3, 180..720 views, detectors 256^2..1024^2) reconstructed
by one hipGraph launch (BF16 storage, 6-product filter), tiles and view chunks
from the batch autotuner; value = aggregate GUPS of one launch, with per-scan
rel-RMSE against the FP64 oracle on the centre slices.

The roofline object covers the dominant kernel (bp_kernel) timed with events
on the stream it runs on; its algorithmic bytes are 8 B per voxel-update (an
FP32 accumulator read + write per update, the reference's per-view
formulation); its traffic is the measured HBM bytes per launch.  cpu_baseline times the CPU oracle's FP32 backprojection on a
bounded sample of slices on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def rel_rmse(a, ref):
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    m = np.isfinite(a) & np.isfinite(ref)
    return float(np.sqrt(np.mean((a[m] - ref[m]) ** 2)) / np.sqrt(np.mean(ref[m] ** 2)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--views", type=int, default=None, help="override the view count (smoke runs)")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the FP64 oracle slices")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-slices", type=int, default=16)
    ap.add_argument("--acc-slices", type=int, default=4)
    ap.add_argument("--kernel-only", action="store_true",
                    help="profiling run: a 1 GiB copy, then only the backprojection kernel")
    ap.add_argument("--kernel-form", choices=("fp32", "bf16", "f16"), default="bf16",
                    help="storage form of the --kernel-only backprojection")
    ap.add_argument("--kernel-tile", default=None,
                    help="zc,ty tile of the --kernel-only backprojection (default: the autotuner's choice)")
    ap.add_argument("--kernel-filter", action="store_true",
                    help="with --kernel-only: time the headline chain's filter (pre-weighted operand in, "
                         "staged strips out) instead of the backprojection")
    ap.add_argument("--workload", choices=("dental512", "batch16", "offset_dental512"), default="dental512")
    ap.add_argument("--branches", type=int, default=4, help="parallel graph branches (batch16)")
    ap.add_argument("--batch-storage", choices=("bf16", "f16", "f32"), default="bf16",
                    help="projection storage of the batch16 scans")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark measures one GPU (the path shards by z-slabs, one process "
                                   "per GPU, not built here); --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from jenhine_amd import configs
    from jenhine_amd.recon import FDK

    if not torch.cuda.is_available():
        sys.exit("bench.py needs a GPU")
    dev = torch.device("cuda:0")
    if args.workload == "batch16":
        return run_batch(args, dev)
    if args.workload == "offset_dental512":
        return run_offset(args, dev)
    w = configs.DENTAL512
    if args.views:
        w = w.scaled(n_views=args.views)
    t0 = time.time()
    d = w.data()
    log(f"synthetic data {w.n_views} x {w.n_v} x {w.n_u} in {time.time() - t0:.1f} s")
    raw = torch.from_numpy(d["raw"].view(np.int16)).to(dev)
    dark = torch.from_numpy(d["dark"].view(np.int16)).to(dev)
    gain = torch.from_numpy(d["gain"]).to(dev)
    scan, rec = w.scan_params(), w.recon_params()
    forms = {"fp32": FDK(scan, rec, d["ptm"], "fp32", "f32", dev),
             "bf16": FDK(scan, rec, d["ptm"], "mfma6", "bf16", dev),
             "f16": FDK(scan, rec, d["ptm"], "mfma6", "f16", dev)}
    if args.kernel_only:
        # profiling run: a plain copy of known size (scales the HBM counters),
        # then the backprojection kernel of one form `steps` times
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        fdk = forms[args.kernel_form]
        if args.kernel_filter:
            if not fdk.operand_path():
                raise SystemExit("--kernel-filter needs a form whose filter fuses (bf16 or f16)")
            n = raw.shape[0]
            a = fdk.pre.process_operand(raw, fdk.filter, dark, gain)
            q = fdk.filter.process_operand_staged(a, n, fdk.bp, fdk.storage)
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(args.steps):
                fdk.filter.process_operand_staged(a, n, fdk.bp, fdk.storage, out=q)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t) / args.steps * 1e3
            log(f"filter_mfma_kernel (operand in, {fdk.storage} strips out): {ms:.2f} ms/launch")
            print(json.dumps({"kernel_only": True, "filter_ms": ms, "copy_bytes": 2 * src.numel() * 4}))
            return
        q = fdk.bp.stage(fdk.filter.process(fdk.pre.process(raw, dark, gain)), fdk.storage)
        if args.kernel_tile:
            fdk.bp.set_tile(*(int(x) for x in args.kernel_tile.split(",")))
        else:
            fdk.bp.autotune(q, fdk.storage, reps=1)
        vol = fdk.bp.run(q, fdk.storage)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            fdk.bp.run(q, fdk.storage, out=vol)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) / args.steps * 1e3
        log(f"bp_kernel {fdk.storage}: {ms:.2f} ms/launch (tile {fdk.bp.tile()}), "
            f"copy kernel moved {2 * src.numel() * 4} bytes")
        print(json.dumps({"kernel_only": True, "bp_ms": ms, "copy_bytes": 2 * src.numel() * 4}))
        return
    units = w.gups_units

    results = {}
    vols = {}
    tuned = {}
    for name, fdk in forms.items():
        # plan set-up: pick the backprojector's tile shape and schedule on this
        # geometry (not part of the timed step, like any plan creation)
        q0 = fdk.bp.stage(fdk.filter.process(fdk.pre.process(raw, dark, gain)), fdk.storage)
        # per-view cost depends on the view angle: tune on the whole orbit
        table = fdk.bp.autotune(q0, fdk.storage, reps=1)
        tuned[name] = dict(zip(("zc", "ty"), fdk.bp.tile()))
        tuned[name]["fetch"] = fdk.bp.fetch()
        tuned[name]["table_ms"] = {f"{k[0]}_zc{k[1]}_ty{k[2]}": round(v, 3) for k, v in table.items()}
        log(name, "autotune", json.dumps(tuned[name]))
        del q0
        torch.cuda.synchronize()
        for _ in range(args.warmup):
            fdk(raw, dark, gain)
        torch.cuda.synchronize()
        # whole step, and the bp kernel alone (events on the same stream)
        t = time.perf_counter()
        for _ in range(args.steps):
            vol = fdk(raw, dark, gain)
        torch.cuda.synchronize()
        step = (time.perf_counter() - t) / args.steps
        # stages of one more pass, timed with events
        # (filter and staging are one kernel when the plans fuse; then the
        # preprocessing writes the filter's pre-weighted operand)
        fused = fdk.fused()
        operand = fdk.operand_path()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
        ev[0].record()
        p = fdk.pre.process_operand(raw, fdk.filter, dark, gain) if operand else fdk.pre.process(raw, dark, gain)
        ev[1].record()
        if operand:
            f = None
            q = fdk.filter.process_operand_staged(p, raw.shape[0], fdk.bp, fdk.storage)
            ev[2].record()
        elif fused:
            f = None
            q = fdk.filter.process_staged(p, fdk.bp, fdk.storage)
            ev[2].record()
        else:
            f = fdk.filter.process(p)
            ev[2].record()
            q = fdk.bp.stage(f, fdk.storage)
        ev[3].record()
        vol = fdk.bp.run(q, fdk.storage)
        ev[4].record()
        torch.cuda.synchronize()
        st = [ev[i].elapsed_time(ev[i + 1]) for i in range(4)]
        # bp kernel alone, averaged over `steps` launches
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            fdk.bp.run(q, fdk.storage, out=vol)
        e1.record()
        torch.cuda.synchronize()
        bp_ms = e0.elapsed_time(e1) / args.steps
        vols[name] = vol.cpu().numpy()
        results[name] = dict(step_ms=step * 1e3, gups=units / step / 1e9, bp_ms=bp_ms,
                             bp_gups=units / (bp_ms * 1e-3) / 1e9,
                             stage_ms=(dict(pre_operand=st[0], filter_stage=st[1], bp=st[3]) if operand else
                                       dict(pre=st[0], filter_stage=st[1], bp=st[3]) if fused else
                                       dict(pre=st[0], filter=st[1], stage=st[2], bp=st[3])))
        log(name, json.dumps(results[name]))
        del p, f, q

    # accuracy on a few slices against the FP64 oracle, and the CPU baseline
    acc = {}
    cpu = None
    if not (args.no_accuracy and args.no_cpu):
        from oracle import oracle as o
        sc = dict(bright_field=scan.BrightField, hsr=scan.HalfSampleRate, dso=scan.DSO,
                  bh=scan.BeamHardeningParams)
        rc = dict(X=w.X, Y=w.Y, Z=w.Z, pxy=w.pxy, pz=w.pz, cut_start=rec.FilterCutOffStart,
                  cut_end=rec.FilterCutOffEnd, window=rec.FilterWindow)
        nth = min(os.cpu_count() or 1, 16)
        if not args.no_accuracy:
            z0 = w.Z // 2 - args.acc_slices // 2
            t = time.time()
            v64 = o.fdk(d["raw"], d["ptm"], sc, rc, fp64=True, nthreads=nth, z0=z0, nz=args.acc_slices,
                        dark=d["dark"], gain=d["gain"])
            log(f"FP64 oracle slices in {time.time() - t:.1f} s")
            for name in vols:
                g = vols[name][z0:z0 + args.acc_slices]
                acc[name] = dict(rel_rmse_vs_fp64=rel_rmse(g, v64),
                                 max_abs_vs_fp64=float(np.nanmax(np.abs(g.astype(np.float64) - v64))),
                                 slices=[z0, z0 + args.acc_slices])
            log("accuracy", json.dumps(acc))
        if not args.no_cpu:
            # FP32-ref oracle: preprocess + filter all views, then time its
            # backprojection of cpu_slices centre slices
            p32 = o.preprocess(d["raw"], bright_field=sc["bright_field"], bh=sc["bh"], dark=d["dark"],
                               gain=d["gain"], nthreads=nth)
            H = o.filter_design(w.n_u, sc["hsr"], rc["cut_start"], rc["cut_end"], window=rc["window"])
            src, blk = o.preweight_schedule(w.n_views)
            pws = np.stack([o.prefilter_weight(d["ptm"][s], w.n_u, w.n_v, sc["dso"]) for s in src])
            f32 = o.filter_rows(p32, H, pws, blk, nthreads=nth)
            del p32
            z0 = w.Z // 2 - args.cpu_slices // 2
            t = time.perf_counter()
            v32 = o.backproject(f32, d["ptm"], (w.Z, w.Y, w.X), w.pxy, w.pz, z0=z0, nz=args.cpu_slices,
                                nthreads=nth)
            dt = time.perf_counter() - t
            n_upd = w.X * w.Y * args.cpu_slices * w.n_views
            cpu = dict(value=n_upd / dt / 1e9, unit="GUPS", cores=nth, kind="port",
                       sample=f"CPU oracle FP32-ref backprojection of {args.cpu_slices} centre slices "
                              f"(512 x 512 x {args.cpu_slices}) x {w.n_views} views, {nth} threads, "
                              f"{dt:.1f} s")
            acc["fp32_vs_fp32ref_rel_rmse"] = rel_rmse(vols["fp32"][z0:z0 + args.cpu_slices], v32)
            log("cpu", json.dumps(cpu), "fp32 vs FP32-ref", acc["fp32_vs_fp32ref_rel_rmse"])

    head = results["bf16"]
    bp_bytes = 8.0 * units
    achieved = bp_bytes / (head["bp_ms"] * 1e-3) / 1e9
    # HBM bytes per launch of this kernel from the committed rocprofv3 --pmc
    # summary (FETCH_SIZE and WRITE_SIZE, each scaled by what the counter
    # reads for a 1 GiB device copy); null when the summary is absent
    traffic = None
    tf = os.path.join(ROOT, "profiles", "bp_dental512_bf16_traffic.json")
    if os.path.exists(tf):
        with open(tf) as fh:
            traffic = json.load(fh)["hbm_bytes"]
    line = {
        "metric": "backprojection GUPS at 512^3 x 600 views; FP32 vs BF16/MFMA path at <=1e-3 rel-RMSE",
        "value": head["gups"],
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["step_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16-storage/f32-accumulate",
        "data": "synthetic (seeded ellipsoid phantom, analytic line integrals, Poisson counts)",
        "config": {"workload": f"dental512: FDK {w.X}x{w.Y}x{w.Z} from {w.n_views} views of {w.n_u}x{w.n_v} u16",
                   "global_batch": 1, "views": w.n_views, "detector": [w.n_u, w.n_v],
                   "volume": [w.X, w.Y, w.Z], "parallelism": "single-gpu"},
        "forms": results,
        "autotune": tuned,
        "accuracy": acc,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved / 8000.0, "traffic": traffic,
                     "kernel": "bp_kernel (bf16 strip storage)", "bytes_per_update": 8,
                     "note": "achieved counts the per-view formulation's 8 B/voxel-update (FP32 accumulator "
                             "read + write); frac > 1 means the register-accumulating kernel beats that roof. "
                             "traffic: measured HBM bytes per launch (FETCH_SIZE + WRITE_SIZE). Measured "
                             "limiter: the vector-memory path's per-instruction rate (21 L1 accesses per "
                             "16-byte gather, ~27 cycles per gather per CU; VALU 58 % busy, "
                             "profiles/bp_dental512_pmc.md)"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def run_offset(args, dev):
    """dental512 on a half-fan detector through the composite chain
    (preprocess -> composite -> 2W -> W mirrored filter -> stage -> BP)."""
    import torch
    from jenhine_amd import configs
    from jenhine_amd.recon import FDK

    w = configs.OFFSET_DENTAL512
    if args.views:
        w = w.scaled(n_views=args.views)
    t0 = time.time()
    d = w.data()
    log(f"synthetic half-fan data {w.n_views} x {w.n_v} x {w.n_u} in {time.time() - t0:.1f} s")
    raw = torch.from_numpy(d["raw"].view(np.int16)).to(dev)
    dark = torch.from_numpy(d["dark"].view(np.int16)).to(dev)
    gain = torch.from_numpy(d["gain"]).to(dev)
    scan, rec = w.scan_params(), w.recon_params()
    units = w.gups_units
    results, vols = {}, {}
    for name, (mode, storage) in {"fp32": ("fp32", "f32"), "bf16": ("mfma6", "bf16")}.items():
        fdk = FDK(scan, rec, d["ptm"], mode, storage, dev, composite=True)
        _, st0 = fdk(raw, dark, gain, return_stages=True)
        fdk.bp.autotune(st0["q"], storage, reps=1)
        del st0
        for _ in range(args.warmup):
            fdk(raw, dark, gain)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            vol = fdk(raw, dark, gain)
        torch.cuda.synchronize()
        step = (time.perf_counter() - t) / args.steps
        fused = fdk.fused()
        operand = fdk.composite_operand_path()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
        ev[0].record()
        p = fdk.pre.process(raw, dark, gain)
        ev[1].record()
        if operand:
            c = fdk.composite.process_operand(p, fdk.filter)
        else:
            c = fdk.composite.process(p)
        ev[2].record()
        if operand:
            f = None
            q = fdk.filter.process_operand_staged(c, w.n_views, fdk.bp, storage)
            ev[3].record()
        elif fused:
            f = None
            q = fdk.filter.process_staged(c, fdk.bp, storage)
            ev[3].record()
        else:
            f = fdk.filter.process(c)
            ev[3].record()
            q = fdk.bp.stage(f, storage)
        ev[4].record()
        vol = fdk.bp.run(q, storage)
        ev[5].record()
        torch.cuda.synchronize()
        stt = [ev[i].elapsed_time(ev[i + 1]) for i in range(5)]
        vols[name] = vol.cpu().numpy()
        results[name] = dict(step_ms=step * 1e3, gups=units / step / 1e9,
                             stage_ms=(dict(pre=stt[0], composite_operand=stt[1], filter_stage=stt[2], bp=stt[4])
                                       if operand else
                                       dict(pre=stt[0], composite=stt[1], filter_stage=stt[2], bp=stt[4]) if fused
                                       else dict(pre=stt[0], composite=stt[1], filter=stt[2], stage=stt[3],
                                                 bp=stt[4])))
        log(name, json.dumps(results[name]))
        del p, c, f, q, fdk
    acc = {}
    if not args.no_accuracy:
        from oracle import oracle as o
        nth = min(os.cpu_count() or 1, 16)
        sc = dict(bright_field=scan.BrightField, hsr=scan.HalfSampleRate, dso=scan.DSO, bh=scan.BeamHardeningParams)
        rc = dict(X=w.X, Y=w.Y, Z=w.Z, pxy=w.pxy, pz=w.pz, cut_start=rec.FilterCutOffStart,
                  cut_end=rec.FilterCutOffEnd, window=rec.FilterWindow)
        z0 = w.Z // 2 - args.acc_slices // 2
        t = time.time()
        v64 = o.fdk(d["raw"], d["ptm"], sc, rc, fp64=True, nthreads=nth, z0=z0, nz=args.acc_slices,
                    dark=d["dark"], gain=d["gain"], composite_views=True)
        log(f"FP64 oracle slices (composite chain) in {time.time() - t:.1f} s")
        for name in vols:
            g = vols[name][z0:z0 + args.acc_slices]
            acc[name] = dict(rel_rmse_vs_fp64=rel_rmse(g, v64),
                             max_abs_vs_fp64=float(np.nanmax(np.abs(g.astype(np.float64) - v64))),
                             slices=[z0, z0 + args.acc_slices])
        log("accuracy", json.dumps(acc))
    line = {
        "metric": "half-fan FDK GUPS at 512^3 x 600 views (composite chain)",
        "value": results["bf16"]["gups"], "unit": "GUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": results["bf16"]["step_ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16-storage/f32-accumulate",
        "data": "synthetic (seeded ellipsoid phantom, analytic line integrals, Poisson counts)",
        "config": {"workload": f"offset_dental512: composite FDK {w.X}x{w.Y}x{w.Z} from {w.n_views} views of "
                               f"{w.n_u}x{w.n_v} u16, central ray at column {w.n_u / 2 - 1 - w.u_shift:.0f}",
                   "parallelism": "single-gpu"},
        "forms": results, "accuracy": acc,
    }
    print(json.dumps(line), flush=True)


def run_batch(args, dev):
    """16 uneven scans, one hipGraph launch per step."""
    import torch
    from jenhine_amd import configs
    from jenhine_amd.recon import FDK, BatchGraph

    t0 = time.time()
    entries = []
    for w in configs.batch16():
        d = w.data()
        del d["lineint"]
        e = dict(w=w, d=d, raw=torch.from_numpy(d["raw"].view(np.int16)).to(dev),
                 dark=torch.from_numpy(d["dark"].view(np.int16)).to(dev), gain=torch.from_numpy(d["gain"]).to(dev),
                 vol=torch.empty((w.Z, w.Y, w.X), dtype=torch.float32, device=dev))
        e["fdk"] = FDK(w.scan_params(), w.recon_params(), d["ptm"], "fp32" if args.batch_storage == "f32" else "mfma6",
                       args.batch_storage, dev)
        entries.append(e)
    units = sum(e["w"].gups_units for e in entries)
    log(f"16 scans generated in {time.time() - t0:.1f} s; {units / 1e9:.1f} G voxel-updates per launch")
    g = BatchGraph(device=dev)
    for e in entries:
        g.add(e["fdk"], e["raw"], e["vol"], e["dark"], e["gain"])
    t = time.time()
    g.autotune(reps=1)
    log(f"autotune {time.time() - t:.1f} s")
    g.instantiate(branches=args.branches)
    for _ in range(args.warmup):
        g.launch()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        g.launch()
    torch.cuda.synchronize()
    step = (time.perf_counter() - t) / args.steps
    log(f"graph launch {step * 1e3:.1f} ms, {units / step / 1e9:.1f} GUPS")
    # the same scans one after another through the one-shot chain (no graph)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for e in entries:
        e["fdk"](e["raw"], e["dark"], e["gain"])
    torch.cuda.synchronize()
    seq = time.perf_counter() - t
    log(f"sequential one-shot FDK {seq * 1e3:.1f} ms, {units / seq / 1e9:.1f} GUPS")

    per_scan = []
    if not args.no_accuracy:
        from oracle import oracle as o
        nth = min(os.cpu_count() or 1, 16)
        t = time.time()
        for i, e in enumerate(entries):
            w, d = e["w"], e["d"]
            scan, rec = w.scan_params(), w.recon_params()
            sc = dict(bright_field=scan.BrightField, hsr=scan.HalfSampleRate, dso=scan.DSO,
                      bh=scan.BeamHardeningParams)
            rc = dict(X=w.X, Y=w.Y, Z=w.Z, pxy=w.pxy, pz=w.pz, cut_start=rec.FilterCutOffStart,
                      cut_end=rec.FilterCutOffEnd, window=rec.FilterWindow)
            z0 = w.Z // 2 - 1
            v64 = o.fdk(d["raw"], d["ptm"], sc, rc, fp64=True, nthreads=nth, z0=z0, nz=2, dark=d["dark"],
                        gain=d["gain"])
            got = e["vol"][z0:z0 + 2].cpu().numpy()
            chunk, zc, ty = g.tuning(i)
            per_scan.append(dict(volume=w.X, views=w.n_views, detector=w.n_u, chunk=chunk, tile=[zc, ty],
                                 rel_rmse_vs_fp64=rel_rmse(got, v64),
                                 max_abs_vs_fp64=float(np.nanmax(np.abs(got.astype(np.float64) - v64)))))
            log("scan", i, json.dumps(per_scan[-1]))
        log(f"FP64 oracle slices in {time.time() - t:.1f} s")
    line = {
        "metric": "batched FDK: aggregate GUPS of 16 uneven scans in one hipGraph launch",
        "value": units / step / 1e9,
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"{args.batch_storage}-storage/f32-accumulate",
        "data": "synthetic (seeded ellipsoid phantoms, analytic line integrals, Poisson counts)",
        "config": {"workload": "batch16: 16 scans, volumes 128^3..400^3, 180..720 views, detectors 256^2..1024^2",
                   "scans": len(entries), "voxel_updates": units, "branches": args.branches,
                   "parallelism": "single-gpu"},
        "sequential_oneshot_ms": seq * 1e3,
        "per_scan": per_scan,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
