#!/usr/bin/env python3
"""Benchmark of the FDK back end on one MI355X.

Workload (dental512): 512^3 FP32 volume at 0.15 mm from 600 views of 768 x 768
uint16 at 0.2 mm, SOD 350 mm, SDD 600 mm, Shepp-Logan window, cubic beam
hardening, seeded 40-ellipsoid + tooth phantom, per-pixel gain and dark.

One step = the whole reconstruction on the device with the raw counts already
resident in HBM: preprocess -> row filter -> stage -> backprojection ->
normalised volume.  Three forms are timed:
  * fp32: FP32 Toeplitz filter on MFMA, FP32 projection storage;
  * bf16: row-scaled hi/lo FP16 Toeplitz filter on MFMA (3 products, FP32
    accumulation; --filter-form selects the split-BF16 forms), BF16
    projection storage (headline);
  * f16: the same filter, FP16 projection storage.
value = voxel-updates per second of the headline form over the whole step
(X*Y*Z*N / step time), in GUPS.

--workload batch16 times the batched form instead: 16 scans of uneven sizes
(volumes 128^3..400^3, 180..720 views, detectors 256^2..1024^2) reconstructed
by one hipGraph launch (FP16 storage with its range exponent, the headline
filter form), tiles and view chunks from the batch autotuner; value =
aggregate GUPS of one launch, with per-scan rel-RMSE against the FP64 oracle
on the centre slices.

--workload adjoint256 times the differentiation form: the forward projector
A and the un-normalised backprojector A^T on 256^3 from 360 views of 512^2,
with the dot-product test of the pair in the same run.

--workload pipeline_dental512 times dental512 through the C++ agent graph
(jh_fdk_pipeline) with the raw counts and the slices in host memory: PCIe
transfers included.

The roofline object covers the dominant kernel (bp_kernel), timed with events
on the stream it runs on.  Its roof is the lowest of three, each from the
kernel's own schedule and the counters committed under profiles/ (rocprofv3
--pmc of the same kernel):
  * HBM: the bytes the schedule must move at least once (the volume written
    once, every staged view read once) at 8 TB/s;
  * VALU: the kernel's VALU instructions per voxel-update by counter class
    (SQ_INSTS_VALU_*), each at the clocks per wave instruction and SIMD the
    valu_rate probe measured for its type (profiles/valu_rate.json: FP32
    add/FMA ~2.2, permutes, BF16 dots, integer, conversions ~4.1), over 1024
    SIMDs at the published 2.4 GHz peak clock;
  * vector-memory issue: the bytes each update requests through the texture
    path at 64 B per clock per CU (a 64-lane 16-byte gather issues in 16
    clocks), or from LDS at 128 B per clock for the LDS-staged fetch;
frac = achieved GUPS / the lowest roof.  The reference's per-view formulation
(an FP32 accumulator read + write per update, 8 B) is kept as a separate
field, per_view_rmw_8B.  cpu_baseline times the CPU oracle (the reference
algorithm in portable C, FP32) on the host cores: preprocessing and filtering
of every view and the backprojection of a bounded slab of slices.
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
    ap.add_argument("--kernel-fetch", choices=("gather", "lds"), default=None,
                    help="sample fetch of the --kernel-only backprojection (default: the autotuner's choice)")
    ap.add_argument("--kernel-filter", action="store_true",
                    help="with --kernel-only: time the headline chain's filter (pre-weighted operand in, "
                         "staged strips out) instead of the backprojection")
    ap.add_argument("--filter-form", choices=("mfma", "mfma6", "mfma16"), default="mfma16",
                    help="MFMA filter of the bf16 and f16 chains: hi/lo BF16 (3 products), hi/mid/lo BF16 "
                         "(6 products) or row-scaled hi/lo FP16 (3 products)")
    ap.add_argument("--workload", choices=("dental512", "batch16", "offset_dental512", "adjoint256",
                                           "pipeline_dental512"),
                    default="dental512")
    ap.add_argument("--chunk", type=int, default=64, help="views per chunk of the agent graph (pipeline_dental512)")
    ap.add_argument("--branches", type=int, default=4, help="parallel graph branches (batch16)")
    ap.add_argument("--summarize-counters", metavar="DIR", default=None,
                    help="no GPU work: read the rocprofv3 --pmc CSV files under DIR (kernel-only runs of one "
                         "backprojection tile) and print the counter summary the roofline reads")
    ap.add_argument("--summarize-valu", metavar="DIR", default=None,
                    help="no GPU work: read the rocprofv3 --pmc CSV of a valu_rate run (SQ_INSTS_VALU and "
                         "GRBM_GUI_ACTIVE) and write the clocks per wave instruction of each stream to "
                         "profiles/valu_rate.json")
    ap.add_argument("--tile", default=None, help="zc,ty of the profiled kernel (with --summarize-counters)")
    ap.add_argument("--fetch", choices=("gather", "lds"), default="gather",
                    help="fetch form of the profiled kernel (with --summarize-counters)")
    ap.add_argument("--form", choices=("fp32", "bf16", "f16"), default="bf16",
                    help="storage form of the profiled kernel (with --summarize-counters)")
    ap.add_argument("--counters-out", default=None,
                    help="with --summarize-counters: merge the summary into this JSON file (e.g. "
                         "profiles/bp_dental512_counters.json)")
    ap.add_argument("--pipeline-args", metavar="DIR", default=None,
                    help="no GPU work: write the pipeline_dental512 input files to DIR and print the "
                         "jh_fdk_pipeline argument list (for running the app under a profiler)")
    ap.add_argument("--summarize-timeline", metavar="DIR", default=None,
                    help="no GPU work: read a rocprofv3 --runtime-trace CSV output under DIR (one jh_fdk_pipeline "
                         "run) and print its kernel, copy and API summary and a 5 ms timeline")
    ap.add_argument("--batch-storage", choices=("bf16", "f16", "f32"), default="f16",
                    help="projection storage of the batch16 scans (FP16 with its range exponent: BF16 storage "
                         "misses 1e-3 on noise-dominated Ram-Lak scans)")
    args = ap.parse_args()
    if args.summarize_counters:
        return summarize_counters(args.summarize_counters, args.tile, args.fetch, args.form, args.counters_out)
    if args.summarize_valu:
        return summarize_valu(args.summarize_valu)
    if args.pipeline_args:
        print(" ".join(pipeline_command(args, args.pipeline_args)[1:]))
        return
    if args.summarize_timeline:
        return summarize_timeline(args.summarize_timeline)
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
    if args.workload == "adjoint256":
        return run_adjoint(args, dev)
    if args.workload == "pipeline_dental512":
        return run_pipeline(args)
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
             "bf16": FDK(scan, rec, d["ptm"], args.filter_form, "bf16", dev),
             "f16": FDK(scan, rec, d["ptm"], args.filter_form, "f16", dev)}
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
            if args.kernel_fetch:
                fdk.bp.set_fetch(args.kernel_fetch)
        else:
            fdk.bp.autotune(q, fdk.storage, reps=1)
        vol = fdk.bp.run(q, fdk.storage)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            fdk.bp.run(q, fdk.storage, out=vol)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) / args.steps * 1e3
        log(f"bp_kernel {fdk.storage}: {ms:.2f} ms/launch (tile {fdk.bp.tile()}, {fdk.bp.fetch()} set, "
            f"{fdk.bp.fetch_in_use(fdk.storage)} in use), copy kernel moved {2 * src.numel() * 4} bytes")
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
            cpu, v32, z0 = cpu_baseline(o, d, w, sc, rc, nth, args.cpu_slices)
            acc["fp32_vs_fp32ref_rel_rmse"] = rel_rmse(vols["fp32"][z0:z0 + args.cpu_slices], v32)
            log("cpu", json.dumps(cpu), "fp32 vs FP32-ref", acc["fp32_vs_fp32ref_rel_rmse"])

    head = results["bf16"]
    stage_all = forms["bf16"].bp.stage_bytes("bf16", w.n_views)
    roof = bp_roofline(head["bp_gups"], units, w, stage_all, tuned["bf16"])
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
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


# Published MI355X peaks used by the roofline: HBM 8 TB/s; FP32 vector 157.3
# TFLOP/s = 256 CUs x 4 SIMDs x 2.4 GHz x 32 FMA lanes per clock, i.e. one
# wave64 VALU instruction per SIMD every 2 clocks (1.229e12 per second).
HBM_PEAK = 8.0e12
CLOCK = 2.4e9
CUS = 256
SIMDS = 1024
VALU_RATE_FILE = os.path.join("profiles", "valu_rate.json")
# VALU instruction classes of the SQ_INSTS_VALU_* counters -> the valu_rate
# stream whose clocks per wave instruction they take; the rest of the
# instructions (v_perm_b32, v_dot2_f32_bf16, selects, compares: no class)
# take those of v_perm_b32 (v_dot2 measures the same)
VALU_CLASSES = {"valu_fma_f32": "fma_f32", "valu_add_f32": "add_f32", "valu_mul_f32": "fma_f32",
                "valu_trans_f32": "perm_b32", "valu_cvt": "cvt", "valu_int32": "xad_u32"}
TA_BYTES_PER_CLOCK = 64.0    # per CU: four lanes of a 16-byte gather per clock
LDS_BYTES_PER_CLOCK = 128.0  # per CU: LDS read bandwidth
COUNTERS_FILE = os.path.join("profiles", "bp_dental512_counters.json")


def counter_key(form, fetch, zc, ty):
    return f"{form},{fetch},{zc},{ty}"


def bp_roofline(gups, units, w, stage_bytes, tuned, form="bf16"):
    """The backprojector's roofline from its schedule and the committed
    counters of the same kernel (profiles/bp_dental512_counters.json, keyed by
    storage form, fetch form and tile)."""
    nvox = w.X * w.Y * w.Z
    fetch = tuned["fetch"]
    # compulsory bytes of the schedule: one launch writes the volume once and
    # reads every view's staged frame once (register accumulation over all
    # views of the launch: no accumulator traffic)
    b_sched = 4.0 * nvox + stage_bytes
    gups_hbm = HBM_PEAK / (b_sched / units) / 1e9
    # bytes each update reads as a sample window: one 16-byte window per pair
    # of z samples (column strips) -> 8 B, through the texture path (gather
    # form) or from LDS (LDS-staged form, whose copies into LDS pass the
    # texture path once per footprint pixel and workgroup)
    req_per_update = 8.0
    rate = TA_BYTES_PER_CLOCK if fetch == "gather" else LDS_BYTES_PER_CLOCK
    gups_win = CUS * CLOCK * rate / req_per_update / 1e9
    win = "vmem_issue" if fetch == "gather" else "lds_read"
    roof = {"kernel": f"{'bp_kernel' if fetch == 'gather' else 'bp_strip_lds_kernel'} ({form.upper()} column strips, "
                      f"{fetch} fetch, zc {tuned['zc']}, {tuned['ty']} sampling waves per workgroup)",
            "achieved": gups, "unit": "GUPS",
            "compulsory_bytes": b_sched, "gups_hbm_sched": gups_hbm,
            "window_bytes_per_update": req_per_update, f"gups_{win}": gups_win,
            "per_view_rmw_8B": {"bytes_per_update": 8, "achieved_GBps": 8.0 * gups,
                                  "peak_GBps": HBM_PEAK / 1e9, "frac": 8.0 * gups * 1e9 / HBM_PEAK,
                                  "note": "the reference's per-view formulation (FP32 accumulator read + write "
                                          "per update); this kernel accumulates in registers and never moves "
                                          "those bytes, so this is not a roof it runs against"}}
    cf = os.path.join(ROOT, COUNTERS_FILE)
    c = None
    if os.path.exists(cf):
        with open(cf) as fh:
            c = json.load(fh)
    key = counter_key(form, fetch, tuned["zc"], tuned["ty"])
    allc = (c or {}).get("kernels", {})
    c = allc.get(key)
    gups_valu = None
    vc = valu_cycles(c) if c is not None else None
    if vc:
        valu_per_update = c["valu_insts"] * 64.0 / units
        # SIMD clocks of VALU issue per voxel-update, at the published peak
        # clock: the class model, calibrated so that no profiled kernel issues
        # faster than it (the model adds each instruction's clocks as measured
        # alone; some kernels beat that sum by a few percent)
        cal = valu_calibration(allc)
        cyc_per_update = vc / units / cal
        gups_valu = SIMDS * CLOCK / cyc_per_update / 1e9
        roof.update(valu_lane_instr_per_update=valu_per_update, valu_simd_clocks_per_update=cyc_per_update,
                    gups_valu=gups_valu, valu_calibration=cal,
                    valu_model=f"SQ_INSTS_VALU_* classes at the clocks per instruction of {VALU_RATE_FILE}, "
                               f"divided by {cal:.3f}: the largest model/measured VALU time of the profiled "
                               f"kernels in {COUNTERS_FILE}",
                    traffic=c["hbm_bytes"], traffic_over_compulsory=c["hbm_bytes"] / b_sched,
                    l2_hit=c["l2_hit"], valu_busy=c["valu_busy"], ta_busy=c["ta_busy"],
                    counters=f"{COUNTERS_FILE} [{key}]")
    else:
        roof.update(traffic=c.get("hbm_bytes") if c else None,
                    counters=f"no VALU class counters for {key}: VALU roof not evaluated")
    roofs = {"hbm": gups_hbm, win: gups_win}
    if gups_valu:
        roofs["valu"] = gups_valu
    bound = min(roofs, key=roofs.get)
    roof.update(bound=bound, peak=roofs[bound], frac=gups / roofs[bound], roofs=roofs)
    return roof


def valu_rates():
    path = os.path.join(ROOT, VALU_RATE_FILE)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh)["cycles_per_wave_instr"]


def valu_calibration(rows):
    """Largest ratio of modelled VALU time (valu_cycles) to the launch's
    measured cycles over the profiled kernels, at least 1: the model's clocks
    divided by it bound every measured kernel."""
    cal = 1.0
    for r in rows.values():
        vc = valu_cycles(r)
        cyc = r.get("gui_active_cycles_per_xcd")
        if vc and cyc:
            cal = max(cal, vc / (SIMDS * cyc))
    return cal


def valu_cycles(c):
    """SIMD clocks of VALU issue of one launch (summed over the SIMDs): each
    SQ_INSTS_VALU_* class at the clocks per wave instruction valu_rate measured
    for its stream, the unclassified rest at v_perm_b32's; None without the
    class counters or the rates."""
    r = valu_rates()
    if r is None or not c.get("valu_insts") or any(c.get(k) is None for k in VALU_CLASSES):
        return None
    classified = sum(c[k] for k in VALU_CLASSES)
    return sum(c[k] * r[VALU_CLASSES[k]] for k in VALU_CLASSES) + (c["valu_insts"] - classified) * r["perm_b32"]


def summarize_valu(folder):
    """Clocks per wave instruction and SIMD of each valu_rate stream from a
    rocprofv3 --pmc run with SQ_INSTS_VALU and GRBM_GUI_ACTIVE (the second,
    warm launch of each stream); written to profiles/valu_rate.json."""
    import csv
    import glob
    names = ["fma_f32", "perm_b32", "dot2_f32_bf16", "xad_u32", "add_f32", "cvt", "fract_f32", "med3_f32",
             "mad_u32_u24", "cmp_f32"]
    per = {}  # stream -> [(dispatch, instr, gui cycles summed over XCDs, ns)]
    for path in glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as fh:
            for row in csv.DictReader(fh):
                kn = row["Kernel_Name"]
                if not kn.startswith("void stream<"):
                    continue
                op = int(kn[len("void stream<"):].split(">")[0])
                d = per.setdefault(names[op], {}).setdefault(int(row["Dispatch_Id"]), {})
                d[row["Counter_Name"]] = d.get(row["Counter_Name"], 0.0) + float(row["Counter_Value"])
                d["ns"] = int(row["End_Timestamp"]) - int(row["Start_Timestamp"])
    out = {"source": "rocprofv3 --pmc SQ_INSTS_VALU GRBM_GUI_ACTIVE of jenhine_amd/build/valu_rate "
                     "(profiles/valu_rate.md)", "cycles_per_wave_instr": {}, "clock_ghz": {}}
    for name in names:
        if name not in per:
            continue
        warm = per[name][max(per[name])]  # the later launch of the stream
        cyc = warm["GRBM_GUI_ACTIVE"] / 8.0  # per XCD
        out["cycles_per_wave_instr"][name] = cyc / (warm["SQ_INSTS_VALU"] / SIMDS)
        out["clock_ghz"][name] = cyc / warm["ns"]
    print(json.dumps(out, indent=1))
    with open(os.path.join(ROOT, VALU_RATE_FILE), "w") as fh:
        json.dump(out, fh, indent=1)
        fh.write("\n")


def summarize_counters(folder, tile, fetch, form, out_file):
    """Per-launch counters of the backprojection kernel from rocprofv3 --pmc
    CSV output (one counter set per run), with the HBM byte counters scaled by
    what they read for the 1 GiB device copy that --kernel-only runs first;
    merged into out_file under "form,fetch,zc,ty"."""
    import csv
    import glob
    vals = {}   # (kernel, counter) -> {dispatch: value}
    span = {}   # kernel -> {dispatch: ns}
    for path in glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as fh:
            for row in csv.DictReader(fh):
                d = (path, row.get("Dispatch_Id"))
                k = (row["Kernel_Name"], row["Counter_Name"])
                vals.setdefault(k, {}).setdefault(d, 0.0)
                vals[k][d] += float(row["Counter_Value"])
                span.setdefault(row["Kernel_Name"], {})[d] = int(row["End_Timestamp"]) - int(row["Start_Timestamp"])

    def per_dispatch(pred, counter):
        xs = [v for (kn, cn), dd in vals.items() if pred(kn) and cn == counter for v in dd.values()]
        return sum(xs) / len(xs) if xs else None
    is_bp = lambda kn: "bp_kernel" in kn or "bp_strip_lds_kernel" in kn
    is_copy = lambda kn: "copy" in kn.lower() and "bp_" not in kn
    gib = float(1 << 30)
    def largest(pred, counter):  # the 1 GiB copy, not the small uploads of the same copy kernel
        xs = [v for (kn, cn), dd in vals.items() if pred(kn) and cn == counter for v in dd.values()]
        return max(xs) if xs else None
    fetch_copy, write_copy = largest(is_copy, "FETCH_SIZE"), largest(is_copy, "WRITE_SIZE")
    fetch_scale = gib / (fetch_copy * 1024.0) if fetch_copy else None
    write_scale = gib / (write_copy * 1024.0) if write_copy else None
    fetch_kib, write_kib = per_dispatch(is_bp, "FETCH_SIZE"), per_dispatch(is_bp, "WRITE_SIZE")
    hit, miss = per_dispatch(is_bp, "TCC_HIT_sum"), per_dispatch(is_bp, "TCC_MISS_sum")
    gui = per_dispatch(is_bp, "GRBM_GUI_ACTIVE")
    names = sorted({kn for kn in span if is_bp(kn)})
    ns = [t for kn in names for t in span[kn].values()]
    zc, ty = (int(x) for x in tile.split(","))
    out = {"kernel": names, "form": form, "fetch": fetch, "zc": zc, "ty": ty,
           "ms": sum(ns) / len(ns) / 1e6 if ns else None,
           "fetch_kib": fetch_kib, "write_kib": write_kib, "fetch_scale": fetch_scale, "write_scale": write_scale,
           "hbm_bytes": (fetch_kib * 1024.0 * fetch_scale + write_kib * 1024.0 * write_scale)
           if None not in (fetch_kib, write_kib, fetch_scale, write_scale) else None,
           "l2_hit": hit / (hit + miss) if hit is not None and miss else None,
           "gui_active_cycles_per_xcd": gui / 8.0 if gui else None}
    for name, counter in (("valu_insts", "SQ_INSTS_VALU"), ("vmem_rd_insts", "SQ_INSTS_VMEM_RD"),
                          ("lds_insts", "SQ_INSTS_LDS"), ("waves", "SQ_WAVES"), ("sq_busy_cycles", "SQ_BUSY_CYCLES"),
                          ("sq_wave_cycles", "SQ_WAVE_CYCLES"), ("sq_active_inst_valu", "SQ_ACTIVE_INST_VALU"),
                          ("sq_wait_inst_any", "SQ_WAIT_INST_ANY"), ("sq_active_inst_lds", "SQ_ACTIVE_INST_LDS"),
                          ("sq_wait_inst_lds", "SQ_WAIT_INST_LDS"), ("lds_bank_conflict", "SQ_LDS_BANK_CONFLICT"),
                          ("ta_busy_sum", "TA_TA_BUSY_sum"), ("ta_buffer_read_wavefronts", "TA_BUFFER_READ_WAVEFRONTS_sum"),
                          ("td_busy_sum", "TD_TD_BUSY_sum"), ("tcp_accesses", "TCP_TOTAL_CACHE_ACCESSES_sum"),
                          ("tcp_tcc_read_req", "TCP_TCC_READ_REQ_sum"), ("valu_fma_f32", "SQ_INSTS_VALU_FMA_F32"),
                          ("valu_add_f32", "SQ_INSTS_VALU_ADD_F32"), ("valu_mul_f32", "SQ_INSTS_VALU_MUL_F32"),
                          ("valu_trans_f32", "SQ_INSTS_VALU_TRANS_F32"), ("valu_cvt", "SQ_INSTS_VALU_CVT"),
                          ("valu_int32", "SQ_INSTS_VALU_INT32")):
        out[name] = per_dispatch(is_bp, counter)
    cyc = out["gui_active_cycles_per_xcd"]
    if cyc:
        out["clock_ghz"] = cyc / (out["ms"] * 1e6) if out["ms"] else None
        # the VALU instruction mix at the measured clocks per instruction
        vc = valu_cycles(out)
        out["valu_cycles_per_simd"] = vc / SIMDS if vc else None
        out["valu_busy"] = vc / (SIMDS * cyc) if vc else None
        out["ta_busy"] = out["ta_busy_sum"] / (CUS * cyc) if out["ta_busy_sum"] else None
        out["lds_bank_conflict_share"] = (out["lds_bank_conflict"] / (CUS * cyc)
                                          if out["lds_bank_conflict"] is not None else None)
        out["vmem_cycles_per_insn_per_cu"] = (CUS * cyc / out["vmem_rd_insts"] if out["vmem_rd_insts"] else None)
    else:
        out["clock_ghz"] = out["valu_busy"] = out["ta_busy"] = None
    if out["vmem_rd_insts"] and out["tcp_accesses"]:
        out["tcp_accesses_per_vmem"] = out["tcp_accesses"] / out["vmem_rd_insts"]
    print(json.dumps(out, indent=1))
    if out_file:
        path = out_file if os.path.isabs(out_file) else os.path.join(ROOT, out_file)
        allc = {}
        if os.path.exists(path):
            with open(path) as fh:
                allc = json.load(fh)
        allc.setdefault("source", "rocprofv3 --pmc passes of bench.py --kernel-only (profiles/bp_dental512_pmc.md)")
        allc.setdefault("kernels", {})[counter_key(form, fetch, zc, ty)] = out
        with open(path, "w") as fh:
            json.dump(allc, fh, indent=1)
            fh.write("\n")


def host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return model, os.cpu_count(), avail


def cpu_baseline(o, d, w, sc, rc, nth, nslices):
    """The CPU oracle (FP32-ref, the reference's op order in portable C,
    threaded like the reference: filter rows in tasks, slices in tasks of 10)
    on the host: preprocessing and filtering of every view, then the
    backprojection of nslices centre slices; the full-volume end-to-end time
    is extrapolated from the slab (backprojection cost is linear in slices)."""
    model, ncpu, avail = host_cpu()
    t0 = time.perf_counter()
    p32 = o.preprocess(d["raw"], bright_field=sc["bright_field"], bh=sc["bh"], dark=d["dark"], gain=d["gain"],
                       nthreads=nth)
    t1 = time.perf_counter()
    H = o.filter_design(w.n_u, sc["hsr"], rc["cut_start"], rc["cut_end"], window=rc["window"])
    src, blk = o.preweight_schedule(w.n_views)
    pws = np.stack([o.prefilter_weight(d["ptm"][s], w.n_u, w.n_v, sc["dso"]) for s in src])
    f32 = o.filter_rows(p32, H, pws, blk, nthreads=nth)
    t2 = time.perf_counter()
    del p32
    z0 = w.Z // 2 - nslices // 2
    v32 = o.backproject(f32, d["ptm"], (w.Z, w.Y, w.X), w.pxy, w.pz, z0=z0, nz=nslices, nthreads=nth)
    t3 = time.perf_counter()
    t_pre, t_filt, t_bp = t1 - t0, t2 - t1, t3 - t2
    e2e_full = t_pre + t_filt + t_bp * (w.Z / nslices)
    units = w.X * w.Y * w.Z * w.n_views
    cpu = dict(value=units / e2e_full / 1e9, unit="GUPS", cores=nth, kind="port",
               host={"cpu_model": model, "logical_cpus": ncpu, "cpus_available": avail},
               seconds={"preprocess_all_views": t_pre, "filter_all_views": t_filt,
                        f"backproject_{nslices}_slices": t_bp, "end_to_end_extrapolated": e2e_full},
               bp_only_gups=w.X * w.Y * nslices * w.n_views / t_bp / 1e9,
               sample=f"CPU oracle FP32-ref (portable C, not IPP), {nth} threads: preprocessing and filtering of all "
                      f"{w.n_views} views of {w.n_u}x{w.n_v}, backprojection of {nslices} centre slices "
                      f"({w.X}x{w.Y}x{nslices}); value = whole-volume FDK GUPS with the backprojection "
                      f"extrapolated linearly from the slab ({t_pre + t_filt + t_bp:.1f} s measured)")
    return cpu, v32, z0


def run_adjoint(args, dev):
    """The differentiation form at 256^3 from 360 views of 512^2: forward
    projector A (jh_fp_project) and backprojector A^T (jh_bp_backproject of
    packed FP32 frames), timed with events; dot-product test in the same run."""
    import torch
    from jenhine_amd import configs, synth
    from jenhine_amd.recon import BPCore

    w = configs.ADJOINT256
    if args.views:
        w = w.scaled(n_views=args.views)
    bp = BPCore(w.ptms(), w.n_u, w.n_v, w.X, w.Y, w.Z, w.pxy, w.pz, device=dev)
    x = torch.from_numpy(synth.normal(21, (w.Z, w.Y, w.X))).to(dev)
    y = torch.from_numpy(synth.normal(22, (w.n_views, w.n_v, w.n_u))).to(dev)
    proj = torch.empty((w.n_views, w.n_v, w.n_u), dtype=torch.float32, device=dev)
    back = torch.empty((w.Z, w.Y, w.X), dtype=torch.float32, device=dev)
    q = bp.pack(y, "f32")
    for _ in range(max(1, args.warmup)):
        bp.project(x, out=proj)
        bp.backproject(q, 0, w.n_views, out=back)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ev[0].record()
    for _ in range(args.steps):
        bp.project(x, out=proj)
    ev[1].record()
    for _ in range(args.steps):
        bp.backproject(q, 0, w.n_views, out=back)
    ev[2].record()
    torch.cuda.synchronize()
    fp_ms = ev[0].elapsed_time(ev[1]) / args.steps
    bp_ms = ev[1].elapsed_time(ev[2]) / args.steps
    units = w.gups_units
    lhs = float((proj.double() * y.double()).sum())
    rhs = float((x.double() * back.double()).sum())
    gap = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    log(f"forward {fp_ms:.2f} ms, back {bp_ms:.2f} ms, dot-test gap {gap:.2e}")
    line = {
        "metric": "adjoint pair GUPS at 256^3 x 360 views of 512^2 (forward projector A and backprojector A^T)",
        "value": 2 * units / ((fp_ms + bp_ms) * 1e-3) / 1e9, "unit": "GUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": fp_ms + bp_ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (x, y ~ N(0, 1))",
        "config": {"workload": f"adjoint256: {w.X}x{w.Y}x{w.Z} at {w.pxy} mm, {w.n_views} views of "
                               f"{w.n_u}x{w.n_v} at {w.du} mm", "parallelism": "single-gpu"},
        "forward": {"ms": fp_ms, "gups": units / (fp_ms * 1e-3) / 1e9},
        "back": {"ms": bp_ms, "gups": units / (bp_ms * 1e-3) / 1e9},
        "dot_test_rel_gap": gap,
    }
    print(json.dumps(line), flush=True)


def pipeline_command(args, tmp):
    """dental512 inputs written to tmp and the jh_fdk_pipeline command line
    (BF16 storage, the headline filter form, args.chunk views per chunk)."""
    from jenhine_amd import configs

    app = os.path.join(ROOT, "jenhine_amd", "host", "build", "jh_fdk_pipeline")
    if not os.path.exists(app):
        sys.exit(f"{app} is not built")
    w = configs.DENTAL512
    if args.views:
        w = w.scaled(n_views=args.views)
    d = w.data()
    scan, rec = w.scan_params(), w.recon_params()
    filt = {"mfma": 1, "mfma6": 2, "mfma16": 4}[args.filter_form]
    os.makedirs(tmp, exist_ok=True)
    paths = {}
    for k, dt in (("raw", np.uint16), ("ptm", np.float64), ("dark", np.uint16), ("gain", np.float32)):
        paths[k] = os.path.join(tmp, f"{k}.bin")
        np.ascontiguousarray(d[k], dt).tofile(paths[k])
    return [app, f"views={w.n_views}", f"dso={w.sod}", f"dsd={w.sdd}", f"detu={w.n_u}", f"detv={w.n_v}",
            f"pitch={w.du}", f"bright_field={scan.BrightField}",
            "bh=" + "|".join(str(b) for b in scan.BeamHardeningParams), f"x={w.X}", f"y={w.Y}", f"z={w.Z}",
            f"pitch_xy={w.pxy}", f"pitch_z={w.pz}", f"cut_start={rec.FilterCutOffStart}",
            f"cut_end={rec.FilterCutOffEnd}", f"window={rec.FilterWindow}", f"filter={filt}", "storage=1",
            f"chunk={args.chunk}", f"out={os.path.join(tmp, 'vol.bin')}"] + [f"{k}={v}" for k, v in paths.items()]


def run_pipeline(args):
    """dental512 through the C++ agent graph (jh_fdk_pipeline): raw counts in
    host memory streamed to the device per view chunk, slices downloaded to
    host memory, BF16 storage and the headline filter form; value = GUPS of
    the graph's run (Start to Join), PCIe transfers included."""
    import subprocess
    import tempfile
    from jenhine_amd import configs

    w = configs.DENTAL512
    if args.views:
        w = w.scaled(n_views=args.views)
    runs = []
    with tempfile.TemporaryDirectory() as tmp:
        cmd = pipeline_command(args, tmp)
        for r in range(args.warmup + args.steps):
            out = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
            if out.returncode != 0:
                sys.exit(out.stderr[-2000:])
            info = json.loads(out.stdout.strip().splitlines()[-1])
            log(f"run {r}: {info['seconds']:.3f} s, {w.gups_units / info['seconds'] / 1e9:.0f} GUPS",
                "(warm-up)" if r < args.warmup else "")
            if r >= args.warmup:
                runs.append(info["seconds"])
        agents = [l for l in out.stderr.splitlines() if "occupied" in l]
    sec = float(np.median(runs))
    line = {
        "metric": "FDK GUPS at 512^3 x 600 views through the agent graph, host buffers at both ends",
        "value": w.gups_units / sec / 1e9, "unit": "GUPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sec * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16-storage/f32-accumulate",
        "data": "synthetic (seeded ellipsoid phantom, analytic line integrals, Poisson counts)",
        "config": {"workload": f"pipeline_dental512: jh_fdk_pipeline FDK {w.X}x{w.Y}x{w.Z} from {w.n_views} views "
                               f"of {w.n_u}x{w.n_v} u16 in host memory, slices to host memory, {args.chunk} views "
                               "per chunk", "parallelism": "single-gpu"},
        "seconds_per_run": runs, "agents": agents,
    }
    print(json.dumps(line), flush=True)


def summarize_timeline(d):
    """Kernel, copy and HIP API summary of one jh_fdk_pipeline run traced with
    rocprofv3 --runtime-trace --output-format csv: busy time and span of every
    kernel and copy direction from the first preprocessing kernel, the copies
    and kernels per 5 ms, and the HIP calls that took over 1 ms."""
    import csv
    import glob

    def rows(pattern):
        files = sorted(glob.glob(os.path.join(d, "**", pattern), recursive=True))
        if not files:
            sys.exit(f"no {pattern} under {d}")
        return list(csv.DictReader(open(files[0])))

    ev = [(int(k["Start_Timestamp"]), int(k["End_Timestamp"]), "kernel " + k["Kernel_Name"].split("(")[0])
          for k in rows("*kernel_trace.csv")]
    ev += [(int(m["Start_Timestamp"]), int(m["End_Timestamp"]), "copy " + m["Direction"].replace("MEMORY_COPY_", ""))
           for m in rows("*memory_copy_trace.csv")]
    starts = [e[0] for e in ev if "preprocess" in e[2]]
    if not starts:
        sys.exit("no preprocessing kernel in the trace")
    t0, t1 = min(starts), max(e[1] for e in ev)
    print(f"from the first preprocessing kernel to the last event: {(t1 - t0) / 1e6:.2f} ms")
    agg = {}
    for s, e, n in ev:
        a = agg.setdefault(n, [0, 0.0, s, e])
        a[0] += 1
        a[1] += (e - s) / 1e6
        a[2], a[3] = min(a[2], s), max(a[3], e)
    print("| event | count | busy ms | first start ms | last end ms |")
    print("|---|---|---|---|---|")
    for n, (c, b, s, e) in sorted(agg.items(), key=lambda x: x[1][2]):
        print(f"| `{n}` | {c} | {b:.2f} | {(s - t0) / 1e6:.2f} | {(e - t0) / 1e6:.2f} |")
    step = 5e6
    nb = int((t1 - t0) / step) + 1
    bins = {"kernel": [0.0] * nb, "copy HOST_TO_DEVICE": [0.0] * nb, "copy DEVICE_TO_HOST": [0.0] * nb}
    for s, e, n in ev:
        key = "kernel" if n.startswith("kernel") else n
        if key not in bins or e <= t0:
            continue
        x = max(s, t0)
        while x < e:
            i = int((x - t0) / step)
            y = min(e, t0 + (i + 1) * step)
            bins[key][i] += (y - x) / 1e6
            x = y
    print()
    print("| ms from the first preprocessing kernel | kernel busy ms (all streams) | H2D busy ms | D2H busy ms |")
    print("|---|---|---|---|")
    for i in range(nb):
        print(f"| {i * 5}–{i * 5 + 5} | {bins['kernel'][i]:.2f} | {bins['copy HOST_TO_DEVICE'][i]:.2f} | "
              f"{bins['copy DEVICE_TO_HOST'][i]:.2f} |")
    api = rows("*hip_api_trace.csv")
    slow = [(int(a["Start_Timestamp"]), int(a["End_Timestamp"]), a["Function"], a["Thread_Id"]) for a in api
            if int(a["End_Timestamp"]) - int(a["Start_Timestamp"]) > 1_000_000]
    print()
    print("HIP calls over 1 ms (ms from the first preprocessing kernel, thread):")
    for s, e, f, t in sorted(slow):
        print(f"  {(s - t0) / 1e6:8.2f} .. {(e - t0) / 1e6:8.2f}  {f}  thread {t}")


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
    for name, (mode, storage) in {"fp32": ("fp32", "f32"), "bf16": (args.filter_form, "bf16")}.items():
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
        e["fdk"] = FDK(w.scan_params(), w.recon_params(), d["ptm"],
                       "fp32" if args.batch_storage == "f32" else args.filter_form, args.batch_storage, dev)
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
