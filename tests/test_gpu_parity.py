"""GPU parity of the HIP back end against the CPU oracle, through the C-ABI.

Tolerances:
  * preprocessing: integer steps bit-exact; float result within 2 ulp of the
    FP32-ref oracle (the only non-exact step is ln: OCML logf vs glibc logf);
  * filter FP32 (MFMA and vector ALU) and hi/mid/lo BF16 (6 products):
    rel-RMSE <= 2e-6 vs the FP64 oracle; row-scaled hi/lo FP16 (3 products):
    <= 3e-6 (22 significant bits per operand against FP32's 24, so about
    twice the FP32 forms' ~1e-6); hi/lo BF16 (3 products): <= 3e-4
    (each operand keeps ~17 significant bits, and the ramp filter's
    cancellation, sum|terms| / |result| ~ 20, amplifies that);
  * backprojection FP32: rel-RMSE <= 1e-4 vs the FP32-ref oracle volume (the
    FP32 path's accuracy bar) -- observed ~1e-6; BF16 / FP16 storage: <= 1e-3
    vs FP64.
"""
import numpy as np
import pytest
import torch

from jenhine_amd import configs, geometry, synth
from oracle import oracle as o
from tests.helpers import max_abs, oracle_params, rel_rmse, ulp_diff_f32

pytestmark = pytest.mark.gpu


def _u16(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, np.uint16).view(np.int16)).to(dev)


SMALL = configs.CBCT128.scaled(n_views=24, name="cbct128-24v")


@pytest.fixture(scope="module")
def small_data():
    return SMALL.data()


# --------------------------------------------------------------------------
# Preprocessing
# --------------------------------------------------------------------------

def test_preprocess_all_u16_values(gpu):
    from jenhine_amd.recon import PreProcessingCore
    raw = np.arange(65536, dtype=np.uint16).reshape(1, 256, 256)
    pre = PreProcessingCore(256, 256, bright_field=1000.0, device=gpu)
    out = pre.process(_u16(raw, gpu)).cpu().numpy()
    ref = o.preprocess(raw, bright_field=1000.0)
    assert out.shape == ref.shape
    assert ulp_diff_f32(out, ref).max() <= 2
    # the clamp of LowerBound(1): counts 0 and 1 both give -ln(1/BF)
    assert out[0, 0, 0] == out[0, 0, 1]


def test_preprocess_dark_gain_bh_borders(gpu, small_data):
    from jenhine_amd.recon import PreProcessingCore
    d = small_data
    raw = d["raw"][:6]
    bh = (1.0, 0.03, 0.004)
    borders = (3, 2, 4, 1)
    H, W = raw.shape[1:]
    oh, ow = H - 5, W - 5
    gain = d["gain"][3:3 + oh, 4:4 + ow].copy()
    pre = PreProcessingCore(W, H, borders, bright_field=4000.0, bh=bh, device=gpu)
    out = pre.process(_u16(raw, gpu), _u16(d["dark"], gpu), torch.from_numpy(gain).to(gpu)).cpu().numpy()
    ref = o.preprocess(raw, borders=borders, bright_field=4000.0, bh=bh, dark=d["dark"], gain=gain)
    assert ulp_diff_f32(out, ref).max() <= 4


def test_preprocess_aligned_rows_dark_gain_bh(gpu, small_data):
    """Rows aligned to 8 pixels take the 16-byte-load kernel (dark, gain,
    beam hardening): same bar against the oracle as the per-pixel kernel."""
    from jenhine_amd.recon import PreProcessingCore
    d = small_data
    raw = d["raw"][:6]
    bh = (1.0, 0.03, 0.004)
    H, W = raw.shape[1:]
    assert W % 8 == 0
    pre = PreProcessingCore(W, H, bright_field=4000.0, bh=bh, device=gpu)
    out = pre.process(_u16(raw, gpu), _u16(d["dark"], gpu), torch.from_numpy(d["gain"]).to(gpu)).cpu().numpy()
    ref = o.preprocess(raw, bright_field=4000.0, bh=bh, dark=d["dark"], gain=d["gain"])
    assert ulp_diff_f32(out, ref).max() <= 4


def test_preprocess_dark_below_raw_clamps(gpu):
    from jenhine_amd.recon import PreProcessingCore
    raw = np.full((1, 8, 16), 100, np.uint16)
    dark = np.full((8, 16), 150, np.uint16)
    pre = PreProcessingCore(16, 8, bright_field=50.0, device=gpu)
    out = pre.process(_u16(raw, gpu), _u16(dark, gpu)).cpu().numpy()
    # raw - dark < 1 is clamped to 1 (LowerBound, PreProcessingCore.cpp:74)
    assert np.all(out == out.flat[0])
    ref = o.preprocess(raw, bright_field=50.0, dark=dark)
    assert ulp_diff_f32(out, ref).max() <= 2
    assert abs(out.flat[0] - np.log(50.0)) < 1e-6


def test_preprocess_resize(gpu):
    from jenhine_amd.recon import PreProcessingCore
    raw = synth.uniform(3, (2, 40, 48), 50, 3000).astype(np.uint16)
    pre = PreProcessingCore(48, 40, (2, 2, 4, 4), out_w=20, out_h=18, bright_field=4000.0, device=gpu)
    out = pre.process(_u16(raw, gpu)).cpu().numpy()
    ref = o.preprocess(raw, out_hw=(18, 20), borders=(2, 2, 4, 4), bright_field=4000.0)
    np.testing.assert_allclose(out, ref, rtol=2e-6, atol=1e-6)


# --------------------------------------------------------------------------
# Filter
# --------------------------------------------------------------------------

def _filter_case(gpu, small_data, mode):
    from jenhine_amd.recon import FilterCore
    w = SMALL
    scan, rec = w.scan_params(), w.recon_params()
    pre = o.preprocess(small_data["raw"], bright_field=scan.BrightField, dark=small_data["dark"],
                       gain=small_data["gain"], fp64=True)
    ptm = small_data["ptm"]
    flt = FilterCore(w.n_u, w.n_v, ptm, scan.DSO, scan.HalfSampleRate, rec.FilterCutOffStart,
                     rec.FilterCutOffEnd, window=rec.FilterWindow, mode=mode, device=gpu)
    out = flt.process(torch.from_numpy(pre.astype(np.float32)).to(gpu)).cpu().numpy()
    H = o.filter_design(w.n_u, scan.HalfSampleRate, rec.FilterCutOffStart, rec.FilterCutOffEnd,
                        window=rec.FilterWindow, fp64=True)
    src, blk = o.preweight_schedule(w.n_views)
    pws = np.stack([o.prefilter_weight(ptm[s], w.n_u, w.n_v, scan.DSO, fp64=True) for s in src])
    ref = o.filter_rows(pre, H, pws, blk, fp64=True)
    return flt, out, ref, H, pws


def test_filter_fp32_vs_fp64_oracle(gpu, small_data):
    flt, out, ref, H, pws = _filter_case(gpu, small_data, "fp32")
    Hd, L = flt.response()
    assert L == o.fft_length(SMALL.n_u)
    np.testing.assert_allclose(Hd, H, rtol=1e-9, atol=1e-12 * np.abs(H).max())
    pw = flt.preweight(13).cpu().numpy()
    np.testing.assert_allclose(pw, pws[13 // 9][:, :SMALL.n_u], rtol=2e-7)
    e = rel_rmse(out, ref)
    print("filter fp32 rel-RMSE", e, "max-abs", max_abs(out, ref))
    assert e <= 2e-6


@pytest.mark.parametrize("mode,tol", [("mfma", 3e-4), ("mfma6", 2e-6), ("fp32valu", 2e-6), ("mfma16", 3e-6)])
def test_filter_other_forms_vs_fp64_oracle(gpu, small_data, mode, tol):
    flt, out, ref, H, pws = _filter_case(gpu, small_data, mode)
    e = rel_rmse(out, ref)
    print("filter", mode, "rel-RMSE", e, "max-abs", max_abs(out, ref))
    assert e <= tol


def test_filter_mfma_partial_column_tile(gpu):
    # W = 200: the last 128-column tile is partial and K = 203 -> 224
    from jenhine_amd.recon import FilterCore
    w = SMALL.scaled(n_views=3, n_u=200, n_v=40)
    scan = w.scan_params()
    frames = np.ascontiguousarray(synth.uniform(8, (3, 40, 200), 0.0, 3.0))
    ptm = w.ptms()
    H = o.filter_design(200, scan.HalfSampleRate, window=1, fp64=True)
    src, blk = o.preweight_schedule(3)
    pws = np.stack([o.prefilter_weight(ptm[s], 200, 40, scan.DSO, fp64=True) for s in src])
    ref = o.filter_rows(frames.astype(np.float64), H, pws, blk, fp64=True)
    for mode, tol in (("fp32", 2e-6), ("mfma6", 2e-6), ("mfma", 3e-4), ("mfma16", 3e-6)):
        flt = FilterCore(200, 40, ptm, scan.DSO, scan.HalfSampleRate, window=1, mode=mode, device=gpu)
        out = flt.process(torch.from_numpy(frames).to(gpu)).cpu().numpy()
        assert rel_rmse(out, ref) <= tol, mode


def test_filter_preweight_schedule_on_a_non_circular_orbit(gpu):
    """Source distance, detector distance and the principal point change from
    view to view, so every 9-view block has its own pre-weight map
    (FilterAgent/FilterAgent.cpp:129-145): the device filter must pick the
    same source view per block as the oracle's schedule."""
    from jenhine_amd.recon import FilterCore
    n, W, V, dso = 25, 96, 40, 300.0

    def view(k):
        P = geometry.circular_ptms(1, dso + 80.0 * np.sin(0.7 * k), 450.0 + 60.0 * np.cos(0.3 * k), W, V, 0.6,
                                   start_deg=360.0 * k / n, u_shift=25.0 * np.sin(k))
        return geometry.move_v(P, 12.0 * np.cos(1.3 * k))

    P = np.concatenate([view(k) for k in range(n)])
    frames = np.ascontiguousarray(synth.uniform(11, (n, V, W), 0.0, 3.0))
    H = o.filter_design(W, 0.5, window=1, fp64=True)
    src, blk = o.preweight_schedule(n)
    assert len(src) == 3
    pws = np.stack([o.prefilter_weight(P[s], W, V, dso, fp64=True) for s in src])
    ref = o.filter_rows(frames.astype(np.float64), H, pws, blk, fp64=True)
    # the same rows filtered with another view of each block as the source
    # differ by far more than the bar, so a schedule mistake cannot pass
    for shift in (-4, -3, 1):
        other = np.stack([o.prefilter_weight(P[s + shift], W, V, dso, fp64=True) for s in src])
        assert rel_rmse(o.filter_rows(frames.astype(np.float64), H, other, blk, fp64=True), ref) > 1e-4
    for mode, tol in (("fp32", 2e-6), ("mfma6", 2e-6), ("mfma16", 3e-6)):
        flt = FilterCore(W, V, P, dso, 0.5, window=1, mode=mode, device=gpu)
        out = flt.process(torch.from_numpy(frames).to(gpu)).cpu().numpy()
        e = rel_rmse(out, ref)
        print("non-circular filter", mode, e)
        assert e <= tol, mode


# --------------------------------------------------------------------------
# Backprojection
# --------------------------------------------------------------------------

def _bp_gpu(gpu, ptm, filt, w, dtype="f32", chunks=None, tile=None):
    from jenhine_amd.recon import BPCore
    bp = BPCore(ptm, filt.shape[2], filt.shape[1], w[0], w[1], w[2], w[3], w[4], device=gpu)
    if tile:
        bp.set_tile(*tile)
    f = torch.from_numpy(np.ascontiguousarray(filt, np.float32)).to(gpu)
    if chunks is None:
        q = bp.stage(f, dtype)
        return bp.run(q, dtype).cpu().numpy(), bp
    bp.reset()
    for v0 in range(0, filt.shape[0], chunks):
        n = min(chunks, filt.shape[0] - v0)
        q = bp.stage(f[v0:v0 + n], dtype, view0=v0)
        bp.accumulate(q, dtype, v0, n)
    return bp.finalize().cpu().numpy(), bp


def test_bp_uweights_match_oracle(gpu, small_data):
    from jenhine_amd.recon import BPCore
    w = SMALL
    bp = BPCore(small_data["ptm"], w.n_u, w.n_v, w.X, w.Y, w.Z, w.pxy, w.pz, device=gpu)
    ref = o.uweights(small_data["ptm"], w.n_u, fp64=True)
    np.testing.assert_allclose(bp.uweights(), ref, rtol=1e-7, atol=1e-7)
    assert bp.geometry_mode() == 0


@pytest.mark.parametrize("tile", [(32, 4), (16, 2), (8, 1), (32, 1)])
def test_bp_fp32_vs_fp32_ref(gpu, tile):
    # random filtered projections on the cbct128 geometry at 24 views
    w = SMALL
    ptm = w.ptms()
    filt = synth.normal(11, (w.n_views, w.n_v, w.n_u))
    shape = (w.X, w.Y, w.Z, w.pxy, w.pz)
    vol, _ = _bp_gpu(gpu, ptm, filt, shape, tile=tile)
    ref = o.backproject(filt, ptm, (w.Z, w.Y, w.X), w.pxy, w.pz)
    assert np.mean(np.isnan(vol) != np.isnan(ref)) < 1e-4
    e = rel_rmse(vol, ref)
    print("bp fp32 tile", tile, "rel-RMSE", e, "max-abs", max_abs(vol, ref))
    assert e <= 1e-4


def _streamed(bp, q, dtype, sizes):
    """Streaming accumulation of the staged views in launches of the given sizes."""
    fb = q.numel() // sum(sizes)
    bp.reset()
    v0 = 0
    for n in sizes:
        bp.accumulate(q[v0 * fb:(v0 + n) * fb], dtype, v0, n)
        v0 += n
    return bp.finalize().cpu().numpy()


# splits of 24 views into launches: whole, one view each, even, uneven
SPLITS24 = [[24], [1] * 24, [5, 5, 5, 5, 4], [3, 10, 11], [7, 17]]


@pytest.mark.parametrize("weight", ["count", "fdk"])
@pytest.mark.parametrize("fetch", ["gather", "lds"])
@pytest.mark.parametrize("dtype", ["f32", "bf16", "f16"])
def test_bp_streaming_equals_one_shot(gpu, dtype, fetch, weight):
    """The streaming form continues the running sums of the earlier launches
    (accumulators, the outside-row weights and the columns' count weights), so
    every split of the views into launches gives the one-launch volume to the
    bit: geometry class 0, gather and LDS-staged fetch, count-normalised and
    distance-weighted."""
    from jenhine_amd.recon import BPCore
    w = SMALL
    filt = torch.from_numpy(synth.normal(12, (w.n_views, w.n_v, w.n_u))).to(gpu)
    bp = BPCore(w.ptms(), w.n_u, w.n_v, w.X, w.Y, w.Z, w.pxy, w.pz, device=gpu)
    assert bp.geometry_mode() == 0
    bp.set_tile(16, 4)
    bp.set_fetch(fetch)
    if weight == "fdk":
        bp.set_weight("fdk")
    q = bp.stage(filt, dtype)
    one = bp.run(q, dtype).cpu().numpy()
    print(dtype, fetch, weight, "fetch in use", bp.fetch_in_use(dtype))
    assert np.isfinite(one).mean() > 0.5
    for sizes in SPLITS24:
        np.testing.assert_array_equal(_streamed(bp, q, dtype, sizes), one, err_msg=str(sizes))


def _slab_tail(bp, q, dtype, head, chunk, slab):
    """The first `head` views streamed over the whole volume in launches of
    `chunk`, the rest slab by slab (jh_bp_accumulate_slab), each slab
    finalised before the next one is accumulated."""
    n = bp.n_views
    fb = q.numel() // n
    bp.reset()
    for v0 in range(0, head, chunk):
        m = min(chunk, head - v0)
        bp.accumulate(q[v0 * fb:(v0 + m) * fb], dtype, v0, m)
    out = []
    for z0 in range(0, bp.Z, slab):
        nz = min(slab, bp.Z - z0)
        if head < n:
            bp.accumulate(q[head * fb:], dtype, head, n - head, z0=z0, nz=nz)
        out.append(bp.finalize(z0, nz).cpu().numpy())
    return np.concatenate(out)


@pytest.mark.parametrize("weight", ["count", "fdk"])
@pytest.mark.parametrize("fetch", ["gather", "lds"])
@pytest.mark.parametrize("dtype", ["f32", "bf16", "f16"])
def test_bp_slab_tail_equals_one_shot(gpu, dtype, fetch, weight):
    """Views streamed over the whole volume, then the remaining views z-slab
    by z-slab with each slab finalised in between (the agent's download
    overlap): every voxel still takes its views in order and the column count
    sums take each view once, so the volume is the one-launch volume to the
    bit -- slabs of one and two z tiles and a partial last slab, no streamed
    head, all views in the head."""
    from jenhine_amd.recon import BPCore
    w = SMALL
    filt = torch.from_numpy(synth.normal(14, (w.n_views, w.n_v, w.n_u))).to(gpu)
    bp = BPCore(w.ptms(), w.n_u, w.n_v, w.X, w.Y, w.Z, w.pxy, w.pz, device=gpu)
    assert bp.geometry_mode() == 0
    bp.set_tile(8, 4)
    bp.set_fetch(fetch)
    if weight == "fdk":
        bp.set_weight("fdk")
    q = bp.stage(filt, dtype)
    one = bp.run(q, dtype).cpu().numpy()
    n = w.n_views
    for head, chunk, slab in ((n // 2, 5, 8), (7, 7, 16), (0, 1, 8), (n, 6, 8), (n - 1, 4, 24)):
        got = _slab_tail(bp, q, dtype, head, chunk, slab)
        np.testing.assert_array_equal(got, one, err_msg=f"head {head} chunk {chunk} slab {slab}")


@pytest.mark.parametrize("dtype", ["f32", "bf16", "f16"])
def test_bp_slab_tail_general_geometry(gpu, dtype):
    """The slab tail in geometry classes 1 and 2 (count summed per sample)."""
    from jenhine_amd.recon import BPCore
    P = np.array([-1.33897, -7.81194, 0.01236, 575.79443, -0.97515, -0.03862, 7.51861, 346.26642,
                  -0.00226, -0.00004, 0.0, 1.0], np.float64)
    tilt = P.copy()
    tilt[10] = 0.0003
    for base, mode in ((P, 1), (tilt, 2)):
        ptm = np.stack([base * (1.0 + 1e-4 * k) for k in range(6)])
        filt = torch.from_numpy(synth.normal(13, (6, 686, 644))).to(gpu)
        bp = BPCore(ptm, 644, 686, 96, 80, 20, 0.25, 1.0, device=gpu)
        assert bp.geometry_mode() == mode
        zc, _ = bp.tile()
        q = bp.stage(filt, dtype)
        one = bp.run(q, dtype).cpu().numpy()
        for head in (0, 2, 5):
            got = _slab_tail(bp, q, dtype, head, 2, zc)
            np.testing.assert_array_equal(got, one, err_msg=f"mode {mode} head {head}")


@pytest.mark.parametrize("dtype", ["f32", "bf16", "f16"])
def test_bp_streaming_equals_one_shot_general_geometry(gpu, dtype):
    """As above in geometry classes 1 (u depends on z) and 2 (w too), where
    the count is summed per sample: the reference test's matrix
    (UTBPCUDACore.cpp:20-32) and tilted copies, 6 views."""
    from jenhine_amd.recon import BPCore
    P = np.array([-1.33897, -7.81194, 0.01236, 575.79443, -0.97515, -0.03862, 7.51861, 346.26642,
                  -0.00226, -0.00004, 0.0, 1.0], np.float64)
    tilt = P.copy()
    tilt[10] = 0.0003
    for base, mode in ((P, 1), (tilt, 2)):
        ptm = np.stack([base * (1.0 + 1e-4 * k) for k in range(6)])
        filt = torch.from_numpy(synth.normal(13, (6, 686, 644))).to(gpu)
        bp = BPCore(ptm, 644, 686, 96, 80, 20, 0.25, 1.0, device=gpu)
        assert bp.geometry_mode() == mode
        q = bp.stage(filt, dtype)
        one = bp.run(q, dtype).cpu().numpy()
        assert np.isfinite(one).mean() > 0.5
        for sizes in ([1] * 6, [2, 4], [5, 1]):
            np.testing.assert_array_equal(_streamed(bp, q, dtype, sizes), one, err_msg=f"mode {mode} {sizes}")


def test_bp_general_geometry_modes(gpu):
    # the reference's own test matrix (UTBPCUDACore.cpp:20-32) has u depending
    # on z (mode 1); a tilted copy also makes w depend on z (mode 2)
    P = np.array([-1.33897, -7.81194, 0.01236, 575.79443, -0.97515, -0.03862, 7.51861, 346.26642,
                  -0.00226, -0.00004, 0.0, 1.0], np.float64)
    tilt = P.copy()
    tilt[10] = 0.0003
    for ptm, mode in ((np.stack([P, P * 1.0001]), 1), (np.stack([tilt, P]), 2)):
        ptm = ptm.reshape(-1, 12)
        filt = synth.normal(13, (2, 686, 644))  # the reference test's detector
        shape = (96, 80, 20, 0.25, 1.0)
        vol, bp = _bp_gpu(gpu, ptm, filt, shape)
        assert bp.geometry_mode() == mode
        ref = o.backproject(filt, ptm.astype(np.float32), (20, 80, 96), 0.25, 1.0)
        assert np.mean(np.isnan(vol) != np.isnan(ref)) < 1e-3
        assert np.isfinite(ref).mean() > 0.5
        e = rel_rmse(vol, ref)
        print("mode", mode, "rel-RMSE", e)
        # white-noise input with 2 views: every voxel error is the float
        # rounding of its coordinate (~1e-4 px) times a gradient of ~|q| per
        # pixel, with no averaging over views
        assert e <= 3e-4


@pytest.mark.parametrize("dtype,tol", [("bf16", 5e-3), ("f16", 1e-3)])
def test_bp_general_geometry_modes_16bit_storage(gpu, dtype, tol):
    # the per-sample weight packing of the 16-bit samplers (modes 1 and 2
    # form the u weights sample by sample); white noise, 2 views, no averaging:
    # BF16 rounds each value by up to 2^-9 relative and the near-column weight
    # moves u by up to 2^-10 pixel, FP16 by 2^-12 and 2^-12 pixel
    P = np.array([-1.33897, -7.81194, 0.01236, 575.79443, -0.97515, -0.03862, 7.51861, 346.26642,
                  -0.00226, -0.00004, 0.0, 1.0], np.float64)
    tilt = P.copy()
    tilt[10] = 0.0003
    for ptm, mode in ((np.stack([P, P * 1.0001]), 1), (np.stack([tilt, P]), 2)):
        ptm = ptm.reshape(-1, 12)
        filt = synth.normal(13, (2, 686, 644))
        shape = (96, 80, 20, 0.25, 1.0)
        vol, bp = _bp_gpu(gpu, ptm, filt, shape, dtype=dtype)
        assert bp.geometry_mode() == mode
        ref = o.backproject(filt, ptm.astype(np.float32), (20, 80, 96), 0.25, 1.0)
        assert np.mean(np.isnan(vol) != np.isnan(ref)) < 1e-3
        e = rel_rmse(vol, ref)
        print("mode", mode, dtype, "rel-RMSE vs FP32-ref", e)
        assert e <= tol


def test_bp_reference_gpu_case(gpu):
    # UTBPCUDACore.cpp:16-201: one view, 644 x 686 detector, 500 x 500 x 50
    # volume; projection 0.5 on its first half and 1.0 on the second; at most
    # N/500 voxels may differ from the CPU result by more than 0.05.
    P = np.array([-1.33897, -7.81194, 0.01236, 575.79443, -0.97515, -0.03862, 7.51861, 346.26642,
                  -0.00226, -0.00004, 0.0, 1.0], np.float64).reshape(1, 12)
    U, V = 644, 686
    proj = np.ones((1, V, U), np.float32)
    proj.reshape(-1)[: U * V // 2] = 0.5
    vol, _ = _bp_gpu(gpu, P, proj, (500, 500, 50, 0.25, 1.0))
    ref = o.backproject(proj, P.astype(np.float32), (50, 500, 500), 0.25, 1.0)
    both = np.isfinite(vol) & np.isfinite(ref)
    bad = np.sum(np.abs(vol[both] - ref[both]) > 0.05) + np.sum(np.isfinite(vol) != np.isfinite(ref))
    print("voxels off by more than 0.05 or NaN mismatched:", bad, "of", vol.size)
    assert bad <= vol.size // 500
    assert rel_rmse(vol, ref) <= 1e-5


def test_bp_constant_projection_is_constant_volume(gpu):
    # normalisation invariant: vol = c wherever any view saw the voxel
    w = SMALL
    ptm = w.ptms()
    filt = np.full((w.n_views, w.n_v, w.n_u), 0.75, np.float32)
    vol, _ = _bp_gpu(gpu, ptm, filt, (w.X, w.Y, w.Z, w.pxy, w.pz))
    m = np.isfinite(vol)
    assert m.mean() > 0.9
    np.testing.assert_allclose(vol[m], 0.75, rtol=2e-6)


@pytest.mark.parametrize("dtype,tol", [("bf16", 1e-3), ("f16", 2e-4)])
def test_bp_reduced_storage_vs_fp64(gpu, small_data, dtype, tol):
    w = SMALL
    sc, rc = oracle_params(w)
    vol64, st = o.fdk(small_data["raw"], small_data["ptm"], sc, rc, fp64=True, dark=small_data["dark"],
                      gain=small_data["gain"], return_stages=True)
    filt = st["filt"].astype(np.float32)
    vol, _ = _bp_gpu(gpu, small_data["ptm"], filt, (w.X, w.Y, w.Z, w.pxy, w.pz), dtype=dtype)
    e = rel_rmse(vol, vol64)
    print(dtype, "rel-RMSE vs FP64", e, "max-abs", max_abs(vol, vol64))
    assert e <= tol


# --------------------------------------------------------------------------
# Full chain
# --------------------------------------------------------------------------

@pytest.mark.parametrize("filter_mode,storage,tol64", [("fp32", "f32", 1e-4), ("mfma", "bf16", 1e-3),
                                                        ("mfma6", "bf16", 1e-3), ("mfma6", "f16", 1e-3)])
def test_fdk_cbct128_vs_oracle(gpu, filter_mode, storage, tol64):
    from jenhine_amd.recon import FDK
    w = configs.CBCT128
    d = w.data()
    sc, rc = oracle_params(w)
    fdk = FDK(w.scan_params(), w.recon_params(), d["ptm"], filter_mode, storage, device=gpu)
    vol = fdk(_u16(d["raw"], gpu), _u16(d["dark"], gpu), torch.from_numpy(d["gain"]).to(gpu)).cpu().numpy()
    v64 = o.fdk(d["raw"], d["ptm"], sc, rc, fp64=True, dark=d["dark"], gain=d["gain"])
    e64 = rel_rmse(vol, v64)
    print(filter_mode, storage, "rel-RMSE vs FP64", e64, "max-abs", max_abs(vol, v64))
    assert e64 <= tol64
    if storage == "f32":
        v32 = o.fdk(d["raw"], d["ptm"], sc, rc, fp64=False, dark=d["dark"], gain=d["gain"])
        e32 = rel_rmse(vol, v32)
        print("rel-RMSE vs FP32-ref", e32)
        assert e32 <= 1e-4


def test_bp_autotune_keeps_results(gpu):
    w = SMALL
    ptm = w.ptms()
    filt = torch.from_numpy(synth.normal(14, (w.n_views, w.n_v, w.n_u))).to(gpu)
    from jenhine_amd.recon import BPCore
    bp = BPCore(ptm, w.n_u, w.n_v, w.X, w.Y, w.Z, w.pxy, w.pz, device=gpu)
    q = bp.stage(filt, "bf16")
    before = bp.run(q, "bf16").cpu().numpy()
    table = bp.autotune(q, "bf16", reps=1)
    # every gather tile is timed; an LDS tile is left out where its footprint
    # exceeds the LDS (the launch would gather)
    assert all(("gather", z, t) in table for z in (8, 16, 32) for t in (1, 2, 4))
    assert all(v > 0 for v in table.values())
    after = bp.run(q, "bf16").cpu().numpy()
    print("autotuned", bp.fetch(), "tile", bp.tile(), "rel-RMSE vs default", rel_rmse(after, before))
    # the z tile origin enters the float sample coordinates, so another tile
    # shape rounds them differently; on white-noise input that shows at ~1e-5
    assert rel_rmse(after, before) <= 1e-4


@pytest.mark.parametrize("dtype", ["f32", "bf16", "f16"])
@pytest.mark.parametrize("tile", [(32, 4), (16, 4), (8, 4), (16, 2), (8, 1)])
def test_bp_lds_fetch_matches_gather(gpu, dtype, tile):
    """The LDS-staged fetch samples the same positions from the same values:
    FP32 quads are staged as bilinear coefficients (only the association of
    the bilinear terms differs); BF16 / FP16 column strips are copied into LDS
    as they are and sampled with the gather form's operations (equal to the
    bit).  Strips take the LDS form in workgroups of 4 waves (16 x 16 voxel
    columns) only; at the other tiles the launch gathers."""
    w = SMALL
    ptm = w.ptms()
    filt = torch.from_numpy(synth.normal(21, (w.n_views, w.n_v, w.n_u))).to(gpu)
    from jenhine_amd.recon import BPCore
    bp = BPCore(ptm, w.n_u, w.n_v, w.X, w.Y, w.Z, w.pxy, w.pz, device=gpu)
    assert bp.geometry_mode() == 0
    bp.set_tile(*tile)
    q = bp.stage(filt, dtype)
    res = {}
    for f in ("gather", "lds"):
        bp.set_fetch(f)
        assert bp.fetch() == f
        expect = "lds" if f == "lds" and (dtype == "f32" or tile[1] == 4) else "gather"
        assert bp.fetch_in_use(dtype) == expect
        vol = bp.run(q, dtype).cpu().numpy()
        bp.reset()
        bp.accumulate(q[: q.numel() // 2], dtype, 0, w.n_views // 2)
        bp.accumulate(q[q.numel() // 2:], dtype, w.n_views // 2, w.n_views - w.n_views // 2)
        streamed = bp.finalize().cpu().numpy()
        res[f] = (vol, streamed)
    for k in range(2):
        a, b = res["gather"][k], res["lds"][k]
        np.testing.assert_array_equal(np.isnan(a), np.isnan(b))
        e = rel_rmse(b, a)
        print(dtype, tile, ["one-shot", "streamed"][k], "LDS vs gather rel-RMSE", e)
        assert e <= 1e-6
        if dtype != "f32":
            np.testing.assert_array_equal(b, a)


def test_bp_lds_fetch_falls_back_beyond_budget(gpu):
    """A detector of very fine pixels makes every block's footprint exceed the
    LDS budget: the staged kernel then gathers, with the same result."""
    w = SMALL.scaled(du=0.06, n_u=160, n_v=160, name="fine-detector")
    ptm = w.ptms()
    filt = torch.from_numpy(synth.normal(22, (w.n_views, w.n_v, w.n_u))).to(gpu)
    from jenhine_amd.recon import BPCore
    bp = BPCore(ptm, w.n_u, w.n_v, w.X, w.Y, w.Z, w.pxy, w.pz, device=gpu)
    q = bp.stage(filt, "bf16")
    out = {}
    for f in ("gather", "lds"):
        bp.set_fetch(f)
        out[f] = bp.run(q, "bf16").cpu().numpy()
    assert np.isfinite(out["gather"]).sum() > 1000
    # blocks whose footprint is clipped to a thin strip of the detector may
    # still be staged, so the two agree to FP32 rounding, not to the bit
    np.testing.assert_array_equal(np.isnan(out["gather"]), np.isnan(out["lds"]))
    assert rel_rmse(out["lds"], out["gather"]) <= 1e-6


# --------------------------------------------------------------------------
# ZFilterAgent and CTNumAgent (fused into staging and normalisation)
# --------------------------------------------------------------------------

def test_zfilter_kernel_matches_oracle_bit_exact(gpu):
    from jenhine_amd.recon import zfilter
    x = synth.normal(41, (3, 40, 57)) * 3.0
    got = zfilter(torch.from_numpy(x).to(gpu)).cpu().numpy()
    np.testing.assert_array_equal(got, o.zfilter(x))
    one_row = synth.normal(42, (2, 1, 33))  # V = 1: every neighbour is the row itself
    np.testing.assert_array_equal(zfilter(torch.from_numpy(one_row).to(gpu)).cpu().numpy(), o.zfilter(one_row))


def test_ctnum_kernel_matches_oracle_bit_exact(gpu):
    from jenhine_amd.recon import ctnum
    x = synth.uniform(43, (10000,), -0.01, 0.2)
    x[:3] = [np.nan, 0.0, 1.0]
    for norm0, norm1, mu in ((1.0, 1.0, 0.02), (1.07, 0.62, 0.0193)):
        got = ctnum(torch.from_numpy(x.copy()).to(gpu), norm0, norm1, mu).cpu().numpy()
        np.testing.assert_array_equal(got, o.ctnum(x, norm0, norm1, mu))


@pytest.mark.parametrize("dtype", ["f32", "bf16", "f16"])
def test_stage_with_fused_zfilter_equals_zfilter_then_stage(gpu, dtype):
    w = SMALL
    from jenhine_amd.recon import BPCore, zfilter
    filt = torch.from_numpy(synth.normal(44, (w.n_views, w.n_v, w.n_u))).to(gpu)
    bp = BPCore(w.ptms(), w.n_u, w.n_v, w.X, w.Y, w.Z, w.pxy, w.pz, device=gpu)
    ref = bp.stage(zfilter(filt), dtype).cpu().numpy()
    bp.set_zfilter(True)
    got = bp.stage(filt, dtype).cpu().numpy()
    np.testing.assert_array_equal(got, ref)


def test_bp_fused_ctnum_equals_ctnum_of_volume(gpu):
    w = SMALL
    from jenhine_amd.recon import BPCore, ctnum
    filt = torch.from_numpy(np.abs(synth.normal(45, (w.n_views, w.n_v, w.n_u))) * 0.05).to(gpu)
    bp = BPCore(w.ptms(), w.n_u, w.n_v, w.X, w.Y, w.Z, w.pxy, w.pz, device=gpu)
    q = bp.stage(filt, "bf16")
    ref = ctnum(bp.run(q, "bf16"), 1.07, 0.62, 0.02).cpu().numpy()
    bp.reset()
    bp.accumulate(q, "bf16", 0, w.n_views)
    ref_s = ctnum(bp.finalize(), 1.07, 0.62, 0.02).cpu().numpy()
    bp.set_ctnum(True, 1.07, 0.62, 0.02)
    np.testing.assert_array_equal(bp.run(q, "bf16").cpu().numpy(), ref)
    np.testing.assert_array_equal(bp.finalize().cpu().numpy(), ref_s)
    assert (ref[np.isfinite(ref)] > 300).any()  # the compressed branch is exercised


def test_fdk_production_chain_with_zfilter_and_ct_vs_oracle(gpu):
    """Filter -> ZFilter -> BP -> CTNum (JEngine.cpp:240-305) against the
    oracle chain with the same two stages."""
    from jenhine_amd.recon import FDK
    w = SMALL
    d = small_data_cache()
    sc, rc = oracle_params(w)
    ct = (1.07, 0.62, 0.02)
    fdk = FDK(w.scan_params(), w.recon_params(), d["ptm"], "fp32", "f32", device=gpu, zfilter=True, ct=ct)
    vol = fdk(_u16(d["raw"], gpu), _u16(d["dark"], gpu), torch.from_numpy(d["gain"]).to(gpu)).cpu().numpy()
    v32 = o.fdk(d["raw"], d["ptm"], sc, rc, fp64=False, dark=d["dark"], gain=d["gain"], zfilter_v=True, ct=ct)
    v64 = o.fdk(d["raw"], d["ptm"], sc, rc, fp64=True, dark=d["dark"], gain=d["gain"], zfilter_v=True, ct=ct)
    plain = o.fdk(d["raw"], d["ptm"], sc, rc, fp64=True, dark=d["dark"], gain=d["gain"], ct=ct)
    e32, e64 = rel_rmse(vol, v32), rel_rmse(vol, v64)
    print("zfilter+ct chain: rel-RMSE vs FP32-ref", e32, "vs FP64", e64, "max-abs HU vs FP64", max_abs(vol, v64))
    assert e32 <= 1e-4 and e64 <= 1e-4
    assert rel_rmse(v64, plain) > 1e-3  # the z filter changes the result


_SMALL_DATA = {}


def small_data_cache():
    if not _SMALL_DATA:
        _SMALL_DATA.update(SMALL.data())
    return _SMALL_DATA


# --------------------------------------------------------------------------
# Offset-detector composite (CompositeAgent) and the mirrored filter
# --------------------------------------------------------------------------

def _offset_geometry(right, n_views=16, W=160, V=48):
    return configs.OFFSET128.scaled(n_views=n_views, n_u=W, n_v=V, u_shift=48.0 if right else -48.0)


@pytest.mark.parametrize("right", [True, False])
def test_composite_kernel_vs_fp64_oracle(gpu, right):
    """GPU composite (FP64-composed per-view homography, FP32 per pixel) vs
    the FP64 oracle: direct half bit-exact; conjugate half and transition
    within 1e-5 rel-RMSE (sample positions agree to ~1e-4 px)."""
    from jenhine_amd.recon import CompositeCore
    w = _offset_geometry(right)
    P = w.ptms()
    N, V, W = w.n_views, w.n_v, w.n_u
    pre = (synth.normal(17, (N, V, W)) + 2.0).astype(np.float32)
    comp = CompositeCore(W, V, P, w.sod, device=gpu)
    assert comp.detector_on_right == right == o.detector_on_right(P[0], W)
    got = comp.process(torch.from_numpy(pre).to(gpu)).cpu().numpy()
    ref = o.composite(pre.astype(np.float64), P, w.sod, right, fp64=True)
    dsl = slice(W, 2 * W) if right else slice(0, W)
    csl = slice(0, W) if right else slice(W, 2 * W)
    np.testing.assert_array_equal(got[:, :, dsl], pre)
    e = rel_rmse(got[:, :, csl], ref[:, :, csl])
    both = (got[:, :, csl] != 0) & (ref[:, :, csl] != 0)
    d = np.abs(got[:, :, csl] - ref[:, :, csl])[both]
    print("composite", "right" if right else "left", "rel-RMSE", e, "max-abs (both sampled)", d.max())
    assert e < 1e-5 and d.max() < 1e-3
    # the zero band (samples beyond the conjugate frame) agrees up to pixels
    # whose source coordinate sits within rounding of the frame border
    assert np.count_nonzero((got[:, :, csl] == 0) != (ref[:, :, csl] == 0)) <= N * V
    # pipe layout: frames interleaved (direct, conjugate) as InputAgent emits them
    pairs = np.empty((N, 2, V, W), np.float32)
    for v in range(N):
        pairs[v, 0] = pre[v]
        pairs[v, 1] = pre[(v + N // 2) % N]
    tp = torch.from_numpy(pairs).to(gpu)
    out = comp.process_pairs(tp[:, 0], tp[:, 1], 0)
    np.testing.assert_array_equal(out.cpu().numpy(), got)


@pytest.mark.parametrize("right", [True, False])
@pytest.mark.parametrize("mode,tol", [("fp32", 2e-6), ("mfma6", 2e-6), ("mfma", 3e-4), ("fp32valu", 2e-6)])
def test_filter_widthout_and_mirroring_vs_fp64_oracle(gpu, right, mode, tol):
    """FilterCore(width = 2W, widthOut = W, yMirroring = detectorOnTheRight)
    of the offset-detector chain, against the FP64 oracle fed the oracle's own
    pre-weight maps (shifted u axis)."""
    from jenhine_amd.recon import FilterCore
    w = _offset_geometry(right, n_views=12)
    P = w.ptms()
    N, V, W = w.n_views, w.n_v, w.n_u
    scan = w.scan_params()
    frames = np.ascontiguousarray(synth.uniform(19, (N, V, 2 * W), 0.0, 3.0))
    H = o.filter_design(2 * W, scan.HalfSampleRate, window=1, fp64=True)
    src, blk = o.preweight_schedule(N)
    pws = np.stack([o.prefilter_weight(P[s], 2 * W, V, scan.DSO, detector_on_right=right, fp64=True) for s in src])
    ref = o.filter_rows(frames.astype(np.float64), H, pws, blk, fp64=True, w_out=W, mirror=right)
    flt = FilterCore(2 * W, V, P, scan.DSO, scan.HalfSampleRate, window=1, mode=mode, device=gpu, w_out=W,
                     y_mirroring=right)
    pw = flt.preweight(10).cpu().numpy()
    np.testing.assert_allclose(pw, pws[10 // 9][:, :2 * W], rtol=2e-7)
    out = flt.process(torch.from_numpy(frames).to(gpu)).cpu().numpy()
    assert out.shape == (N, V, W)
    e = rel_rmse(out, ref)
    print("filter", mode, "mirror" if right else "plain", "2W->W rel-RMSE", e)
    assert e <= tol


@pytest.mark.parametrize("form,storage,tol", [("fp32", "f32", 1e-4), ("mfma6", "bf16", 1e-3), ("mfma6", "f16", 1e-3)])
def test_fdk_offset_detector_chain_vs_oracle(gpu, form, storage, tol):
    """The offset-detector chain Preprocess -> Composite -> Filter(2W -> W,
    yMirroring) -> BP (JEngine.cpp:152-305) against the CPU oracle."""
    from jenhine_amd.recon import FDK
    w = configs.OFFSET128.scaled(n_views=64)
    d = w.data()
    sc, rc = oracle_params(w)
    fdk = FDK(w.scan_params(), w.recon_params(), d["ptm"], form, storage, device=gpu, composite=True)
    assert fdk.composite.detector_on_right
    vol = fdk(_u16(d["raw"], gpu), _u16(d["dark"], gpu), torch.from_numpy(d["gain"]).to(gpu)).cpu().numpy()
    z0, nz = w.Z // 2 - 4, 8
    v64 = o.fdk(d["raw"], d["ptm"], sc, rc, fp64=True, dark=d["dark"], gain=d["gain"], composite_views=True,
                z0=z0, nz=nz)
    e64 = rel_rmse(vol[z0:z0 + nz], v64)
    msg = f"offset chain {form}/{storage}: rel-RMSE vs FP64 {e64:.2e}"
    if storage == "f32":
        v32 = o.fdk(d["raw"], d["ptm"], sc, rc, fp64=False, dark=d["dark"], gain=d["gain"], composite_views=True,
                    z0=z0, nz=nz)
        e32 = rel_rmse(vol[z0:z0 + nz], v32)
        msg += f", vs FP32-ref {e32:.2e}"
        assert e32 <= tol
    print(msg, "max-abs", max_abs(vol[z0:z0 + nz], v64))
    assert e64 <= tol


# --------------------------------------------------------------------------
# Staged layouts of BF16 / FP16 projections
# --------------------------------------------------------------------------

# the dental geometry scaled down: 0.6 mm * (600 / 350) / 0.8 mm = 1.29 rows
# per z voxel, the headline workload's row step (the HI = 1 pair branch)
DENTAL_STEP = configs.DENTAL512.scaled(n_views=60, n_u=192, n_v=192, du=0.8, X=128, Y=128, Z=128, pxy=0.6,
                                       pz=0.6, name="dental-step-1.29")


@pytest.mark.parametrize("dtype", ["bf16", "f16"])
@pytest.mark.parametrize("tile", [(8, 4), (16, 4), (32, 4), (16, 2), (32, 1)])
@pytest.mark.parametrize("geom", ["cbct-step-0.83", "dental-step-1.29"])
def test_bp_strip_layout_equals_quad_layout_bit_exact(gpu, dtype, tile, geom):
    """The column-strip layout's paired gathers read the same four values per
    sample and interpolate them with the same operations as the quad layout's
    one gather per sample: the volumes agree to the bit (one-shot and
    streamed), including the columns that leave the detector rows (partial
    path), and so does the LDS-staged strip fetch.  Row steps 0.83 (second sample zero or
    one row down) and 1.29 (one or two rows down)."""
    from jenhine_amd.recon import BPCore
    w = SMALL if geom.startswith("cbct") else DENTAL_STEP
    step = w.ptms()[0][6] * w.pz / w.ptms()[0][11]
    assert (step < 1.0) == geom.startswith("cbct") and 0.5 < step < 1.984375
    filt = torch.from_numpy(synth.normal(31, (w.n_views, w.n_v, w.n_u))).to(gpu)
    bp = BPCore(w.ptms(), w.n_u, w.n_v, w.X, w.Y, w.Z, w.pxy, w.pz, device=gpu)
    assert bp.layout() == "strip"
    bp.set_tile(*tile)
    out = {}
    for layout in ("quad", "strip"):
        bp.set_layout(layout)
        q = bp.stage(filt, dtype)
        # 8 B per pixel; strips pad each row with one zero pixel
        assert q.numel() == w.n_views * w.n_v * (w.n_u + (layout == "strip")) * 8
        vol = bp.run(q, dtype).cpu().numpy()
        bp.reset()
        h = q.numel() // 2
        bp.accumulate(q[:h], dtype, 0, w.n_views // 2)
        bp.accumulate(q[h:], dtype, w.n_views // 2, w.n_views - w.n_views // 2)
        out[layout] = (vol, bp.finalize().cpu().numpy())
        if layout == "strip":
            bp.set_fetch("lds")
            print(tile, geom, "strip fetch in use:", bp.fetch_in_use(dtype))
            vol = bp.run(q, dtype).cpu().numpy()
            bp.reset()
            bp.accumulate(q[:h], dtype, 0, w.n_views // 2)
            bp.accumulate(q[h:], dtype, w.n_views // 2, w.n_views - w.n_views // 2)
            out["lds"] = (vol, bp.finalize().cpu().numpy())
            bp.set_fetch("gather")
    for k in range(2):
        np.testing.assert_array_equal(out["strip"][k], out["quad"][k])
        # the LDS-staged strips: the same windows, the same operations
        np.testing.assert_array_equal(out["lds"][k], out["quad"][k])


def test_bp_layouts_agree_on_a_fine_detector(gpu):
    """A fine detector (3.5 rows per z voxel) rules the paired gathers out:
    every sample takes the one-gather path, from quads or from column strips
    (plan default for geometry class 0), with bit-identical volumes."""
    from jenhine_amd.recon import BPCore
    w = SMALL.scaled(du=0.12, n_u=200, n_v=200, name="fine-detector")
    filt = torch.from_numpy(synth.normal(32, (w.n_views, w.n_v, w.n_u))).to(gpu)
    bp = BPCore(w.ptms(), w.n_u, w.n_v, w.X, w.Y, w.Z, w.pxy, w.pz, device=gpu)
    assert bp.geometry_mode() == 0 and bp.layout() == "strip"
    strip = bp.run(bp.stage(filt, "bf16"), "bf16").cpu().numpy()
    bp.set_layout("quad")
    quad = bp.run(bp.stage(filt, "bf16"), "bf16").cpu().numpy()
    assert np.isfinite(quad).sum() > 1000
    np.testing.assert_array_equal(strip, quad)


def _h16_bits(x, dtype):
    """16-bit patterns of float32 values, rounded to nearest-even."""
    x = np.ascontiguousarray(x, np.float32)
    if dtype == "f16":
        return x.astype(np.float16).view(np.uint16).astype(np.uint32)
    b = x.view(np.uint32).astype(np.uint64)
    return (((b + 0x7FFF + ((b >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint32)


def _staged_reference(filt, uw, dtype, layout, zf):
    """The staged layouts built from their definitions (jhip.h): q = frame * uW
    (after the 1-2-1 v smoothing with clamped neighbours when zf), values of
    rows beyond the frame and of the column right of the last one are 0."""
    N, V, W = filt.shape
    x = filt.astype(np.float32)
    if zf:
        xp = np.concatenate([x[:, :1], x, x[:, -1:]], axis=1)
        x = (xp[:, 1:-1] * np.float32(0.5) + xp[:, :-2] * np.float32(0.25)) + xp[:, 2:] * np.float32(0.25)
    q = np.zeros((N, V + 4, W + 1), np.float32)
    q[:, :V, :W] = x * uw[:, None, :].astype(np.float32)
    if dtype == "f32":
        a0, b0, a1, b1 = q[:, :V, :W], q[:, :V, 1:], q[:, 1:V + 1, :W], q[:, 1:V + 1, 1:]
        return np.stack([a0, b0 - a0, a1, b1 - a1], axis=-1).view(np.uint32)
    h = lambda a: _h16_bits(a, dtype)
    if layout == "strip":
        w0 = h(q[:, 0:V]) | (h(q[:, 1:V + 1]) << 16)
        w1 = h(q[:, 2:V + 2]) | (h(q[:, 3:V + 3]) << 16)
        out = np.stack([w0, w1], axis=-1)
        out[:, :, W] = 0
        return out
    ox = h(q[:, 0:V, :W]) | (h(q[:, 0:V, 1:]) << 16)
    oy = h(q[:, 1:V + 1, :W]) | (h(q[:, 1:V + 1, 1:]) << 16)
    return np.stack([ox, oy], axis=-1)


@pytest.mark.parametrize("dtype,layout", [("f32", "quad"), ("bf16", "quad"), ("bf16", "strip"), ("f16", "quad"),
                                          ("f16", "strip")])
@pytest.mark.parametrize("zf", [False, True])
def test_stage_layouts_bit_exact(gpu, dtype, layout, zf):
    """jh_bp_stage against the layouts' definitions, to the bit, on a frame
    whose width and height are not multiples of the kernel's column block
    (64) and row run (16): edge columns, the strip padding column and the
    rows beyond the frame."""
    from jenhine_amd.recon import BPCore
    w = SMALL.scaled(n_views=5, n_u=50, n_v=70, name="ragged-frame")
    bp = BPCore(w.ptms(), w.n_u, w.n_v, w.X, w.Y, w.Z, w.pxy, w.pz, device=gpu)
    if dtype != "f32":
        bp.set_layout(layout)
    bp.set_zfilter(zf)
    filt = synth.normal(33, (w.n_views, w.n_v, w.n_u)).astype(np.float32)
    got = bp.stage(torch.from_numpy(filt).to(gpu), dtype).cpu().numpy().view(np.uint32)
    ref = _staged_reference(filt, bp.uweights(), dtype, layout, zf)
    assert got.size == ref.size
    np.testing.assert_array_equal(got, ref.reshape(-1))


@pytest.mark.parametrize("dtype", ["bf16", "f16"])
@pytest.mark.parametrize("case", ["views-in-one-row-tile", "multiple-of-128-columns", "composite-2W-to-W"])
def test_filter_run_staged_equals_filter_then_stage(gpu, dtype, case):
    """jh_filter_run_staged writes, to the bit, what jh_filter_run followed by
    jh_bp_stage writes: with frames shorter than the 128-row tile (tiles and
    strips across view boundaries), with a width that is a multiple of the
    128-column tile (the padding column comes from the last column), through
    the composite chain's 2W -> W mirrored filter, and for a view range other
    than the first (pre-weight blocks and uW of those views)."""
    from jenhine_amd.recon import BPCore, FilterCore
    mirror = False
    if case == "views-in-one-row-tile":
        w = SMALL.scaled(n_views=5, n_u=200, n_v=70, name="short-frames")
    elif case == "multiple-of-128-columns":
        w = SMALL.scaled(n_views=5, n_u=256, n_v=150, name="256-columns")
    else:
        w = _offset_geometry(True, n_views=6)
        mirror = True
    N, V, W = w.n_views, w.n_v, w.n_u
    win = 2 * W if mirror else W
    ptm = w.ptms()
    scan = w.scan_params()
    filt = FilterCore(win, V, ptm, scan.DSO, scan.HalfSampleRate, mode="mfma6", device=gpu, w_out=W,
                      y_mirroring=mirror)
    bp = BPCore(ptm, W, V, w.X, w.Y, w.Z, w.pxy, w.pz, device=gpu)
    assert bp.layout() == "strip" and filt.stage_fusable(bp, dtype)
    x = torch.from_numpy(synth.normal(41, (N, V, win))).to(gpu)
    ref = bp.stage(filt.process(x), dtype)
    got = filt.process_staged(x, bp, dtype)
    assert got.numel() == ref.numel() == N * V * (W + 1) * 8
    assert torch.equal(got, ref)
    ref2 = bp.stage(filt.process(x[1:4], view0=2), dtype, view0=2)
    got2 = filt.process_staged(x[1:4], bp, dtype, view0=2)
    assert torch.equal(got2, ref2)


@pytest.mark.parametrize("W,V,mirror", [(256, 150, False), (512, 40, False), (768, 70, False), (1024, 33, False),
                                         (256, 48, True)])
def test_filter_generated_operand_equals_packed(gpu, W, V, mirror):
    """The split-FP16 filter reads its matrix operand from the Toeplitz
    diagonals kept in LDS (eight shifted copies) instead of copying the packed
    matrix per k step: filtered rows, fused staging and the pre-weighted
    operand path (preprocessing written into the operand) are the packed
    form's to the bit -- widths 256 to 1024, frames shorter and longer than
    the 128-row tile, and the composite chain's mirrored 2W -> W filter."""
    from jenhine_amd.recon import BPCore, FilterCore, PreProcessingCore
    if mirror:
        w = _offset_geometry(True, n_views=6, W=W, V=V)
    else:
        w = SMALL.scaled(n_views=5, n_u=W, n_v=V, name=f"filter-{W}")
    N = w.n_views
    win = 2 * W if mirror else W
    ptm = w.ptms()
    scan = w.scan_params()
    filt = FilterCore(win, V, ptm, scan.DSO, scan.HalfSampleRate, mode="mfma16", device=gpu, w_out=W,
                      y_mirroring=mirror)
    assert filt.generated()
    bp = BPCore(ptm, W, V, w.X, w.Y, w.Z, w.pxy, w.pz, device=gpu)
    x = torch.from_numpy(synth.normal(43, (N, V, win))).to(gpu)
    rng = np.random.default_rng(W)
    raw = torch.from_numpy(rng.integers(200, 4000, (N, V, win), dtype=np.uint16).view(np.int16)).to(gpu)
    pre = PreProcessingCore(win, V, device=gpu)
    op = pre.process_operand(raw, filt)
    outs = {}
    for gen in (True, False):
        filt.set_generated(gen)
        rows = filt.process(x)
        staged = filt.process_staged(x, bp, "f16") if not mirror else None
        from_op = filt.process_operand_staged(op, N, bp, "f16") if not mirror else filt.process(x[1:3], view0=1)
        outs[gen] = (rows, staged, from_op)
    for a, b in zip(outs[True], outs[False]):
        if a is not None:
            assert torch.equal(a, b)


def test_filter_generated_operand_refused_where_absent(gpu):
    """Widths that are not a multiple of 32 keep the packed operand only."""
    from jenhine_amd.recon import FilterCore
    w = SMALL.scaled(n_views=3, n_u=200, n_v=40, name="filter-200")
    scan = w.scan_params()
    filt = FilterCore(200, 40, w.ptms(), scan.DSO, scan.HalfSampleRate, mode="mfma16", device=gpu)
    assert not filt.generated()
    with pytest.raises(RuntimeError, match="no generated operand"):
        filt.set_generated(True)
    filt.set_generated(False)


def test_filter_run_staged_refuses_plans_that_do_not_fuse(gpu):
    from jenhine_amd.recon import BPCore, FilterCore
    w = SMALL.scaled(n_views=4, n_u=64, n_v=40, name="small-frames")
    ptm = w.ptms()
    scan = w.scan_params()
    bp = BPCore(ptm, w.n_u, w.n_v, w.X, w.Y, w.Z, w.pxy, w.pz, device=gpu)
    mfma = FilterCore(w.n_u, w.n_v, ptm, scan.DSO, scan.HalfSampleRate, mode="mfma6", device=gpu)
    valu = FilterCore(w.n_u, w.n_v, ptm, scan.DSO, scan.HalfSampleRate, mode="fp32valu", device=gpu)
    assert mfma.stage_fusable(bp, "bf16") and not mfma.stage_fusable(bp, "f32")
    assert not valu.stage_fusable(bp, "bf16")
    bp.set_zfilter(True)
    assert not mfma.stage_fusable(bp, "bf16")
    bp.set_zfilter(False)
    bp.set_layout("quad")
    assert not mfma.stage_fusable(bp, "bf16")
    x = torch.zeros((w.n_views, w.n_v, w.n_u), dtype=torch.float32, device=gpu)
    with pytest.raises(RuntimeError, match="do not fuse"):
        mfma.process_staged(x, bp, "bf16")


@pytest.mark.parametrize("mode", ["mfma6", "mfma16"])
@pytest.mark.parametrize("dtype", ["bf16", "f16"])
@pytest.mark.parametrize("width", [200, 196])  # 16-byte loads, and the per-pixel path (rows not 8-aligned)
def test_pre_operand_path_equals_two_kernel_chain(gpu, dtype, width, mode):
    """jh_pre_run_operand + jh_filter_run_operand_staged against jh_pre_run +
    jh_filter_run_staged, to the bit: a frame width that is not a multiple of
    the 32-wide k step, a row count that is not a multiple of the 16-row
    block, dark and gain correction and beam hardening, a view range other
    than the first; and the operand itself against its definition (pixel
    times pre-weight, the three unweighted edge columns, zeros beyond, in the
    MFMA tile order; the split-FP16 form also keeps each row's largest
    magnitude per 128-column k tile)."""
    from jenhine_amd.recon import BPCore, FilterCore, PreProcessingCore
    w = SMALL.scaled(n_views=5, n_u=width, n_v=70, name="ragged-operand")
    N, V, W = w.n_views, w.n_v, w.n_u
    d = w.data()
    ptm = w.ptms()
    scan = w.scan_params()
    pre = PreProcessingCore(W, V, bright_field=4000.0, bh=(1.0, 0.03, 0.004), device=gpu)
    filt = FilterCore(W, V, ptm, scan.DSO, scan.HalfSampleRate, mode=mode, device=gpu)
    bp = BPCore(ptm, W, V, w.X, w.Y, w.Z, w.pxy, w.pz, device=gpu)
    raw, dark = _u16(d["raw"], gpu), _u16(d["dark"], gpu)
    gain = torch.from_numpy(d["gain"]).to(gpu)
    p = pre.process(raw, dark, gain)
    ref = filt.process_staged(p, bp, dtype)
    a = pre.process_operand(raw, filt, dark, gain)
    got = filt.process_operand_staged(a, N, bp, dtype)
    assert torch.equal(got, ref)
    ref2 = filt.process_staged(p[1:4], bp, dtype, view0=2)
    got2 = filt.process_operand_staged(pre.process_operand(raw[1:4], filt, dark, gain, view0=2), 3, bp, dtype,
                                       view0=2)
    assert torch.equal(got2, ref2)
    # the operand against its definition
    Kp = (W + 3 + 31) // 32 * 32
    nsteps, ntk = Kp // 32, (Kp + 127) // 128
    extra = 16 * ntk if mode == "mfma16" else 0   # the split-FP16 form's k-tile row maxima
    assert filt.operand_bytes(16) == 4 * V * (16 * Kp + extra)  # 16 views: V blocks of 16 rows
    pn = p.cpu().numpy().reshape(N * V, W)
    pw = np.stack([filt.preweight(v).cpu().numpy() for v in range(N)]).reshape(N * V, W)
    rows = N * V
    nb = (rows + 15) // 16
    full = np.zeros((nb * 16, Kp), np.float32)
    full[:rows, :W] = pn * pw
    full[:rows, W:W + 3] = pn[:, 0:3]
    ab = a.cpu().numpy()[: nb * (16 * Kp + extra)].reshape(nb, 16 * Kp + extra)
    # tile order: ((r16 * nsteps + s) * 2 + h) * 64 + 16 g + i, 4 floats
    t = full.reshape(-1, 16, nsteps, 4, 2, 4)           # r16, i, s, g, h, e
    t = t.transpose(0, 2, 4, 3, 1, 5).reshape(nb, -1)   # r16, s, h, g, i, e
    np.testing.assert_array_equal(ab[:, :16 * Kp], t)
    if extra:
        # per block: tile kt's maxima of |value| over columns 128 kt .. 128 kt + 127, 16 rows each
        fp = np.zeros((nb * 16, ntk * 128), np.float32)
        fp[:, :Kp] = np.abs(full)
        want = fp.reshape(nb, 16, ntk, 128).max(axis=3).transpose(0, 2, 1).reshape(nb, -1)
        np.testing.assert_array_equal(ab[:, 16 * Kp:], want)


@pytest.mark.parametrize("mode", ["mfma6", "mfma16"])
@pytest.mark.parametrize("dtype", ["bf16", "f16"])
@pytest.mark.parametrize("right", [True, False])
def test_composite_operand_path_equals_composite_then_filter(gpu, dtype, right, mode):
    """jh_composite_run_operand + jh_filter_run_operand_staged against
    jh_composite_run + jh_filter_run_staged, to the bit, through the 2W -> W
    filter with and without yMirroring: each half of the scan is one call
    and a half ends inside a 16-row block (3 views of 42 rows), so the two
    calls share a block; and a view range other than the first."""
    from jenhine_amd.recon import BPCore, CompositeCore, FilterCore
    w = _offset_geometry(right, n_views=6, W=160, V=42)
    N, V, W = w.n_views, w.n_v, w.n_u
    ptm = w.ptms()
    scan = w.scan_params()
    comp = CompositeCore(W, V, ptm, w.sod, device=gpu)
    assert comp.detector_on_right == right
    filt = FilterCore(2 * W, V, ptm, scan.DSO, scan.HalfSampleRate, mode=mode, device=gpu, w_out=W,
                      y_mirroring=right)
    bp = BPCore(ptm, W, V, w.X, w.Y, w.Z, w.pxy, w.pz, device=gpu)
    assert filt.stage_fusable(bp, dtype)
    frames = torch.from_numpy((synth.normal(43, (N, V, W)) + 2.0).astype(np.float32)).to(gpu)
    c = comp.process(frames)
    ref = filt.process_staged(c, bp, dtype)
    a = comp.process_operand(frames, filt)
    got = filt.process_operand_staged(a, N, bp, dtype)
    assert torch.equal(got, ref)
    # views 1 and 2 (conjugates 4 and 5) alone: an operand that starts at view 1
    half = N // 2
    a2 = comp.process_pairs_operand(frames[1:half], frames[half + 1:], filt, view0=1, aview0=1)
    ref2 = filt.process_staged(c[1:half], bp, dtype, view0=1)
    got2 = filt.process_operand_staged(a2, half - 1, bp, dtype, view0=1)
    assert torch.equal(got2, ref2)


@pytest.mark.parametrize("mode", ["mfma6", "mfma16"])
def test_fdk_composite_operand_path_equals_stage_by_stage(gpu, mode):
    """The offset-detector production chain (pre -> composite into the
    operand -> filter with fused staging -> bp) gives the volume of the
    stage-by-stage chain to the bit."""
    from jenhine_amd.recon import FDK
    w = configs.OFFSET128.scaled(n_views=16)
    d = w.data()
    fdk = FDK(w.scan_params(), w.recon_params(), d["ptm"], mode, "bf16", device=gpu, composite=True)
    assert fdk.composite_operand_path()
    raw, dark = _u16(d["raw"], gpu), _u16(d["dark"], gpu)
    gain = torch.from_numpy(d["gain"]).to(gpu)
    vol = fdk(raw, dark, gain).cpu().numpy()
    ref, _ = fdk(raw, dark, gain, return_stages=True)
    np.testing.assert_array_equal(vol, ref.cpu().numpy())


@pytest.mark.parametrize("mode", ["mfma6", "mfma16"])
def test_fdk_operand_path_equals_stage_by_stage(gpu, mode):
    """The production chain (pre -> operand -> filter with fused staging ->
    bp) gives the volume of the stage-by-stage chain to the bit."""
    from jenhine_amd.recon import FDK
    w = SMALL
    d = small_data_cache()
    sc, rc = w.scan_params(), w.recon_params()
    fdk = FDK(sc, rc, d["ptm"], mode, "bf16", gpu)
    assert fdk.operand_path()
    raw = _u16(d["raw"], gpu)
    vol = fdk(raw).cpu().numpy()
    ref, _ = fdk(raw, return_stages=True)
    np.testing.assert_array_equal(vol, ref.cpu().numpy())
