"""The FDK distance-weighted backprojection (jh_bp_set_weight, JH_WEIGHT_FDK):
vol = scale * sum_v bilinear(q_v) / w_v^2, the weighting the reference
replaces by its count normalisation (BPAgent/BPCore.cpp:147-151,198-205;
BPCUDAAgent/BPCUDACore.cu:62).  Checked against the oracle's restatement
(jo_bp_fdkw, pinned by a closed form in test_oracle.py), against the closed
form itself on the device, in every geometry class and storage type, one-shot
and streamed, and through the full chain.  Tolerances as the count form's:
FP32 <= 1e-4 vs FP32-ref, BF16 <= 1e-3 and FP16 <= 2e-4 vs FP64."""
import numpy as np
import pytest
import torch

from jenhine_amd import configs, geometry, synth
from oracle import oracle as o
from tests.helpers import max_abs, oracle_params, rel_rmse

pytestmark = pytest.mark.gpu

SMALL = configs.CBCT128.scaled(n_views=24, name="cbct128-24v")


def _bp(gpu, ptm, filt, shape, dtype="f32", chunks=None, scale=0.0, tile=None):
    from jenhine_amd.recon import BPCore
    bp = BPCore(ptm, filt.shape[2], filt.shape[1], *shape, device=gpu)
    bp.set_weight("fdk", scale)
    if tile:
        bp.set_tile(*tile)
    f = torch.from_numpy(np.ascontiguousarray(filt, np.float32)).to(gpu)
    if chunks is None:
        return bp.run(bp.stage(f, dtype), dtype).cpu().numpy(), bp
    bp.reset()
    for v0 in range(0, filt.shape[0], chunks):
        n = min(chunks, filt.shape[0] - v0)
        bp.accumulate(bp.stage(f[v0:v0 + n], dtype, view0=v0), dtype, v0, n)
    return bp.finalize().cpu().numpy(), bp


def test_weight_api_defaults(gpu):
    from jenhine_amd.recon import BPCore
    w = SMALL
    bp = BPCore(w.ptms(), w.n_u, w.n_v, 8, 8, 8, w.pxy, w.pz, device=gpu)
    assert bp.weight() == ("count", pytest.approx(2.0 / w.n_views))
    bp.set_weight("fdk", 0.0)
    assert bp.weight() == ("fdk", pytest.approx(2.0 / w.n_views))
    bp.set_weight("fdk", 0.25)
    assert bp.weight() == ("fdk", 0.25)
    from jenhine_amd import _lib
    with pytest.raises(_lib.JhError):
        bp.set_weight("fdk", float("nan"))


def test_fdk_weight_single_view_closed_form_on_device(gpu):
    """One view whose weighted image is affine: every voxel projecting onto
    the detector holds scale * q(su, sv) / w^2 (FP64 from the matrix alone),
    every other voxel 0."""
    W, V = 64, 48
    ptm = geometry.circular_ptms(1, 300.0, 500.0, W, V, 0.5, start_deg=17.0)
    a, b, c = 0.3, 0.01, -0.02
    uw = o.uweights(ptm, W, fp64=True)[0]
    uu, vv = np.meshgrid(np.arange(W), np.arange(V))
    with np.errstate(divide="ignore", invalid="ignore"):
        proj = np.where(uu >= 1, (a + b * uu + c * vv) / uw[uu], 0.0)[None]
    shape = (10, 60, 6, 0.4, 0.4)   # X, Y, Z: y extends past both detector edges
    vol, bp = _bp(gpu, ptm, proj, shape, scale=0.7)
    assert bp.geometry_mode() == 0
    ax = lambda n, p: (np.arange(n) - (n - 1) / 2) * p
    zz, yy, xx = np.meshgrid(ax(6, 0.4), ax(60, 0.4), ax(10, 0.4), indexing="ij")
    uvw = np.c_[xx.ravel(), yy.ravel(), zz.ravel(), np.ones(xx.size)] @ ptm[0].reshape(3, 4).T
    su = (uvw[:, 0] / uvw[:, 2] + 0.5).reshape(vol.shape)
    sv = (uvw[:, 1] / uvw[:, 2] + 0.5).reshape(vol.shape)
    w = uvw[:, 2].reshape(vol.shape)
    # away from the detector edges by more than float rounding of (su, sv)
    inside = (su >= 1.001) & (su <= W - 1.001) & (sv >= 0.001) & (sv <= V - 1.001)
    outside = (su < -0.001) | (su > W - 0.999) | (sv < -0.001) | (sv > V - 0.999)
    assert inside.sum() > 500 and outside.sum() > 100
    exp = 0.7 * (a + b * su + c * sv) / w ** 2
    np.testing.assert_allclose(vol[inside], exp[inside], rtol=2e-6, atol=1e-7)
    assert (vol[outside] == 0).all()


@pytest.mark.parametrize("chunks", [None, 5])
def test_fdk_weight_fp32_vs_fp32_ref(gpu, chunks):
    w = SMALL
    ptm = w.ptms()
    filt = synth.normal(41, (w.n_views, w.n_v, w.n_u))
    vol, _ = _bp(gpu, ptm, filt, (w.X, w.Y, w.Z, w.pxy, w.pz), chunks=chunks)
    ref = o.backproject(filt, ptm, (w.Z, w.Y, w.X), w.pxy, w.pz, fdk_scale=2.0 / w.n_views)
    e = rel_rmse(vol, ref)
    print("fdk weight fp32", chunks, "rel-RMSE vs FP32-ref", e, "max-abs", max_abs(vol, ref))
    assert np.isfinite(vol).all()
    assert e <= 1e-4


def test_fdk_weight_general_geometry_modes(gpu):
    P = np.array([-1.33897, -7.81194, 0.01236, 575.79443, -0.97515, -0.03862, 7.51861, 346.26642,
                  -0.00226, -0.00004, 0.0, 1.0], np.float64)
    tilt = P.copy()
    tilt[10] = 0.0003
    for ptm, mode in ((np.stack([P, P * 1.0001]), 1), (np.stack([tilt, P]), 2)):
        ptm = ptm.reshape(-1, 12)
        filt = synth.normal(43, (2, 686, 644))
        vol, bp = _bp(gpu, ptm, filt, (96, 80, 20, 0.25, 1.0), scale=1.0)
        assert bp.geometry_mode() == mode
        ref = o.backproject(filt, ptm.astype(np.float32), (20, 80, 96), 0.25, 1.0, fdk_scale=1.0)
        e = rel_rmse(vol, ref)
        print("fdk weight mode", mode, "rel-RMSE", e)
        assert e <= 3e-4   # as the count form's test: 2 views of white noise


@pytest.mark.parametrize("dtype,tol", [("bf16", 1e-3), ("f16", 2e-4)])
def test_fdk_weight_reduced_storage_vs_fp64(gpu, dtype, tol):
    w = SMALL
    d = w.data()
    sc, rc = oracle_params(w)
    v64, st = o.fdk(d["raw"], d["ptm"], sc, rc, fp64=True, dark=d["dark"], gain=d["gain"], return_stages=True,
                    fdk_scale=2.0 / w.n_views)
    vol, bp = _bp(gpu, d["ptm"], st["filt"].astype(np.float32), (w.X, w.Y, w.Z, w.pxy, w.pz), dtype=dtype)
    assert bp.layout() == "strip"
    e = rel_rmse(vol, v64)
    print("fdk weight", dtype, "rel-RMSE vs FP64", e)
    assert e <= tol


@pytest.mark.parametrize("dtype", ["bf16", "f16"])
def test_fdk_weight_strips_equal_quads_bit_exact(gpu, dtype):
    """The weighted paired-gather path interpolates the same operands as the
    one-gather-per-sample quad path."""
    from jenhine_amd.recon import BPCore
    w = SMALL
    filt = torch.from_numpy(synth.normal(47, (w.n_views, w.n_v, w.n_u))).to(gpu)
    bp = BPCore(w.ptms(), w.n_u, w.n_v, w.X, w.Y, w.Z, w.pxy, w.pz, device=gpu)
    bp.set_weight("fdk")
    out = {}
    for layout in ("quad", "strip"):
        bp.set_layout(layout)
        q = bp.stage(filt, dtype)
        out[layout] = bp.run(q, dtype).cpu().numpy()
    np.testing.assert_array_equal(out["strip"], out["quad"])


@pytest.mark.parametrize("filter_mode,storage,tol64", [("fp32", "f32", 1e-4), ("mfma6", "bf16", 1e-3)])
def test_fdk_weight_chain_cbct128_vs_oracle(gpu, filter_mode, storage, tol64):
    """The whole chain with the distance weight, against the oracle chain with
    the same weight; the count and distance forms are reported side by side
    (informational: they differ by the 1/w^2 factor off the iso-centre)."""
    from jenhine_amd.recon import FDK
    w = configs.CBCT128
    d = w.data()
    sc, rc = oracle_params(w)
    raw = torch.from_numpy(d["raw"].view(np.int16)).to(gpu)
    dark = torch.from_numpy(d["dark"].view(np.int16)).to(gpu)
    gain = torch.from_numpy(d["gain"]).to(gpu)
    fdk = FDK(w.scan_params(), w.recon_params(), d["ptm"], filter_mode, storage, device=gpu, weight="fdk")
    vol = fdk(raw, dark, gain).cpu().numpy()
    v64 = o.fdk(d["raw"], d["ptm"], sc, rc, fp64=True, dark=d["dark"], gain=d["gain"], fdk_scale=2.0 / w.n_views)
    e = rel_rmse(vol, v64)
    cnt = FDK(w.scan_params(), w.recon_params(), d["ptm"], filter_mode, storage, device=gpu)(raw, dark, gain)
    print(filter_mode, storage, "fdk weight rel-RMSE vs FP64", e, "; count vs distance form",
          rel_rmse(cnt.cpu().numpy(), vol))
    assert e <= tol64
