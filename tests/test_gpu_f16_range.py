"""FP16 storage with a range exponent (jh_bp_set_f16_exponent): staged
values are q * 2^-k and outputs are scaled back by 2^k.  Both scalings are
exact, so in FP16's normal range the volume is the k = 0 volume to the bit;
a k chosen from the a-priori bounds (jh_pre_output_bound,
jh_filter_output_bound) keeps projections that would overflow FP16 finite.
The bounds themselves are checked against the values they bound."""
import numpy as np
import pytest
import torch

from jenhine_amd import configs, synth
from jenhine_amd._lib import JhError
from jenhine_amd.recon import f16_exponent_for
from oracle import oracle as o
from tests.helpers import rel_rmse

pytestmark = pytest.mark.gpu

SMALL = configs.CBCT128.scaled(n_views=24, X=64, Y=64, Z=48, pxy=0.5, pz=0.5, name="cbct-24v-64")


def _bp(gpu, w):
    from jenhine_amd.recon import BPCore
    return BPCore(w.ptms(), w.n_u, w.n_v, w.X, w.Y, w.Z, w.pxy, w.pz, device=gpu)


def _normal_range_frames(w, seed):
    # |values| in [2^-8, 2^8]: with |k| <= 4 every scaled value stays normal
    rng = np.random.default_rng(seed)
    mag = np.exp2(rng.uniform(-8.0, 8.0, (w.n_views, w.n_v, w.n_u)))
    return (mag * rng.choice([-1.0, 1.0], mag.shape)).astype(np.float32)


@pytest.mark.parametrize("layout", ["strip", "quad"])
@pytest.mark.parametrize("k", [-4, 3])
def test_f16_exponent_is_an_exact_scaling(gpu, layout, k):
    w = SMALL
    f = torch.from_numpy(_normal_range_frames(w, 5)).to(gpu)
    bp = _bp(gpu, w)
    bp.set_layout(layout)
    out = {}
    for e in (0, k):
        bp.set_f16_exponent(e)
        q = bp.pack(f, "f16")
        one = bp.run(q, "f16").cpu().numpy()
        bp.reset()
        h = w.n_views // 2
        bp.accumulate(bp.pack(f[:h], "f16"), "f16", 0, h)
        bp.accumulate(bp.pack(f[h:], "f16", view0=h), "f16", h, w.n_views - h)
        out[e] = (one, bp.finalize().cpu().numpy(), q)
    np.testing.assert_array_equal(out[k][0], out[0][0])
    np.testing.assert_array_equal(out[k][1], out[0][1])
    # the staged bits differ by the scaling only: 2^-k applied exactly
    a = out[0][2].view(torch.float16).float().cpu().numpy()
    b = out[k][2].view(torch.float16).float().cpu().numpy()
    np.testing.assert_array_equal(b, a * np.float32(2.0 ** -k))


def test_f16_exponent_keeps_large_projections_finite(gpu):
    w = SMALL
    filt = synth.normal(9, (w.n_views, w.n_v, w.n_u)) * np.float32(4e5)
    f = torch.from_numpy(filt).to(gpu)
    bp = _bp(gpu, w)
    plain = bp.run(bp.stage(f, "f16"), "f16").cpu().numpy()
    assert not np.isfinite(plain).all()   # FP16 overflows without the exponent
    k = f16_exponent_for(float(np.abs(filt).max()))
    assert k > 0
    bp.set_f16_exponent(k)
    vol = bp.run(bp.stage(f, "f16"), "f16").cpu().numpy()
    ref = o.backproject(filt, w.ptms(), (w.Z, w.Y, w.X), w.pxy, w.pz)
    assert np.isfinite(ref).all() and np.isfinite(vol).all()
    e = rel_rmse(vol, ref)
    print("k", k, "rel-RMSE vs FP32-ref", e)
    assert e <= 2e-3   # white noise, 24 views: FP16 rounding of each sample, little averaging


def test_fused_staging_applies_the_exponent(gpu):
    """Preprocessing into the operand, filter with fused FP16 staging: with an
    exponent, the staged strips are those of filter-then-stage with the same
    exponent (bit-exact), and the volume is the k = 0 volume up to values
    that leave FP16's normal range."""
    from jenhine_amd.recon import FDK
    w = SMALL
    d = w.data()
    raw = torch.from_numpy(d["raw"].view(np.int16)).to(gpu)
    dark = torch.from_numpy(d["dark"].view(np.int16)).to(gpu)
    gain = torch.from_numpy(d["gain"]).to(gpu)
    fdk = FDK(w.scan_params(), w.recon_params(), d["ptm"], "mfma6", "f16", device=gpu)
    assert fdk.operand_path()
    base = None
    for k in (0, 5, -3):
        fdk.bp.set_f16_exponent(k)
        fdk._f16_gain = (gain,)  # keep the exponent set here
        a = fdk.pre.process_operand(raw, fdk.filter, dark, gain)
        q1 = fdk.filter.process_operand_staged(a, w.n_views, fdk.bp, "f16")
        q2 = fdk.bp.stage(fdk.filter.process(fdk.pre.process(raw, dark, gain)), "f16")
        assert torch.equal(q1, q2)
        vol = fdk(raw, dark, gain).cpu().numpy()
        if base is None:
            base = vol
        else:
            # values below FP16's normal range round differently at other k
            assert rel_rmse(vol, base) <= 1e-5


def test_prepare_picks_the_exponent_from_the_bounds(gpu):
    """The preprocessing bound holds every preprocessed value, the filter
    bound every filtered value, and FDK.prepare sets the exponent that maps
    the product into [0, 2^15]."""
    from jenhine_amd.recon import FDK
    w = configs.DENTAL512.scaled(n_views=36, n_u=256, n_v=192, X=96, Y=96, Z=64, pxy=0.4, pz=0.4,
                                 name="dental-small")
    d = w.data()
    gain = torch.from_numpy(d["gain"]).to(gpu)
    fdk = FDK(w.scan_params(), w.recon_params(), d["ptm"], "mfma6", "f16", device=gpu)
    _, st = fdk(torch.from_numpy(d["raw"].view(np.int16)).to(gpu),
                torch.from_numpy(d["dark"].view(np.int16)).to(gpu), gain, return_stages=True)
    pb = fdk.pre.output_bound(float(d["gain"].min()), float(d["gain"].max()))
    fb = fdk.filter.output_bound(pb)
    pmax = float(st["pre"].abs().max())
    fmax = float(st["filt"].abs().max())
    print("preprocessed max", pmax, "bound", pb, "; filtered max", fmax, "bound", fb)
    assert pmax <= pb and fmax <= fb
    assert pb < 100 * pmax            # not vacuous
    k = fdk.bp.f16_exponent()
    assert k == f16_exponent_for(fb) and fb * 2.0 ** -k < 2 ** 15


def test_mixed_exponents_since_reset_are_refused(gpu):
    w = SMALL
    f = torch.from_numpy(_normal_range_frames(w, 6)).to(gpu)
    bp = _bp(gpu, w)
    bp.reset()
    bp.accumulate(bp.pack(f[:4], "f16"), "f16", 0, 4)
    bp.set_f16_exponent(2)
    with pytest.raises(JhError):
        bp.accumulate(bp.pack(f[4:8], "f16", view0=4), "f16", 4, 4)
    bp.reset()
    bp.accumulate(bp.pack(f[4:8], "f16", view0=4), "f16", 4, 4)
