"""Row-scaled hi/lo FP16 filter form (JH_FILTER_MFMA_F16X3) beyond the
parity cases the other forms share:

* every operand row is scaled by its own power of two before the FP16 split,
  so rows of any magnitude FP32 holds keep the form's accuracy (an unscaled
  FP16 split would overflow above 65504 and lose everything below 2^-24);
* the filter from the pre-weighted operand takes each row's scale from the
  operand's per-k-tile row maxima, the filter from FP32 rows from a pass
  over the row: on wide frames (many k tiles, spread over the lanes of a
  row) both must give the same result to the bit.

Reference semantics: FilterAgent/FilterCore.cpp:189-222 (row filter),
FilterAgent/FilterAgent.cpp:129-145 (pre-weight per 9-view block)."""
import numpy as np
import pytest
import torch

from jenhine_amd import geometry, synth
from oracle import oracle as o

pytestmark = pytest.mark.gpu


def _u16(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, np.uint16).view(np.int16)).to(dev)


def test_f16_split_keeps_rows_of_any_magnitude(gpu):
    from jenhine_amd.recon import FilterCore
    n, W, V, dso = 3, 96, 24, 300.0
    ptm = geometry.circular_ptms(n, dso, 500.0, W, V, 0.6)
    base = synth.uniform(5, (n, V, W), 0.0, 3.0)
    mag = 2.0 ** np.linspace(-60.0, 60.0, n * V).reshape(n, V, 1)
    frames = np.ascontiguousarray((base * mag).astype(np.float32))
    H = o.filter_design(W, 0.5, window=1, fp64=True)
    src, blk = o.preweight_schedule(n)
    pws = np.stack([o.prefilter_weight(ptm[s], W, V, dso, fp64=True) for s in src])
    ref = o.filter_rows(frames.astype(np.float64), H, pws, blk, fp64=True).reshape(n * V, W)
    flt = FilterCore(W, V, ptm, dso, 0.5, window=1, mode="mfma16", device=gpu)
    out = flt.process(torch.from_numpy(frames).to(gpu)).cpu().numpy().reshape(n * V, W).astype(np.float64)
    assert np.all(np.isfinite(out))
    err = np.linalg.norm(out - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print("per-row rel error: max", err.max(), "median", np.median(err))
    assert err.max() <= 1e-5


@pytest.mark.parametrize("dtype", ["bf16", "f16"])
def test_wide_frames_operand_row_scales_equal_row_pass(gpu, dtype):
    """W = 2048: 17 k tiles of 128 columns per row."""
    from jenhine_amd.recon import BPCore, FilterCore, PreProcessingCore
    n, W, V, dso = 3, 2048, 20, 300.0
    ptm = geometry.circular_ptms(n, dso, 500.0, W, V, 0.1)
    raw = synth.uniform(9, (n, V, W), 200.0, 3800.0).astype(np.uint16)
    dark = synth.uniform(10, (V, W), 90.0, 110.0).astype(np.uint16)
    gain = synth.uniform(11, (V, W), 0.9, 1.1).astype(np.float32)
    pre = PreProcessingCore(W, V, bright_field=4000.0, bh=(1.0, 0.02, 0.003), device=gpu)
    filt = FilterCore(W, V, ptm, dso, 0.5, window=1, mode="mfma16", device=gpu)
    bp = BPCore(ptm, W, V, 16, 16, 8, 0.2, 0.2, device=gpu)
    assert filt.stage_fusable(bp, dtype)
    r, d, g = _u16(raw, gpu), _u16(dark, gpu), torch.from_numpy(gain).to(gpu)
    ref = filt.process_staged(pre.process(r, d, g), bp, dtype)
    got = filt.process_operand_staged(pre.process_operand(r, filt, d, g), n, bp, dtype)
    assert torch.equal(got, ref)
