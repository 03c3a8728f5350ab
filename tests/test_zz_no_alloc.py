"""jh_bp_accumulate_noalloc, the allocation-free variant of the streaming
backprojection (jh_bp_reset / jh_bp_accumulate / jh_bp_finalize), called
through the C-ABI: its accumulators live in the library's static device
array.  Two sequences in a row, the second on a smaller volume with fewer
views, each equal to the plan-buffer entry points' volume to the bit; the
size limit and the sequence rules are refused with an error."""
import ctypes

import numpy as np
import pytest
import torch

from jenhine_amd import _lib, configs, synth
from jenhine_amd._lib import call, DTYPES

pytestmark = pytest.mark.gpu


def _s():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _noalloc(bp, chunks, dtype, flags_last=True, out=None):
    """Accumulate the staged chunks [(q, view0, n)] in one sequence."""
    vol = out if out is not None else torch.full((bp.Z, bp.Y, bp.X), -7.0, dtype=torch.float32, device="cuda")
    for i, (q, v0, n) in enumerate(chunks):
        flags = (_lib.JH_NOALLOC_FIRST if i == 0 else 0) | (_lib.JH_NOALLOC_LAST if i == len(chunks) - 1 else 0)
        call("jh_bp_accumulate_noalloc", bp._h, _p(q), DTYPES[dtype], v0, n, flags, _p(vol), _s())
    return vol


def _plan_buffers(bp, chunks, dtype):
    bp.reset()
    for q, v0, n in chunks:
        bp.accumulate(q, dtype, v0, n)
    return bp.finalize()


def _case(gpu, w, dtype, chunk, seed, weight="count"):
    from jenhine_amd.recon import BPCore
    bp = BPCore(w.ptms(), w.n_u, w.n_v, w.X, w.Y, w.Z, w.pxy, w.pz, device=gpu)
    if weight != "count":
        bp.set_weight(weight)
    f = torch.from_numpy(synth.normal(seed, (w.n_views, w.n_v, w.n_u))).to(gpu)
    chunks = []
    for v0 in range(0, w.n_views, chunk):
        n = min(chunk, w.n_views - v0)
        chunks.append((bp.stage(f[v0:v0 + n], dtype, view0=v0), v0, n))
    return bp, chunks


@pytest.mark.parametrize("dtype,weight", [("bf16", "count"), ("f32", "count"), ("f16", "fdk")])
def test_noalloc_equals_plan_buffers_on_two_calls(gpu, dtype, weight):
    big = configs.CBCT128.scaled(n_views=40, X=128, Y=128, Z=96, name="noalloc-big")
    small = configs.CBCT128.scaled(n_views=24, X=64, Y=48, Z=40, pxy=0.5, pz=0.5, name="noalloc-small")
    for w, chunk, seed in ((big, 16, 61), (small, 10, 62)):
        bp, chunks = _case(gpu, w, dtype, chunk, seed, weight)
        got = _noalloc(bp, chunks, dtype)
        ref = _plan_buffers(bp, chunks, dtype)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(got.cpu().numpy(), ref.cpu().numpy(), err_msg=w.name)


def test_noalloc_refuses_oversize_and_broken_sequences(gpu):
    from jenhine_amd.recon import BPCore
    w = configs.CBCT128.scaled(n_views=8)
    bp = BPCore(w.ptms(), w.n_u, w.n_v, 256, 256, 129, 0.25, 0.25, device=gpu)  # 8.45 M voxels > 8 Mi
    vol = torch.empty((129, 256, 256), dtype=torch.float32, device=gpu)
    with pytest.raises(_lib.JhError, match="exceeds"):
        call("jh_bp_accumulate_noalloc", bp._h, None, DTYPES["f32"], 0, 0, 3, _p(vol), _s())
    ok = BPCore(w.ptms(), w.n_u, w.n_v, 32, 32, 16, 0.5, 0.5, device=gpu)
    # continuing without an open sequence
    call("jh_bp_accumulate_noalloc", ok._h, None, DTYPES["f32"], 0, 0, 3, _p(vol), _s())  # FIRST|LAST closes it
    with pytest.raises(_lib.JhError, match="no open sequence"):
        call("jh_bp_accumulate_noalloc", ok._h, None, DTYPES["f32"], 0, 0, 0, None, _s())
    with pytest.raises(_lib.JhError, match="unknown flags"):
        call("jh_bp_accumulate_noalloc", ok._h, None, DTYPES["f32"], 0, 0, 4, None, _s())
    torch.cuda.synchronize()
