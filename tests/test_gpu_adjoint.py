"""Differentiation form on the GPU: the forward projector A = B^T against the
un-normalised backprojector B, through the C-ABI.

  * dot-product test <A x, y> = <x, B y> on the adjoint256 geometry
    (256^3, 360 views of 512^2) with x, y ~ N(0, 1): relative gap <= 1e-5;
  * the gradient of 0.5 |A x - p|^2 (computed as B (A x - p) chunk by chunk)
    against central finite differences on 64 random voxels and against the
    CPU oracle's gradient (FP64);
  * 20 SIRT-style preconditioned gradient steps with view-chunk
    recomputation: the loss falls at every step.

The torch.autograd.Function below lives in the test harness only; the library
exposes plain C entry points.
"""
import numpy as np
import pytest
import torch

from jenhine_amd import configs, synth
from oracle import oracle as o

pytestmark = pytest.mark.gpu

ADJ = configs.ADJOINT256


class ProjectionLoss(torch.autograd.Function):
    """0.5 |A x - p|^2 summed over view chunks; backward recomputes A x per
    chunk instead of keeping the sinogram."""

    @staticmethod
    def forward(ctx, x, p, bp, chunk):
        ctx.save_for_backward(x, p)
        ctx.bp, ctx.chunk = bp, chunk
        loss = torch.zeros((), dtype=torch.float64, device=x.device)
        for v0 in range(0, bp.n_views, chunk):
            n = min(chunk, bp.n_views - v0)
            r = bp.project(x, v0, n) - p[v0:v0 + n]
            loss += 0.5 * (r.double() ** 2).sum()
        return loss

    @staticmethod
    def backward(ctx, g):
        x, p = ctx.saved_tensors
        bp, chunk = ctx.bp, ctx.chunk
        grad = torch.zeros_like(x)
        for v0 in range(0, bp.n_views, chunk):
            n = min(chunk, bp.n_views - v0)
            r = (bp.project(x, v0, n) - p[v0:v0 + n]).contiguous()
            grad += bp.backproject(bp.pack(r, "f32", v0), v0, n)
        return (g.float() * grad), None, None, None


def _plan(w, dev):
    from jenhine_amd.recon import BPCore
    return BPCore(w.ptms(), w.n_u, w.n_v, w.X, w.Y, w.Z, w.pxy, w.pz, device=dev)


@pytest.mark.parametrize("weighted", [False, True])
def test_dot_product_adjoint256(gpu, weighted):
    bp = _plan(ADJ, gpu)
    assert bp.geometry_mode() == 0
    x = torch.from_numpy(synth.normal(21, (ADJ.Z, ADJ.Y, ADJ.X))).to(gpu)
    y = torch.from_numpy(synth.normal(22, (ADJ.n_views, ADJ.n_v, ADJ.n_u))).to(gpu)
    Ax = bp.project(x, weighted=weighted)
    q = bp.stage(y, "f32") if weighted else bp.pack(y, "f32")
    By = bp.backproject(q, 0, ADJ.n_views)
    lhs = (Ax.double() * y.double()).sum().item()
    rhs = (x.double() * By.double()).sum().item()
    gap = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print("dot test weighted", weighted, lhs, rhs, gap)
    assert gap <= 1e-5


@pytest.mark.parametrize("weighted", [False, True])
def test_dot_product_sparse_volume_adjoint256(gpu, weighted):
    """The dot test with 4096 random voxels set: each detector pixel then sums
    a few terms, so rounding leaves the gap near 1e-7, and one sample that the
    projector placed differently from the backprojector (another row or
    column, or dropped) would show as a gap of about 1e-4."""
    bp = _plan(ADJ, gpu)
    rng = np.random.default_rng(41)
    x = np.zeros((ADJ.Z, ADJ.Y, ADJ.X), np.float32)
    idx = rng.choice(x.size, 4096, replace=False)
    x.reshape(-1)[idx] = rng.standard_normal(4096).astype(np.float32)
    y = torch.from_numpy(synth.normal(26, (ADJ.n_views, ADJ.n_v, ADJ.n_u))).to(gpu)
    xt = torch.from_numpy(x).to(gpu)
    Ax = bp.project(xt, weighted=weighted)
    q = bp.stage(y, "f32") if weighted else bp.pack(y, "f32")
    By = bp.backproject(q, 0, ADJ.n_views)
    lhs = (Ax.double() * y.double()).sum().item()
    rhs = (xt.double() * By.double()).sum().item()
    gap = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print("sparse dot test weighted", weighted, lhs, rhs, gap)
    assert gap <= 2e-6


def test_projector_matches_cpu_oracle_and_general_modes(gpu):
    from jenhine_amd.recon import BPCore
    w = configs.CBCT128.scaled(n_views=5, n_u=48, n_v=40, X=24, Y=20, Z=14, du=1.0)
    x = synth.normal(23, (w.Z, w.Y, w.X))
    for ptm in (w.ptms(), w.ptms() + np.array([0, 0, 0.02, 0, 0, 0, 0, 0, 0, 0, 1e-4, 0])):
        bp = BPCore(ptm, w.n_u, w.n_v, w.X, w.Y, w.Z, w.pxy, w.pz, device=gpu)
        for weighted in (False, True):
            g = bp.project(torch.from_numpy(x).to(gpu), weighted=weighted).cpu().numpy()
            c = o.forward_project(x, ptm, (w.n_v, w.n_u), w.pxy, w.pz, weighted=weighted)
            err = np.sqrt(np.mean((g - c) ** 2)) / np.sqrt(np.mean(c ** 2))
            print("mode", bp.geometry_mode(), "weighted", weighted, "rel-RMSE vs CPU", err)
            assert err <= 1e-5


@pytest.mark.parametrize("du,zc", [(1.0, 8), (0.25, 16), (2.2, 32)])
def test_projector_row_steps_and_z_tiles(gpu, du, zc):
    """Row steps on both sides of one row per z voxel: 0.38 (du 2.2) and 0.83
    (du 1.0), where the projector's lanes take turns adding to the column
    (several samples per row), and 3.3 (du 0.25), each with another z tile
    of the backprojector (the sample positions follow its tiling): against
    the CPU oracle's transpose and by the dot test with the
    backprojector."""
    from jenhine_amd.recon import BPCore
    w = configs.CBCT128.scaled(n_views=7, n_u=64, n_v=96, X=28, Y=24, Z=40, pxy=0.5, pz=0.5, du=du)
    x = synth.normal(24, (w.Z, w.Y, w.X))
    y = synth.normal(25, (w.n_views, w.n_v, w.n_u))
    bp = BPCore(w.ptms(), w.n_u, w.n_v, w.X, w.Y, w.Z, w.pxy, w.pz, device=gpu)
    bp.set_tile(zc, 4)
    assert bp.geometry_mode() == 0
    g = bp.project(torch.from_numpy(x).to(gpu)).cpu().numpy()
    c = o.forward_project(x, w.ptms(), (w.n_v, w.n_u), w.pxy, w.pz, weighted=False)
    err = np.sqrt(np.mean((g - c) ** 2)) / np.sqrt(np.mean(c ** 2))
    yt = torch.from_numpy(y).to(gpu)
    bty = bp.backproject(bp.pack(yt, "f32"), 0, w.n_views).cpu().numpy()
    lhs, rhs = float((g.astype(np.float64) * y).sum()), float((x.astype(np.float64) * bty).sum())
    gap = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print("du", du, "zc", zc, "projector vs CPU", err, "dot gap", gap)
    assert err <= 1e-5 and gap <= 1e-5


@pytest.mark.parametrize("Z,n_v", [(96, 500), (128, 300), (300, 500), (512, 700), (512, 400), (600, 500)])
def test_projector_column_depths(gpu, Z, n_v):
    """Column depths across the projector's sample-register forms (Z up to
    128, 256, 512 per lane set, and the chunked form beyond), whole columns
    inside the detector (Z 128 and 512 on 300 and 700 rows: no lane masked)
    and columns partly off it (Z 512 on 400 rows, Z 600 on 500 rows), at
    1.19 rows per z voxel: by the dot test with the backprojector, by the
    sparse dot test (exact transposition), and against the CPU oracle's FP64
    transpose within twice the CPU FP32 restatement's own distance from it --
    on detectors of 300 to 700 rows the float sample positions alone put
    both at 1.6e-5 to 4.8e-5, the backprojector's own distance."""
    from jenhine_amd.recon import BPCore
    w = configs.CBCT128.scaled(n_views=5, n_u=48, n_v=n_v, X=12, Y=10, Z=Z, pxy=0.5, pz=0.25, du=0.35)
    x = synth.normal(27, (w.Z, w.Y, w.X))
    y = synth.normal(28, (w.n_views, w.n_v, w.n_u))
    bp = BPCore(w.ptms(), w.n_u, w.n_v, w.X, w.Y, w.Z, w.pxy, w.pz, device=gpu)
    assert bp.geometry_mode() == 0
    g = bp.project(torch.from_numpy(x).to(gpu)).cpu().numpy()
    c = o.forward_project(x, w.ptms(), (w.n_v, w.n_u), w.pxy, w.pz, weighted=False)
    err = np.sqrt(np.mean((g - c) ** 2)) / np.sqrt(np.mean(c ** 2))
    bty = bp.backproject(bp.pack(torch.from_numpy(y).to(gpu), "f32"), 0, w.n_views).cpu().numpy()
    lhs, rhs = float((g.astype(np.float64) * y).sum()), float((x.astype(np.float64) * bty).sum())
    gap = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    rng = np.random.default_rng(Z)
    xs = np.zeros_like(x)
    xs.reshape(-1)[rng.choice(xs.size, 64, replace=False)] = rng.standard_normal(64).astype(np.float32)
    gs = bp.project(torch.from_numpy(xs).to(gpu)).cpu().numpy()
    lhs_s = float((gs.astype(np.float64) * y).sum())
    rhs_s = float((xs.astype(np.float64) * bty).sum())
    gap_s = abs(lhs_s - rhs_s) / max(abs(lhs_s), abs(rhs_s))
    c32 = o.forward_project(x, w.ptms(), (w.n_v, w.n_u), w.pxy, w.pz, weighted=False, fp64=False)
    err32 = np.sqrt(np.mean((c32 - c) ** 2)) / np.sqrt(np.mean(c ** 2))
    print("Z", Z, "V", n_v, "projector vs CPU FP64", err, "CPU FP32 vs FP64", err32, "dot gap", gap,
          "sparse dot gap", gap_s)
    assert err <= 2 * err32 and gap <= 1e-5 and gap_s <= 2e-6


def _loss_setup(w, dev, seed=31, noise=0.01):
    d = w.data(noise=False)
    p = d["lineint"]
    p = p + noise * p.max() * synth.normal(seed, p.shape)
    return torch.from_numpy(np.ascontiguousarray(p, np.float32)).to(dev), d


def test_gradient_vs_finite_differences_and_cpu(gpu):
    w = configs.ADJOINT256.scaled(n_views=90, n_u=256, n_v=256, du=0.8, X=96, Y=96, Z=64, pxy=0.5, pz=0.5)
    bp = _plan(w, gpu)
    p, _ = _loss_setup(w, gpu)
    x = torch.from_numpy(0.01 * synth.uniform(32, (w.Z, w.Y, w.X))).to(gpu).requires_grad_(True)
    loss = ProjectionLoss.apply(x, p, bp, 16)
    loss.backward()
    g = x.grad.double().cpu().numpy()
    # central differences: the loss is exactly quadratic, so any step is exact
    # in exact arithmetic; h = 1 keeps the float rounding of A x negligible
    rng = np.random.default_rng(5)
    idx = rng.integers(0, x.numel(), 64)
    h = 1.0
    fd = []
    with torch.no_grad():
        for t in idx:
            e = torch.zeros_like(x).view(-1)
            e[t] = h
            e = e.view_as(x)
            lp = ProjectionLoss.apply(x + e, p, bp, 16).item()
            lm = ProjectionLoss.apply(x - e, p, bp, 16).item()
            fd.append((lp - lm) / (2 * h))
    fd = np.array(fd)
    gi = g.reshape(-1)[idx]
    err = np.abs(fd - gi) / np.maximum(np.abs(gi), 1e-3 * np.abs(gi).max())
    print("FD max rel err", err.max())
    assert err.max() <= 1e-3
    # CPU oracle gradient B (A x - p) in FP64 on a sub-problem of 12 views
    ws = w.scaled(n_views=12)
    bps = _plan(ws, gpu)
    xs = x.detach()
    ps = p[:12].contiguous()
    xs_ = xs.clone().requires_grad_(True)
    ProjectionLoss.apply(xs_, ps, bps, 5).backward()
    xc = xs.double().cpu().numpy()
    r = o.forward_project(xc, ws.ptms(), (ws.n_v, ws.n_u), ws.pxy, ws.pz, weighted=False) - ps.double().cpu().numpy()
    gc = o.bp_acc(r, ws.ptms(), (ws.Z, ws.Y, ws.X), ws.pxy, ws.pz, weighted=False)
    e = np.sqrt(np.mean((xs_.grad.double().cpu().numpy() - gc) ** 2)) / np.sqrt(np.mean(gc ** 2))
    print("gradient rel-RMSE vs CPU FP64", e)
    assert e <= 1e-5


def test_sirt_steps_with_view_chunks(gpu):
    w = configs.ADJOINT256
    bp = _plan(w, gpu)
    p, d = _loss_setup(w, gpu)
    # SIRT preconditioners: R = 1 / (A 1) per pixel, C = 1 / (B 1) per voxel
    ones_v = torch.ones((w.Z, w.Y, w.X), dtype=torch.float32, device=gpu)
    chunk = 45
    rs = []
    for v0 in range(0, w.n_views, chunk):
        rs.append(bp.project(ones_v, v0, min(chunk, w.n_views - v0)))
    R = torch.cat(rs)
    Rinv = torch.where(R > 0, 1.0 / R, torch.zeros_like(R))
    C = bp.backproject(bp.pack(torch.ones_like(p), "f32"), 0, w.n_views)
    Cinv = torch.where(C > 0, 1.0 / C, torch.zeros_like(C))
    x = torch.zeros((w.Z, w.Y, w.X), dtype=torch.float32, device=gpu)
    losses = []
    for it in range(20):
        g = torch.zeros_like(x)
        loss = 0.0
        for v0 in range(0, w.n_views, chunk):
            n = min(chunk, w.n_views - v0)
            r = bp.project(x, v0, n) - p[v0:v0 + n]
            loss += 0.5 * float((r.double() ** 2).sum())
            g += bp.backproject(bp.pack((Rinv[v0:v0 + n] * r).contiguous(), "f32", v0), v0, n)
        losses.append(loss)
        x = x - Cinv * g
    print("SIRT losses", [f"{l:.4g}" for l in losses])
    assert all(b < a for a, b in zip(losses, losses[1:]))
    assert losses[-1] < 0.2 * losses[0]
    ph = synth.voxelize(d["ellipsoids"], (w.Z, w.Y, w.X), w.pxy, w.pz)
    c = np.corrcoef(x.cpu().numpy().ravel(), ph.ravel())[0, 1]
    print("correlation with the phantom", c)
    assert c > 0.8
