"""Pins of the CPU oracle (runs without a GPU).

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails it:
  * the reference's own golden vectors (UTRealFFT, UTPreFilterWeightGenerator,
    UTImgResizer, UTTransformMatrix, UTProjectionMatrixModifier);
  * closed forms (ln of BF*exp(-p), the BH polynomial, a linear projection
    sampled bilinearly, the Ram-Lak response ~ k/L, a constant volume);
  * brute force on tiny inputs (direct DFT, the filter as an explicit matrix);
  * the FP64 reconstruction of an analytic phantom.
"""
import numpy as np
import pytest

from jenhine_amd import configs, geometry, synth
from oracle import oracle as o
from tests.helpers import oracle_params, rel_rmse, ulp_diff_f32

REF_PTM = np.array([-1.31600, -7.78872, -0.02526, 573.39276, -0.95473, -0.03864, 7.52651, 344.78983,
                    -0.00225, -0.00000, 0.00000, 1.00000], np.float32)


# ---------------------------------------------------------------- FFT -------

def test_rfft_golden_n7():
    # UTRealFFT.cpp:91-102 (MKL output; our radix/DFT rounding is within 2 ulp)
    f = o.rfft(np.arange(1, 8, dtype=np.float32))
    exp = np.array([28.0, 0.0, -3.49999976, 7.26782513, -3.50000000, 2.79115677, -3.50000000, 0.798852026],
                   np.float32)
    assert ulp_diff_f32(f, exp).max() <= 2


@pytest.mark.parametrize("n", [7, 6, 100, 101, 256])
def test_rfft_round_trip_and_brute_force(n):
    x = synth.normal(5, (n,)).astype(np.float64)
    ccs = o.rfft(x, fp64=True)
    X = np.fft.rfft(x)  # independent library routine
    np.testing.assert_allclose(ccs[0::2][: n // 2 + 1], X.real, atol=1e-10)
    np.testing.assert_allclose(ccs[1::2][: n // 2 + 1], X.imag, atol=1e-10)
    np.testing.assert_allclose(o.irfft(ccs, n, fp64=True), x, atol=1e-12)
    x32 = x.astype(np.float32)
    np.testing.assert_allclose(o.irfft(o.rfft(x32), n), x32, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("n", [100, 101])
def test_ramlak_response(n):
    # UTRealFFT.cpp:120-192: Re(rFFT(ramlak)) ~ k/n within 0.003, Im ~ 0
    rl = np.zeros(n, np.float32)
    rl[0] = 0.25
    for i in range(1, n // 2 + 1, 2):
        rl[i] = rl[n - i] = -1.0 / (np.float32(np.pi) ** 2 * i * i)
    ccs = o.rfft(rl)
    k = np.arange(n // 2 + 1)
    np.testing.assert_allclose(ccs[0::2][: n // 2 + 1], k / n, atol=0.003)
    assert np.abs(ccs[1::2][1: n // 2 + 1]).max() < 1e-6


# ------------------------------------------------------- filter design ------

def test_filter_design_ramlak_no_window_matches_closed_form():
    # cut-offs (f_N, 2 f_N): no taper below f_N; spline with no knots = 1;
    # H[k] = 2 pi f_N * Re(rFFT(ramlak))[k] ~ 2 pi f_N * k / L
    W, hsr = 256, 1.5
    H = o.filter_design(W, hsr, hsr, 2 * hsr, fp64=True)
    L = o.fft_length(W)
    assert L == 512
    k = np.arange(L // 2 + 1)
    np.testing.assert_allclose(H / (2 * np.pi * hsr), k / L, atol=0.003)
    assert abs(H[0] - 2 * np.pi * hsr * (0.25 + 2 * sum(-1 / (np.pi ** 2 * m * m) for m in range(1, L // 2, 2)))) < 1e-9


def test_filter_design_windows():
    W, hsr = 128, 2.0
    base = o.filter_design(W, hsr, hsr, 2 * hsr, fp64=True)
    cos = o.filter_design(W, hsr, 0.0, 0.0, fp64=True)
    sl = o.filter_design(W, hsr, 0.0, 0.0, window=1, fp64=True)
    L = o.fft_length(W)
    f = np.arange(L // 2 + 1) * hsr * 2 / L
    np.testing.assert_allclose(cos[:-1], base[:-1] * np.cos(f[:-1] / hsr * np.pi / 2), rtol=1e-12, atol=1e-15)
    assert cos[-1] == 0.0  # f = f_N is cut (FilterCore.cpp:103-110)
    a = np.pi * f / (2 * hsr)
    np.testing.assert_allclose(sl[1:], base[1:] * np.sin(a[1:]) / a[1:], rtol=1e-12)
    # taper between start and end
    tp = o.filter_design(W, hsr, hsr / 2, hsr, fp64=True)
    mid = (f > hsr / 2) & (f < hsr)
    np.testing.assert_allclose(tp[mid], base[mid] * (0.5 + 0.5 * np.cos((f[mid] - hsr / 2) / (hsr / 2) * np.pi)),
                               rtol=1e-12)


def test_filter_design_db_adjust_spline_passes_knots():
    W, hsr = 128, 2.0
    L = o.fft_length(W)
    base = o.filter_design(W, hsr, hsr, 2 * hsr, fp64=True)
    # a knot placed exactly on a frequency sample: gain 10^(6/20) there
    k = 40
    fk = k * hsr * 2 / L
    adj = o.filter_design(W, hsr, hsr, 2 * hsr, adj_pts=[fk], adj_db=[6.0], fp64=True)
    assert abs(adj[k] / base[k] - 10 ** (6 / 20)) < 1e-9
    assert abs(adj[0] / base[0] - 1.0) < 1e-12 if base[0] != 0 else True


# ------------------------------------------------------------ row filter ----

def test_filter_rows_equals_explicit_toeplitz_plus_ramp():
    """The FFT filter equals T (x * pw) + edge * r with T[n][m] = h[|n-m|]."""
    W, V = 24, 3
    L = o.fft_length(W)
    R = (L - W) // 2
    H = o.filter_design(W, 1.7, 0.0, 0.0, fp64=True)
    x = synth.normal(9, (1, V, W)).astype(np.float64) + 2.0
    pw = synth.uniform(9, (1, V, L), 0.5, 1.0).astype(np.float64)
    out = o.filter_rows(x, H, pw, [0], fp64=True)[0]
    Hf = np.concatenate([H, H[1:-1][::-1]])
    h = np.real(np.fft.ifft(Hf))
    for r in range(V):
        T = np.array([[h[abs(n - m)] for m in range(W)] for n in range(W)])
        edge = x[0, r, :3].mean()
        rr = np.array([sum((i / R) * h[(n + R - i) % L] for i in range(R)) for n in range(W)])
        exp = T @ (x[0, r] * pw[0, r, :W]) + edge * rr
        np.testing.assert_allclose(out[r], exp, rtol=1e-10, atol=1e-12)


# -------------------------------------------------------- pre-weight --------

def test_prefilter_weight_goldens():
    # UTPreFilterWeightGenerator.cpp:7-48 (float, LAPACK sgesv in the reference)
    P = np.array([-1.33108, -7.78814, -0.01779, 573.81219, -0.96645, -0.02576, 7.50237, 344.55353,
                  -0.00225, 0.0, 0.0, 1.0], np.float32)
    w = o.prefilter_weight(P, 686, 644, 443.0).ravel()
    assert w.size == 2048 * 644
    np.testing.assert_allclose(w[0], 0.977879, rtol=1e-6)
    np.testing.assert_allclose(w[-1], 0.9202493, rtol=1e-6)
    np.testing.assert_allclose(w[429 * 2048 + 591], 1.0, rtol=1e-6)
    w64 = o.prefilter_weight(P.astype(np.float64), 686, 644, 443.0, fp64=True).ravel()
    np.testing.assert_allclose(w64[[0, -1, 429 * 2048 + 591]], [0.977879, 0.9202493, 1.0], rtol=5e-6)
    assert w64.max() == 1.0


def test_prefilter_weight_is_ray_obliquity_on_a_circular_orbit():
    # for an ideal flat detector the back-projected segment length grows as
    # the ray length: weight = SDD / |ray| (closed form), 1 on the central ray
    w = configs.CBCT128
    ptm = w.ptms()[0]
    pw = o.prefilter_weight(ptm, w.n_u, w.n_v, w.sod, fp64=True)
    u = np.arange(o.fft_length(w.n_u)) + 0.5 - w.n_u / 2.0  # PTM coordinate minus the principal point
    v = np.arange(w.n_v) + 0.5 - w.n_v / 2.0
    du = w.du
    dist = np.sqrt(w.sdd ** 2 + (u[None, :] * du) ** 2 + (v[:, None] * du) ** 2)
    exp = w.sdd / dist
    exp = exp / exp.max() * pw.max()
    np.testing.assert_allclose(pw, exp / exp.max(), rtol=2e-3)


# ---------------------------------------------------------- preprocess ------

def test_preprocess_closed_form_and_bh():
    p = np.linspace(0.0, 4.0, 50).astype(np.float64)
    bf = 4000.0
    raw = np.round(bf * np.exp(-p)).astype(np.uint16).reshape(1, 5, 10)
    out = o.preprocess(raw, bright_field=bf, fp64=True).ravel()
    np.testing.assert_allclose(out, -np.log(raw.ravel().astype(np.float64) / bf), rtol=1e-15)
    bh = (1.0, 0.02, 0.003)
    ob = o.preprocess(raw, bright_field=bf, bh=bh, fp64=True).ravel()
    np.testing.assert_allclose(ob, out * (bh[0] + bh[1] * out + bh[2] * out ** 2), rtol=1e-14)
    # the identity polynomial {1,0,0} is skipped exactly
    np.testing.assert_array_equal(o.preprocess(raw, bright_field=bf, bh=(1.0, 0.0, 0.0)),
                                  o.preprocess(raw, bright_field=bf))


def test_preprocess_u16_exhaustive_and_dark_gain():
    raw = np.arange(65536, dtype=np.uint16).reshape(1, 256, 256)
    out = o.preprocess(raw, bright_field=1000.0)
    ref = -np.log(np.maximum(raw.astype(np.float64), 1.0) / 1000.0)
    np.testing.assert_allclose(out[0], ref[0], rtol=3e-7, atol=3e-7)
    dark = (raw[0] // 7).astype(np.uint16)
    gain = synth.uniform(1, (256, 256), 0.9, 1.1)
    od = o.preprocess(raw, bright_field=1000.0, dark=dark, gain=gain, fp64=True)[0]
    exp = -np.log(np.maximum(raw[0].astype(np.int64) - dark, 1).astype(np.float64) / (1000.0 * gain.astype(np.float64)))
    np.testing.assert_allclose(od, exp, rtol=1e-12)


def test_resize_golden():
    # UTImgResizer.cpp:15-53
    img0 = np.ones((8, 8), np.float32)
    img0[1:7, 1:7] = np.array([[2, 3, 4, 2, 3, 4], [3, 4, 5, 3, 4, 5], [4, 5, 6, 4, 5, 6],
                               [5, 6, 7, 5, 6, 7], [6, 7, 8, 6, 7, 8], [7, 8, 9, 7, 8, 9]])
    out = o.resize_linear(img0, (1, 1, 6, 6), (3, 3))
    np.testing.assert_array_equal(out, [[3.0, 3.5, 4.0], [5.0, 5.5, 6.0], [7.0, 7.5, 8.0]])


# ---------------------------------------------------------- geometry --------

def test_uweights_closed_form():
    W = 100
    ptm = geometry.circular_ptms(3, 300.0, 500.0, W, 50, 0.5)
    uw = o.uweights(ptm, W, fp64=True)
    u0 = int(ptm[0, 3] / ptm[0, 11])  # 49 (<= W/2): left taper of 2*u0 samples
    i = np.arange(W)
    exp = np.where(i < 2 * u0, 0.5 - 0.5 * np.cos(i * np.pi / (2 * u0)), 1.0)
    np.testing.assert_allclose(uw[0], exp, rtol=1e-15)
    # right-offset detector: taper at the right end
    sh = geometry.move_u(ptm, -30.0)
    u0 = int(sh[0, 3] / sh[0, 11])
    nT = 2 * (W - u0)
    exp = np.ones(W)
    exp[W - nT:] = 0.5 + 0.5 * np.cos(np.arange(nT) * np.pi / nT)
    np.testing.assert_allclose(o.uweights(sh, W, fp64=True)[0], exp, rtol=1e-15)


def test_ptm_modifiers_invariants():
    # UTProjectionMatrixModifier.cpp:14-218: cube corners move as expected
    P = REF_PTM.astype(np.float64)
    xyz = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float64)
    u0, v0 = geometry.project(P, xyz)
    u1, v1 = geometry.project(geometry.move_u(P, 2.0), xyz)
    np.testing.assert_allclose(u0, u1 + 2.0, rtol=1e-12)
    np.testing.assert_allclose(v0, v1, rtol=1e-12)
    u1, v1 = geometry.project(geometry.move_v(P, 2.0), xyz)
    np.testing.assert_allclose(v0, v1 + 2.0, rtol=1e-12)
    u1, v1 = geometry.project(geometry.reverse_x(P), xyz * [-1, 1, 1])
    np.testing.assert_allclose(u0, u1, rtol=1e-12)
    u1, v1 = geometry.project(geometry.reverse_y(P), xyz * [1, -1, 1])
    np.testing.assert_allclose(v0, v1, rtol=1e-12)
    u1, v1 = geometry.project(geometry.move_z(P, 0.11), xyz - [0, 0, 0.11])
    np.testing.assert_allclose(u0, u1, rtol=1e-12)
    np.testing.assert_allclose(v0, v1, rtol=1e-12)


def test_ptm_adjust_sequence_closed_forms():
    """JEngine's adjustment (JEngine.cpp:139-150) of the reference test matrix:
    the borders shift (u, v) by (-borderLeft, -borderUp), ReverseX / ReverseY
    mirror x / y, MoveZ(-CenterZ) moves the volume origin to z = -CenterZ;
    unadjust_ptm inverts it."""
    P = REF_PTM.astype(np.float64)
    xyz = np.array([[x, y, z] for x in (-3, 2) for y in (-1, 5) for z in (-2, 1)], np.float64)
    u0, v0 = geometry.project(P, xyz)
    A = geometry.adjust_ptm(P, border_left=7, border_up=3, mirroring_y=True, center_z=0.45)
    u1, v1 = geometry.project(A, xyz * [-1, -1, 1] + [0, 0, 0.45])
    np.testing.assert_allclose(u1, u0 - 7.0, rtol=0, atol=1e-9)
    np.testing.assert_allclose(v1, v0 - 3.0, rtol=0, atol=1e-9)
    A = geometry.adjust_ptm(P)
    u1, v1 = geometry.project(A, xyz * [-1, 1, 1])
    np.testing.assert_allclose(u1, u0, rtol=1e-12)
    np.testing.assert_allclose(v1, v0, rtol=1e-12)
    for mir in (False, True):
        B = geometry.unadjust_ptm(geometry.adjust_ptm(P, 7, 3, mir, 0.45), 7, 3, mir, 0.45)
        np.testing.assert_allclose(B, P, rtol=1e-13, atol=1e-13 * np.abs(P).max())


def test_f16_exponent_for_maps_the_bound_into_fp16_range():
    from jenhine_amd.recon import f16_exponent_for
    for b in (1e-9, 3e-3, 0.7, 1.0, 1000.0, 32768.0, 32769.0, 5e8):
        k = f16_exponent_for(b)
        s = b * 2.0 ** -k
        assert 2 ** 14 <= s < 2 ** 15, (b, k, s)
    assert f16_exponent_for(0.0) == 0


def test_circular_ptm_central_ray_and_magnification():
    ptm = geometry.circular_ptms(8, 350.0, 600.0, 768, 768, 0.2)
    for P in ptm:
        u, v = geometry.project(P, [[0.0, 0.0, 0.0]])
        assert abs(u[0] + 0.5 - 383.5) < 1e-9 and abs(v[0] + 0.5 - 383.5) < 1e-9
        u1, v1 = geometry.project(P, [[0.0, 0.0, 1.0]])
        assert abs((v1[0] - v[0]) - 600.0 / 350.0 / 0.2) < 1e-9


# ---------------------------------------------------------- backprojection --

def test_bp_linear_projection_single_view_is_exact():
    """Bilinear sampling of an affine image is exact.  One view whose weighted
    image q = p * uW is affine, q(u, v) = a + b u + c v, accumulates exactly
    q(su, sv) at every voxel, with (su, sv) the projected voxel centre plus
    0.5; the count is uW interpolated linearly at su (checked with numpy's
    own interpolation)."""
    W, V = 64, 48
    ptm = geometry.circular_ptms(1, 300.0, 500.0, W, V, 0.5, start_deg=17.0)
    a, b, c = 0.3, 0.01, -0.02
    uw = o.uweights(ptm, W, fp64=True)[0]
    uu, vv = np.meshgrid(np.arange(W), np.arange(V))
    with np.errstate(divide="ignore", invalid="ignore"):
        proj = np.where(uu >= 1, (a + b * uu + c * vv) / uw[uu], 0.0)[None]
    vol, acc, cnt = o.backproject(proj, ptm, (6, 10, 12), 0.4, 0.4, fp64=True, return_parts=True)
    ax = lambda n, p: (np.arange(n) - (n - 1) / 2) * p
    zz, yy, xx = np.meshgrid(ax(6, 0.4), ax(10, 0.4), ax(12, 0.4), indexing="ij")
    u, v = geometry.project(ptm[0], np.c_[xx.ravel(), yy.ravel(), zz.ravel()])
    su, sv = (u + 0.5).reshape(vol.shape), (v + 0.5).reshape(vol.shape)
    assert np.isfinite(vol).all() and su.min() >= 1.0
    np.testing.assert_allclose(acc, a + b * su + c * sv, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(cnt, np.interp(su, np.arange(W), uw), rtol=1e-12)
    np.testing.assert_allclose(vol, (a + b * su + c * sv) / cnt, rtol=1e-12)


def test_bp_fdk_weight_single_view_closed_form():
    """The distance-weighted form on one view whose weighted image is affine:
    every voxel holds scale * q(su, sv) / w^2 exactly, with w the third
    homogeneous coordinate P[2] . (x, y, z, 1) of the voxel centre, computed
    here from the matrix alone; voxels projecting off the detector hold 0
    (no renormalisation).  A dropped or squared-twice weight, or w taken at
    the iso-centre instead of the voxel, fails."""
    W, V = 64, 48
    ptm = geometry.circular_ptms(1, 300.0, 500.0, W, V, 0.5, start_deg=17.0)
    a, b, c = 0.3, 0.01, -0.02
    uw = o.uweights(ptm, W, fp64=True)[0]
    uu, vv = np.meshgrid(np.arange(W), np.arange(V))
    with np.errstate(divide="ignore", invalid="ignore"):
        proj = np.where(uu >= 1, (a + b * uu + c * vv) / uw[uu], 0.0)[None]
    shape = (6, 60, 10)   # y extends past both detector edges
    vol = o.backproject(proj, ptm, shape, 0.4, 0.4, fp64=True, fdk_scale=0.7)
    ax = lambda n, p: (np.arange(n) - (n - 1) / 2) * p
    zz, yy, xx = np.meshgrid(ax(6, 0.4), ax(60, 0.4), ax(10, 0.4), indexing="ij")
    X4 = np.c_[xx.ravel(), yy.ravel(), zz.ravel(), np.ones(xx.size)]
    uvw = X4 @ ptm[0].reshape(3, 4).T
    su = (uvw[:, 0] / uvw[:, 2] + 0.5).reshape(shape)
    sv = (uvw[:, 1] / uvw[:, 2] + 0.5).reshape(shape)
    w = uvw[:, 2].reshape(shape)
    inside = (su >= 0) & (su <= W - 1) & (sv >= 0) & (sv <= V - 1)
    affine = inside & (su >= 1.0)          # column 0 of the image is 0, not affine
    assert affine.sum() > 500 and (~inside).sum() > 100
    assert np.abs(w - 1).max() > 1e-3   # the weight is not trivially 1
    np.testing.assert_allclose(vol[affine], (0.7 * (a + b * su + c * sv) / w ** 2)[affine], rtol=1e-12, atol=1e-15)
    assert (vol[~inside] == 0).all()


def test_bp_fdk_weight_fp32_matches_fp64():
    w = configs.CBCT128.scaled(n_views=12, X=24, Y=24, Z=8)
    rng = np.random.default_rng(5)
    proj = rng.standard_normal((12, w.n_v, w.n_u))
    v64 = o.backproject(proj, w.ptms(), (8, 24, 24), w.pxy, w.pz, fp64=True, fdk_scale=2.0 / 12)
    v32 = o.backproject(proj.astype(np.float32), w.ptms(), (8, 24, 24), w.pxy, w.pz, fdk_scale=2.0 / 12)
    # zero-mean random data (as the count form's test): the sums cancel
    assert rel_rmse(v32, v64) < 5e-5


def test_bp_constant_projection_gives_constant_volume():
    w = configs.CBCT128.scaled(n_views=16, X=32, Y=32, Z=16)
    proj = np.full((16, w.n_v, w.n_u), 2.5, np.float32)
    vol = o.backproject(proj, w.ptms(), (16, 32, 32), w.pxy, w.pz)
    assert np.isfinite(vol).all()
    np.testing.assert_allclose(vol, 2.5, rtol=1e-6)


def test_bp_out_of_field_is_nan():
    # a voxel no view sees has count 0 -> 0/0 (BPCore.cpp:198-205)
    ptm = geometry.circular_ptms(4, 300.0, 500.0, 16, 16, 0.5)
    vol = o.backproject(np.ones((4, 16, 16)), ptm, (3, 40, 40), 1.0, 1.0, fp64=True)
    assert np.isnan(vol[:, 0, 0]).all() and np.isfinite(vol[:, 20, 20]).all()


def test_bp_fp32_matches_fp64_on_random_input():
    w = configs.CBCT128.scaled(n_views=12, X=24, Y=24, Z=8)
    proj = synth.normal(3, (12, w.n_v, w.n_u))
    v32 = o.backproject(proj, w.ptms(), (8, 24, 24), w.pxy, w.pz)
    v64 = o.backproject(proj.astype(np.float64), w.ptms(), (8, 24, 24), w.pxy, w.pz, fp64=True)
    # zero-mean random data: the sums cancel, so the relative error is larger
    # than on a phantom (7e-6 there)
    assert rel_rmse(v32, v64) < 5e-5


def test_bp_slices_are_independent():
    w = configs.CBCT128.scaled(n_views=8, X=16, Y=16, Z=12)
    proj = synth.normal(4, (8, w.n_v, w.n_u))
    full = o.backproject(proj, w.ptms(), (12, 16, 16), w.pxy, w.pz)
    part = o.backproject(proj, w.ptms(), (12, 16, 16), w.pxy, w.pz, z0=5, nz=4)
    np.testing.assert_array_equal(full[5:9], part)


# ---------------------------------------------------------- end to end ------

def test_fdk_fp64_reconstructs_the_phantom():
    w = configs.CBCT128.scaled(n_views=96, X=64, Y=64, Z=16, pxy=0.5, pz=0.5)
    d = w.data(noise=False)
    sc, rc = oracle_params(w)
    vol = o.fdk(d["raw"], d["ptm"], sc, rc, fp64=True, dark=d["dark"], gain=d["gain"])
    ph = synth.voxelize(d["ellipsoids"], (w.Z, w.Y, w.X), w.pxy, w.pz)
    c = np.corrcoef(vol.ravel(), ph.ravel())[0, 1]
    assert c > 0.9, c
    assert abs(vol.mean() / ph.mean() - 1.0) < 0.05
    v32 = o.fdk(d["raw"], d["ptm"], sc, rc, fp64=False, dark=d["dark"], gain=d["gain"])
    assert rel_rmse(v32, vol) < 1e-4


# ---------------------------------------------------------- adjoint pair ----

@pytest.mark.parametrize("weighted", [0, 1])
def test_oracle_projector_is_transpose_of_backprojector(weighted):
    w = configs.CBCT128.scaled(n_views=6, n_u=40, n_v=36, X=20, Y=18, Z=12, du=1.2)
    ptm = w.ptms()
    x = synth.normal(1, (12, 18, 20)).astype(np.float64)
    y = synth.normal(2, (6, 36, 40)).astype(np.float64)
    Ax = o.forward_project(x, ptm, (36, 40), w.pxy, w.pz, weighted=weighted)
    By = o.bp_acc(y, ptm, (12, 18, 20), w.pxy, w.pz, weighted=weighted)
    a, b = (Ax * y).sum(), (x * By).sum()
    assert abs(a - b) <= 1e-12 * abs(a)
    # bp_acc with weights equals the accumulator of the normalised backprojector
    _, acc, _ = o.backproject(y, ptm, (12, 18, 20), w.pxy, w.pz, fp64=True, return_parts=True)
    if weighted:
        np.testing.assert_allclose(By, acc, rtol=1e-12, atol=1e-12)


# ------------------------------------------------- ZFilter, CT numbers -------

@pytest.mark.parametrize("fp64", [False, True])
def test_zfilter_impulse_and_edges(fp64):
    # ZFilterAgent.cpp:49-71: kernel (0.25, 0.5, 0.25) along v, edge rows
    # replicated; an impulse in an interior row and in each edge row
    V, W = 7, 3
    x = np.zeros((3, V, W))
    x[0, 3, 1] = 8.0
    x[1, 0, 0] = 8.0
    x[2, V - 1, 2] = 8.0
    z = o.zfilter(x, fp64=fp64)
    exp = np.zeros_like(x)
    exp[0, 2:5, 1] = [2.0, 4.0, 2.0]
    exp[1, 0, 0], exp[1, 1, 0] = 6.0, 2.0          # 0.5*8 + 0.25*8 (clamped) ; 0.25*8
    exp[2, V - 1, 2], exp[2, V - 2, 2] = 6.0, 2.0
    np.testing.assert_array_equal(z, exp)


def test_zfilter_constant_and_linear_in_v():
    # a constant frame is unchanged; a ramp linear in v is unchanged away from
    # the edges (the kernel is symmetric and sums to 1), and the edges see the
    # replicated row: row 0 -> 0.75 r0 + 0.25 r1
    V, W = 9, 4
    c = np.full((1, V, W), 3.25, np.float32)
    np.testing.assert_array_equal(o.zfilter(c), c)
    r = np.tile(np.arange(V, dtype=np.float64)[None, :, None] * 2.0 + 1.0, (1, 1, W))
    z = o.zfilter(r, fp64=True)
    np.testing.assert_allclose(z[0, 1:-1], r[0, 1:-1], rtol=0, atol=1e-12)
    np.testing.assert_allclose(z[0, 0], 0.75 * r[0, 0] + 0.25 * r[0, 1], rtol=0, atol=1e-12)
    np.testing.assert_allclose(z[0, -1], 0.75 * r[0, -1] + 0.25 * r[0, -2], rtol=0, atol=1e-12)


def test_zfilter_fp32_matches_fp64():
    x = synth.normal(31, (2, 40, 24))
    a = o.zfilter(x)
    b = o.zfilter(x.astype(np.float64), fp64=True)
    assert np.abs(a - b).max() <= 2e-7 * np.abs(b).max()


@pytest.mark.parametrize("fp64", [False, True])
def test_ctnum_closed_forms(fp64):
    # CTNumCore.cpp:22-36 with norm0 = 1: water -> 0, air -> -1000; above 300
    # compressed by norm1; saturated at 8191; NaN (no view) stays NaN
    mu = 0.02
    x = np.array([0.0, mu, mu * 1.2, mu * 1.3, mu * 1.9, mu * 100.0, np.nan])
    y = o.ctnum(x, 1.0, 0.5, mu, fp64=fp64)
    exp = [-1000.0, 0.0, 200.0, 300.0, 300.0 + 600.0 * 0.5, 8191.0]
    np.testing.assert_allclose(y[:6], exp, rtol=0, atol=1e-3 if not fp64 else 1e-9)
    assert np.isnan(y[6])


def test_ctnum_inverts_the_reference_metal_threshold():
    # ConfigHelper.cpp:139-142 maps a threshold in CT numbers (> 300) back to an
    # attenuation through the inverse of CTNumCore's map: the forward map must
    # return the threshold
    norm0, norm1, mu = 1.07, 0.62, 0.02
    for hu in (600.0, 2500.0, 4000.0):
        h = (hu - 300.0) / norm1 + 300.0
        h = (h + 1000.0) / norm0 - 1000.0
        att = (h + 1000.0) / 1000.0 * mu
        got = o.ctnum(np.array([att]), norm0, norm1, mu, fp64=True)[0]
        assert abs(got - hu) < 1e-9 * hu


# ------------------------------------------ offset-detector composite -------

def _offset_case(u_shift, n_views=16, W=96, V=24):
    w = configs.OFFSET128.scaled(n_views=n_views, n_u=W, n_v=V, u_shift=u_shift)
    return w, w.ptms()


@pytest.mark.parametrize("mirror", [False, True])
def test_filter_mirroring_is_the_reflection_of_the_plain_filter(mirror):
    """yMirroring (FilterCore.cpp:45-59,205-219) on 2W columns equals the plain
    filter of the reversed row, reversed back: the circular convolution with
    an even kernel commutes with the reflection m -> (Win-1-m) mod L, which
    maps the ramp after the data onto the ramp before it.  Pins the ramp
    direction, the edge columns and the output offset."""
    Win, Wout, V = 40, 20, 3
    L = o.fft_length(Win)
    H = o.filter_design(Win, 1.3, 0.0, 0.0, fp64=True)
    x = synth.normal(21, (1, V, Win)).astype(np.float64) + np.linspace(1.0, 3.0, Win)
    pw = synth.uniform(21, (1, V, L), 0.5, 1.0).astype(np.float64)
    out = o.filter_rows(x, H, pw, [0], fp64=True, w_out=Wout, mirror=mirror)[0]
    full = o.filter_rows(x, H, pw, [0], fp64=True, w_out=Win, mirror=False)[0]
    if not mirror:
        np.testing.assert_allclose(out, full[:, :Wout], rtol=0, atol=1e-13)
        return
    pw_r = pw.copy()
    pw_r[0, :, :Win] = pw[0, :, :Win][:, ::-1]
    ref = o.filter_rows(x[:, :, ::-1].copy(), H, pw_r, [0], fp64=True, w_out=Wout, mirror=False)[0][:, ::-1]
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    # and the edge really is the last three samples: changing x[0..2] leaves
    # the ramp (hence the output far from the data start) unchanged only
    # through the convolution, not through the edge term
    x2 = x.copy()
    x2[0, :, -3:] += 5.0
    out2 = o.filter_rows(x2, H, pw, [0], fp64=True, w_out=Wout, mirror=True)[0]
    assert np.abs(out2 - out).max() > 1e-3


def test_detector_side_rule():
    # JEngine.cpp:153-161: originU < W/2 -> detector on the right
    _, P = _offset_case(30.0)
    assert o.detector_on_right(P[0], 96)
    _, P = _offset_case(-30.0)
    assert not o.detector_on_right(P[0], 96)
    _, P = _offset_case(0.0)  # centred: originU = W/2 - 1 -> right
    assert o.detector_on_right(P[0], 96)


@pytest.mark.parametrize("u_shift", [30.0, -30.0])
def test_composite_circular_orbit_closed_form(u_shift):
    """On a circular orbit the pseudo-depth-0 plane is the iso-centre plane
    (s23 / s20 = 1 / DSO^3 mm from it), so the conjugate view sees the
    virtual detector point (u1, v1) at u2 = 2 cu - u1, v2 = v1 (cu: the
    principal point), up to ~1e-5 px from that plane offset.  With an
    integer 2 cu the composite's conjugate half is the conjugate frame read
    at integer pixels (a random frame, so a one-pixel slip shows as O(1)).  Pins u1's sign and
    offset, the conjugate view index (v + N/2) % N, the TM inverse and the
    half placement."""
    w, P = _offset_case(u_shift)
    N, W, V = w.n_views, w.n_u, w.n_v
    pre = synth.normal(5, (N, V, W)).astype(np.float64)
    right = o.detector_on_right(P[0], W)
    out = o.composite(pre, P, w.sod, right, fp64=True)
    cu = P[0][3] / P[0][11]
    assert abs(2 * cu - round(2 * cu)) < 1e-9
    for v in (0, 3, N - 1):
        c_view = (v + N // 2) % N
        conj_half = out[v][:, :W] if right else out[v][:, W:]
        dirc_half = out[v][:, W:] if right else out[v][:, :W]
        np.testing.assert_array_equal(dirc_half, pre[v])
        checked = 0
        for c in range(W):
            u1 = c - W if right else c + W
            u2 = int(round(2 * cu)) - u1
            in_trans = (c >= W - 40) if right else (c < 40)
            if in_trans:
                continue
            if 0 <= u2 <= W - 1:
                np.testing.assert_allclose(conj_half[:, c], pre[c_view][:, u2], rtol=0, atol=1e-4)
                checked += 1
            elif u2 < -1 or u2 > W:
                assert np.all(conj_half[:, c] == 0.0)
        assert checked > 10


@pytest.mark.parametrize("right", [True, False])
def test_composite_transition_closed_form(right):
    """Constant frames: direct value a, conjugate value b.  The 40 columns next
    to the direct half are b * maskConj + edge * maskDirc with edge = a, the
    masks 0.5 -/+ 0.5 cos(i pi / 40) (CompositeCore.cpp:202-231); with a ramp
    in the direct frame the edge is the mean of its 5 border pixels."""
    W, V, N = 96, 8, 16
    w = configs.OFFSET128.scaled(n_views=N, n_u=W, n_v=V, u_shift=40.0 if right else -40.0)
    P = w.ptms()
    a, b = 2.0, 5.0
    pre = np.empty((N, V, W))
    pre[:] = b
    pre[0] = a + 0.01 * np.arange(W)[None, :]  # direct view 0 with a ramp in u
    out = o.composite(pre, P, w.sod, right, fp64=True)[0]
    md, mc = o.composite_masks(right, fp64=True)
    i = np.arange(40)
    np.testing.assert_allclose(md, 0.5 - 0.5 * np.cos(i * np.pi / 40) if right else 0.5 + 0.5 * np.cos(i * np.pi / 40),
                               atol=1e-15)
    np.testing.assert_allclose(md + mc, 1.0, atol=1e-15)
    edge = (a + 0.01 * (np.arange(5) if right else np.arange(W - 5, W))).mean()
    trans = out[:, W - 40:W] if right else out[:, W:W + 40]
    conj_sample = trans - edge * md[None, :]
    # the conjugate half reads the constant b wherever the sample is inside
    inside = np.abs(conj_sample) > 1e-12
    np.testing.assert_allclose(conj_sample[inside], (b * mc[None, :].repeat(V, 0))[inside], rtol=1e-12)
    assert inside[:, :30].all() if right else inside[:, 10:].all()


def test_composite_fp32_matches_fp64():
    w, P = _offset_case(30.0)
    pre = (synth.normal(6, (w.n_views, w.n_v, w.n_u)) + 3.0).astype(np.float64)
    c64 = o.composite(pre, P, w.sod, True, fp64=True)
    c32 = o.composite(pre.astype(np.float32), P.astype(np.float32), w.sod, True, fp64=False)
    assert rel_rmse(c32, c64) < 1e-5


def test_fdk_offset_detector_reconstructs_the_phantom():
    """End to end: the composite chain (JEngine.cpp:152-238) reconstructs the
    analytic phantom from a half-fan scan better than the plain chain on the
    same data."""
    w = configs.OFFSET128.scaled(n_views=96, X=64, Y=64, Z=8, pxy=0.5, pz=0.5)
    d = w.data(noise=False)
    sc, rc = oracle_params(w)
    ph = synth.voxelize(d["ellipsoids"], (w.Z, w.Y, w.X), w.pxy, w.pz)
    vol = o.fdk(d["raw"], d["ptm"], sc, rc, fp64=True, dark=d["dark"], gain=d["gain"], composite_views=True)
    c = np.corrcoef(vol.ravel(), ph.ravel())[0, 1]
    assert c > 0.9, c
    assert abs(vol.mean() / ph.mean() - 1.0) < 0.1
    # the plain chain on the same half-fan data: the uW taper alone handles
    # the redundancy, but every row is filtered truncated at the detector
    # edge, which leaves a low-frequency bias the composite removes
    plain = o.fdk(d["raw"], d["ptm"], sc, rc, fp64=True, dark=d["dark"], gain=d["gain"])
    e_comp, e_plain = rel_rmse(vol, ph), rel_rmse(plain, ph)
    assert e_comp < 0.7 * e_plain, (e_comp, e_plain)
    v32 = o.fdk(d["raw"], d["ptm"], sc, rc, fp64=False, dark=d["dark"], gain=d["gain"], composite_views=True)
    assert rel_rmse(v32, vol) < 1e-4
