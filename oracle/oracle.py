"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of bench.py
import this module, and only to check results against; the shipped library never does.
Every function here is a thin wrapper around ``libjh_oracle.so`` (built from
jh_oracle.c by the Makefile beside it).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libjh_oracle.so")
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _dt(fp64):
    return np.float64 if fp64 else np.float32


def _sfx(fp64):
    return "_f64" if fp64 else "_f32"


def _real(fp64):
    return ctypes.c_double if fp64 else ctypes.c_float


def fft_length(W):
    return int(lib().jo_fft_length(ctypes.c_int(W)))


def axis(n, pitch, fp64=False):
    out = np.zeros(n, _dt(fp64))
    getattr(lib(), "jo_axis" + _sfx(fp64))(ctypes.c_int(n), _real(fp64)(pitch), _p(out))
    return out


def uweights(ptm, W, fp64=False):
    ptm = np.ascontiguousarray(ptm, _dt(fp64)).reshape(-1, 12)
    out = np.zeros((ptm.shape[0], W), _dt(fp64))
    getattr(lib(), "jo_uweights" + _sfx(fp64))(_p(ptm), ctypes.c_int(ptm.shape[0]), ctypes.c_int(W), _p(out))
    return out


def rfft(x, fp64=False):
    x = np.ascontiguousarray(x, _dt(fp64))
    n = x.size
    out = np.zeros(n // 2 * 2 + 2, _dt(fp64))
    getattr(lib(), "jo_rfft" + _sfx(fp64))(ctypes.c_int(n), _p(x), _p(out))
    return out


def irfft(ccs, n, fp64=False):
    ccs = np.ascontiguousarray(ccs, _dt(fp64))
    out = np.zeros(n, _dt(fp64))
    getattr(lib(), "jo_irfft" + _sfx(fp64))(ctypes.c_int(n), _p(ccs), _p(out))
    return out


def resize_linear(src, sroi, dshape, fp64=False):
    """src full image, sroi=(x, y, w, h); returns a dshape=(h, w) image."""
    src = np.ascontiguousarray(src, _dt(fp64))
    x, y, w, h = sroi
    dh, dw = dshape
    out = np.zeros((dh, dw), _dt(fp64))
    base = src.ctypes.data + (y * src.shape[1] + x) * src.itemsize
    getattr(lib(), "jo_resize_linear" + _sfx(fp64))(
        ctypes.c_void_p(base), ctypes.c_int(src.shape[1]), ctypes.c_int(w), ctypes.c_int(h),
        _p(out), ctypes.c_int(dw), ctypes.c_int(dw), ctypes.c_int(dh))
    return out


def filter_design(W, half_sample_rate, cut_start=0.0, cut_end=0.0, adj_pts=(), adj_db=(),
                  window=0, fp64=False):
    L = fft_length(W)
    H = np.zeros(L // 2 + 1, _dt(fp64))
    pts = np.ascontiguousarray(adj_pts, _dt(fp64))
    db = np.ascontiguousarray(adj_db, _dt(fp64))
    assert pts.size == db.size
    R = _real(fp64)
    getattr(lib(), "jo_filter_design" + _sfx(fp64))(
        ctypes.c_int(W), R(half_sample_rate), R(cut_start), R(cut_end),
        _p(pts) if pts.size else None, _p(db) if db.size else None, ctypes.c_int(pts.size),
        ctypes.c_int(window), _p(H))
    return H


def prefilter_weight(ptm, W, V, dso, detector_on_right=False, fp64=False):
    ptm = np.ascontiguousarray(ptm, _dt(fp64)).reshape(12)
    L = fft_length(W)
    out = np.zeros((V, L), _dt(fp64))
    getattr(lib(), "jo_prefilter_weight" + _sfx(fp64))(
        _p(ptm), ctypes.c_int(W), ctypes.c_int(V), ctypes.c_int(int(detector_on_right)),
        _real(fp64)(dso), _p(out))
    return out


def preweight_schedule(nviews):
    nb = (nviews + 8) // 9
    src = np.zeros(nb, np.int32)
    blk = np.zeros(nviews, np.int32)
    n = lib().jo_preweight_schedule(ctypes.c_int(nviews), _p(src), _p(blk))
    assert n == nb
    return src, blk


def preprocess(raw, out_hw=None, borders=(0, 0, 0, 0), bright_field=4000.0, bh=(1.0, 0.0, 0.0),
               dark=None, gain=None, fp64=False, nthreads=8):
    """raw: (N, inH, inW) uint16. borders = (up, down, left, right)."""
    raw = np.ascontiguousarray(raw, np.uint16)
    N, inH, inW = raw.shape
    bu, bd, bl, br = borders
    if out_hw is None:
        out_hw = (inH - bu - bd, inW - bl - br)
    outH, outW = out_hw
    out = np.zeros((N, outH, outW), _dt(fp64))
    bhv = np.ascontiguousarray(bh, np.float64)
    if dark is not None:
        dark = np.ascontiguousarray(dark, np.uint16)
    if gain is not None:
        gain = np.ascontiguousarray(gain, np.float32)
    rc = lib().jo_preprocess(_p(raw), _p(dark), _p(gain), inW, inH, bu, bd, bl, br, outW, outH,
                             ctypes.c_double(bright_field), _p(bhv), _p(out), N, int(fp64), nthreads)
    assert rc == 0
    return out


def filter_rows(frames, H, preweights=None, pw_index=None, fp64=False, nthreads=8, w_out=None, mirror=False):
    """frames: (N, V, Win). preweights: (nblocks, V, L) or None.  w_out and
    mirror are FilterCore's widthOut and yMirroring (default: Win, False)."""
    frames = np.ascontiguousarray(frames, _dt(fp64))
    N, V, W = frames.shape
    w_out = W if w_out is None else int(w_out)
    H = np.ascontiguousarray(H, _dt(fp64))
    out = np.zeros((N, V, w_out), _dt(fp64))
    pw = None if preweights is None else np.ascontiguousarray(preweights, _dt(fp64))
    idx = None if pw_index is None else np.ascontiguousarray(pw_index, np.int32)
    rc = lib().jo_filter_ex(_p(frames), _p(pw), _p(idx), _p(H), W, w_out, int(bool(mirror)), V, N, _p(out),
                            int(fp64), nthreads)
    assert rc == 0
    return out


def detector_on_right(ptm0, n_used_u):
    """JEngine/JEngine.cpp:153-161: the detector is on the right when the
    iso-centre projects left of the detector middle (float arithmetic)."""
    P = np.asarray(ptm0, np.float32).reshape(12)
    origin_u = np.float32(P[3] / P[11])
    return bool(origin_u < np.float32(n_used_u) / np.float32(2.0))


def composite_masks(det_right, fp64=False):
    """Transition masks (dirc, conj) of CompositeCore over 40 columns."""
    md = np.zeros(40, _dt(fp64))
    mc = np.zeros(40, _dt(fp64))
    getattr(lib(), "jo_composite_masks" + _sfx(fp64))(int(bool(det_right)), _p(md), _p(mc))
    return md, mc


def composite(pre, ptm, dso, det_right, fp64=False, nthreads=8):
    """CompositeCore over a full rotation: (N, V, W) -> (N, V, 2W), view v
    composed with its conjugate (v + N/2) % N."""
    pre = np.ascontiguousarray(pre, _dt(fp64))
    N, V, W = pre.shape
    P = np.ascontiguousarray(np.asarray(ptm).reshape(N, 12), _dt(fp64))
    out = np.zeros((N, V, 2 * W), _dt(fp64))
    rc = lib().jo_composite(_p(pre), _p(P), N, W, V, ctypes.c_double(float(np.float32(dso)) if not fp64 else dso),
                            int(bool(det_right)), _p(out), int(fp64), nthreads)
    assert rc == 0
    return out


def backproject(proj, ptm, vol_shape, pitch_xy, pitch_z, z0=0, nz=None, fp64=False, nthreads=8,
                return_parts=False, fdk_scale=None):
    """proj: (N, V, W) filtered projections; vol_shape = (Z, Y, X).
    Returns vol[z0:z0+nz] (and acc, count when return_parts).  With
    fdk_scale the distance-weighted form: vol = fdk_scale * sum_v
    remap(q_v) / w^2 (jo_bp_fdkw; count returned as zeros)."""
    proj = np.ascontiguousarray(proj, _dt(fp64))
    N, V, W = proj.shape
    ptm = np.ascontiguousarray(ptm, _dt(fp64)).reshape(N, 12)
    Z, Y, X = vol_shape
    if nz is None:
        nz = Z - z0
    vol = np.zeros((nz, Y, X), _dt(fp64))
    acc = np.zeros_like(vol)
    cnt = np.zeros_like(vol)
    R = _real(fp64)
    if fdk_scale is not None:
        rc = getattr(lib(), "jo_bp_fdkw" + _sfx(fp64))(
            _p(proj), _p(ptm), N, W, V, X, Y, Z, R(pitch_xy), R(pitch_z), z0, nz, R(fdk_scale),
            _p(vol), _p(acc), nthreads)
    else:
        rc = getattr(lib(), "jo_bp" + _sfx(fp64))(
            _p(proj), _p(ptm), N, W, V, X, Y, Z, R(pitch_xy), R(pitch_z), z0, nz,
            _p(vol), _p(acc), _p(cnt), nthreads)
    assert rc == 0
    if return_parts:
        return vol, acc, cnt
    return vol


def fdk(raw, ptm, scan, recon, fp64=False, nthreads=8, z0=0, nz=None, dark=None, gain=None,
        return_stages=False, zfilter_v=False, ct=None, composite_views=False, fdk_scale=None):
    """Full oracle chain PreProcessingCore [-> CompositeCore] -> FilterCore
    [-> ZFilterAgent] -> BPCore [-> CTNumCore].

    scan: dict with W, V, bright_field, bh, hsr, dso, borders (optional).
    recon: dict with X, Y, Z, pxy, pz, cut_start, cut_end, window.
    composite_views: the offset-detector chain of JEngine.cpp:152-238 (each
    view composed with its conjugate into a 2W-wide frame, filtered with
    widthOut = W and yMirroring = detectorOnTheRight);
    zfilter_v: apply the 1-2-1 v smoothing between filter and backprojection
    (JEngine.cpp:240-245); ct: (norm0, norm1, mu_water) for the CT-number map
    (JEngine.cpp:297-305)."""
    p = preprocess(raw, borders=scan.get("borders", (0, 0, 0, 0)), bright_field=scan["bright_field"],
                   bh=scan.get("bh", (1.0, 0.0, 0.0)), dark=dark, gain=gain, fp64=fp64,
                   nthreads=nthreads)
    N, V, W = p.shape
    ptm = np.asarray(ptm).reshape(N, 12)
    det_right = detector_on_right(ptm[0], W) if composite_views else False
    fin = composite(p, ptm, scan["dso"], det_right, fp64=fp64, nthreads=nthreads) if composite_views else p
    Win = fin.shape[2]
    H = filter_design(Win, scan["hsr"], recon.get("cut_start", 0.0), recon.get("cut_end", 0.0),
                      recon.get("adj_pts", ()), recon.get("adj_db", ()), recon.get("window", 0),
                      fp64=fp64)
    src, blk = preweight_schedule(N)
    pws = np.stack([prefilter_weight(ptm[s], Win, V, scan["dso"], detector_on_right=det_right, fp64=fp64)
                    for s in src])
    f = filter_rows(fin, H, pws, blk, fp64=fp64, nthreads=nthreads, w_out=W, mirror=det_right)
    if zfilter_v:
        f = zfilter(f, fp64=fp64)
    vol = backproject(f, ptm, (recon["Z"], recon["Y"], recon["X"]), recon["pxy"], recon["pz"],
                      z0=z0, nz=nz, fp64=fp64, nthreads=nthreads, fdk_scale=fdk_scale)
    if ct is not None:
        vol = ctnum(vol, *ct, fp64=fp64)
    if return_stages:
        return vol, dict(pre=p, comp=fin if composite_views else None, det_right=det_right, H=H, pw=pws,
                         pw_index=blk, filt=f)
    return vol


def zfilter(frames, fp64=False):
    """ZFilterAgent: 1-2-1 smoothing along v with clamped edges, (N, V, W)."""
    x = np.ascontiguousarray(frames, _dt(fp64))
    if x.ndim == 2:
        x = x[None]
    N, V, W = x.shape
    out = np.empty_like(x)
    getattr(lib(), "jo_zfilter" + _sfx(fp64))(_p(x), N, W, V, _p(out))
    return out


def ctnum(vol, norm0, norm1, mu_water, fp64=False):
    """CTNumCore: attenuation -> CT number (returns a new array)."""
    x = np.array(vol, _dt(fp64), copy=True, order="C")
    R = _real(fp64)
    getattr(lib(), "jo_ctnum" + _sfx(fp64))(_p(x), ctypes.c_size_t(x.size), R(norm0), R(norm1), R(mu_water))
    return x


def bp_acc(proj, ptm, vol_shape, pitch_xy, pitch_z, weighted=True, fp64=True):
    """Un-normalised backprojection sum_v remap(proj_v [* uW_v]) (small sizes)."""
    proj = np.ascontiguousarray(proj, _dt(fp64))
    N, V, W = proj.shape
    ptm = np.ascontiguousarray(ptm, _dt(fp64)).reshape(N, 12)
    Z, Y, X = vol_shape
    acc = np.zeros((Z, Y, X), _dt(fp64))
    R = _real(fp64)
    getattr(lib(), "jo_bp_acc" + _sfx(fp64))(_p(proj), _p(ptm), N, W, V, X, Y, Z, R(pitch_xy), R(pitch_z),
                                              int(weighted), _p(acc))
    return acc


def forward_project(x, ptm, det_shape, pitch_xy, pitch_z, weighted=True, fp64=True):
    """Transpose of bp_acc: (N, V, W) frames from a (Z, Y, X) volume."""
    x = np.ascontiguousarray(x, _dt(fp64))
    Z, Y, X = x.shape
    ptm = np.ascontiguousarray(ptm, _dt(fp64)).reshape(-1, 12)
    N = ptm.shape[0]
    V, W = det_shape
    out = np.zeros((N, V, W), _dt(fp64))
    R = _real(fp64)
    getattr(lib(), "jo_fp" + _sfx(fp64))(_p(x), _p(ptm), N, W, V, X, Y, Z, R(pitch_xy), R(pitch_z),
                                         int(weighted), _p(out))
    return out
