"""Python host of the HIP back end: the three reconstruction cores over the
C-ABI, with torch tensors as device memory.

The class names follow the reference cores they replace:

* ``PreProcessingCore`` -- PreProcessingAgent/PreProcessingCore.cpp
* ``FilterCore``        -- FilterAgent/FilterCore.cpp (+ PreFilterWeightGenerator)
* ``BPCore``            -- BPAgent/BPCore.cpp and BPCUDAAgent/BPCUDACore.cpp
* ``CompositeCore``     -- CompositeAgent/CompositeCore.cpp (offset detectors)
* ``FDK``               -- the chain PreProcessingCore [-> CompositeCore] ->
                           FilterCore -> BPCore

All work is enqueued on torch's current stream.  Every method raises on a
failed call; none falls back to a CPU path.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import call, DTYPES, JH_FILTER_FP32, JH_FILTER_MFMA_BF16X3


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _dev_index(device):
    d = torch.device(device)
    return d.index if d.index is not None else torch.cuda.current_device()


def _require_cuda(t, name, dtype=None):
    if t is None:
        return
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if dtype is not None and t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")


class PreProcessingCore:
    """u16 raw frames -> line integrals (PreProcessingCore.cpp:15-86)."""

    def __init__(self, in_w, in_h, borders=(0, 0, 0, 0), out_w=None, out_h=None, bright_field=4000.0,
                 bh=(1.0, 0.0, 0.0), device="cuda:0"):
        bu, bd, bl, br = borders
        self.borders = tuple(borders)
        self.out_w = out_w if out_w is not None else in_w - bl - br
        self.out_h = out_h if out_h is not None else in_h - bu - bd
        self.in_w, self.in_h = in_w, in_h
        self.device = torch.device(device)
        h = ctypes.c_void_p()
        bhv = (ctypes.c_float * 3)(*bh)
        call("jh_pre_create", ctypes.byref(h), _dev_index(device), in_w, in_h, bu, bd, bl, br, self.out_w,
             self.out_h, ctypes.c_float(bright_field), bhv)
        self._h = h

    def process(self, raw, dark=None, gain=None, out=None):
        """raw: (N, in_h, in_w) int16/uint16 device tensor holding u16 counts."""
        _require_cuda(raw, "raw")
        if raw.element_size() != 2:
            raise ValueError("raw must hold 16-bit counts")
        n = raw.shape[0]
        if out is None:
            out = torch.empty((n, self.out_h, self.out_w), dtype=torch.float32, device=raw.device)
        _require_cuda(out, "out", torch.float32)
        _require_cuda(dark, "dark")
        _require_cuda(gain, "gain", torch.float32)
        call("jh_pre_run", self._h, _ptr(raw), _ptr(dark), _ptr(gain), _ptr(out), n, _stream())
        return out

    def output_bound(self, gain_min=1.0, gain_max=1.0):
        """Upper bound of |preprocessed value| for gains in [gain_min, gain_max]."""
        b = ctypes.c_double()
        call("jh_pre_output_bound", self._h, ctypes.c_double(gain_min), ctypes.c_double(gain_max), ctypes.byref(b))
        return b.value

    def resizes(self):
        bu, bd, bl, br = self.borders
        return (self.in_w - bl - br, self.in_h - bu - bd) != (self.out_w, self.out_h)

    def process_operand(self, raw, filt, dark=None, gain=None, view0=0, out=None):
        """Preprocessing written straight into `filt`'s pre-weighted MFMA
        operand (jh_pre_run_operand): feed it to FilterCore.process_operand_staged."""
        _require_cuda(raw, "raw")
        if raw.element_size() != 2:
            raise ValueError("raw must hold 16-bit counts")
        n = raw.shape[0]
        if out is None:
            out = torch.empty(filt.operand_bytes(n) // 4, dtype=torch.float32, device=raw.device)
        _require_cuda(dark, "dark")
        _require_cuda(gain, "gain", torch.float32)
        call("jh_pre_run_operand", self._h, filt._h, _ptr(raw), _ptr(dark), _ptr(gain), view0, n, _ptr(out),
             _stream())
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            try:
                _lib.lib().jh_pre_destroy(h)
            except Exception:  # interpreter shutdown: the library may be gone already
                pass


class FilterCore:
    """Row ramp filter with pre-filter weight (FilterCore.cpp:13-222).

    W is the input width (FilterCore's ``width``); w_out and y_mirroring are
    its ``widthOut`` and ``yMirroring`` (defaults: W, False -- the plain
    chain; the offset-detector chain uses 2 NumUsedDetsU in, NumUsedDetsU
    out, yMirroring = detectorOnTheRight, JEngine.cpp:223-238)."""

    def __init__(self, W, V, ptm, dso, half_sample_rate, cut_start=0.0, cut_end=0.0, adj_pts=(),
                 adj_db=(), window=0, mode="fp32", device="cuda:0", w_out=None, y_mirroring=False):
        self.W, self.V = W, V
        self.w_out = W if w_out is None else int(w_out)
        self.y_mirroring = bool(y_mirroring)
        ptm = np.ascontiguousarray(ptm, np.float64).reshape(-1, 12)
        self.n_views = ptm.shape[0]
        self.mode = mode
        m = {"fp32": _lib.JH_FILTER_FP32, "mfma": _lib.JH_FILTER_MFMA_BF16X3, "mfma6": _lib.JH_FILTER_MFMA_BF16X6,
             "fp32valu": _lib.JH_FILTER_FP32_VALU, "mfma16": _lib.JH_FILTER_MFMA_F16X3}[mode]
        n = len(adj_pts)
        pts = (ctypes.c_float * max(n, 1))(*adj_pts)
        db = (ctypes.c_float * max(n, 1))(*adj_db)
        h = ctypes.c_void_p()
        call("jh_filter_create_ex", ctypes.byref(h), _dev_index(device), W, V, self.w_out, int(self.y_mirroring),
             self.n_views, ptm.ctypes.data_as(ctypes.c_void_p), ctypes.c_float(dso),
             ctypes.c_float(half_sample_rate), ctypes.c_float(cut_start), ctypes.c_float(cut_end),
             pts if n else None, db if n else None, n, window, m)
        self._h = h

    def set_generated(self, enable=True):
        """Split-FP16 form: filter-matrix operand from its diagonals in LDS
        (True) or copied from the packed matrix (False); same result."""
        call("jh_filter_set_generated", self._h, int(bool(enable)))

    def generated(self):
        v = ctypes.c_int()
        call("jh_filter_get_generated", self._h, ctypes.byref(v))
        return bool(v.value)

    def output_bound(self, in_bound):
        """Upper bound of |filtered value| for inputs bounded by in_bound."""
        b = ctypes.c_double()
        call("jh_filter_output_bound", self._h, ctypes.c_double(in_bound), ctypes.byref(b))
        return b.value

    def response(self):
        L = ctypes.c_int()
        H = np.zeros(4 * self.W + 8, np.float64)
        call("jh_filter_response", self._h, H.ctypes.data_as(ctypes.c_void_p), ctypes.byref(L))
        return H[: L.value // 2 + 1].copy(), L.value

    def preweight(self, view):
        out = torch.empty((self.V, self.W), dtype=torch.float32, device="cuda")
        call("jh_filter_preweight", self._h, view, _ptr(out), _stream())
        return out

    def process(self, frames, view0=0, out=None):
        _require_cuda(frames, "frames", torch.float32)
        if frames.dim() != 3 or frames.shape[1:] != (self.V, self.W):
            raise ValueError(f"frames must be (N, {self.V}, {self.W})")
        n = frames.shape[0]
        if out is None:
            out = torch.empty((n, self.V, self.w_out), dtype=torch.float32, device=frames.device)
        _require_cuda(out, "out", torch.float32)
        call("jh_filter_run", self._h, _ptr(frames), _ptr(out), view0, n, _stream())
        return out

    def stage_fusable(self, bp, dtype):
        """Whether process_staged can feed `bp` in `dtype` (see jh_filter_stage_fusable)."""
        ok = ctypes.c_int()
        call("jh_filter_stage_fusable", self._h, bp._h, DTYPES[dtype], ctypes.byref(ok))
        return bool(ok.value)

    def operand_bytes(self, n_views):
        b = ctypes.c_size_t()
        call("jh_filter_operand_bytes", self._h, n_views, ctypes.byref(b))
        return b.value

    def process_operand_staged(self, operand, n_views, bp, dtype, view0=0, out=None):
        """process_staged() from the pre-weighted operand of
        PreProcessingCore.process_operand."""
        _require_cuda(operand, "operand", torch.float32)
        if operand.numel() * 4 < self.operand_bytes(n_views):
            raise ValueError("operand is smaller than operand_bytes(n_views)")
        if out is None:
            out = torch.empty(bp.stage_bytes(dtype, n_views), dtype=torch.uint8, device=operand.device)
        call("jh_filter_run_operand_staged", self._h, bp._h, _ptr(operand), view0, n_views, DTYPES[dtype], _ptr(out),
             _stream())
        return out

    def process_staged(self, frames, bp, dtype, view0=0, out=None):
        """process() followed by bp.stage() in one pass (bit-identical); the
        filtered FP32 frames are not written."""
        _require_cuda(frames, "frames", torch.float32)
        if frames.dim() != 3 or frames.shape[1:] != (self.V, self.W):
            raise ValueError(f"frames must be (N, {self.V}, {self.W})")
        n = frames.shape[0]
        if out is None:
            out = torch.empty(bp.stage_bytes(dtype, n), dtype=torch.uint8, device=frames.device)
        call("jh_filter_run_staged", self._h, bp._h, _ptr(frames), view0, n, DTYPES[dtype], _ptr(out), _stream())
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            try:
                _lib.lib().jh_filter_destroy(h)
            except Exception:  # interpreter shutdown: the library may be gone already
                pass


def detector_on_right(ptm0, n_used_u):
    """JEngine.cpp:153-161 (through the library's jh_detector_on_right)."""
    P = np.ascontiguousarray(ptm0, np.float64).reshape(12)
    return bool(_lib.lib().jh_detector_on_right(P.ctypes.data_as(ctypes.c_void_p), int(n_used_u)))


class CompositeCore:
    """Offset-detector composite (CompositeAgent/CompositeCore.cpp:13-150):
    each view (W x V) is widened to 2W columns with its conjugate view
    (v + N/2) % N re-sampled through the iso-centre plane."""

    def __init__(self, W, V, ptm, dso, detector_on_right=None, device="cuda:0"):
        ptm = np.ascontiguousarray(ptm, np.float64).reshape(-1, 12)
        self.W, self.V, self.n_views = W, V, ptm.shape[0]
        side = -1 if detector_on_right is None else int(bool(detector_on_right))
        h = ctypes.c_void_p()
        call("jh_composite_create", ctypes.byref(h), _dev_index(device), W, V, self.n_views,
             ptm.ctypes.data_as(ctypes.c_void_p), ctypes.c_float(dso), side)
        self._h = h
        s = ctypes.c_int()
        call("jh_composite_side", h, ctypes.byref(s))
        self.detector_on_right = bool(s.value)

    def process_pairs(self, dirc, conj, view0=0, out=None):
        """dirc, conj: (n, V, W) frames of views view0.. and their conjugates.
        Each frame must be contiguous; the frames may be strided alike (e.g.
        the two halves of a (n, 2, V, W) pipe of interleaved pairs)."""
        for t, name in ((dirc, "dirc"), (conj, "conj")):
            if t.dim() != 3 or t.shape[1:] != (self.V, self.W):
                raise ValueError(f"{name} must be (n, {self.V}, {self.W})")
            if not t.is_cuda or t.dtype != torch.float32:
                raise ValueError(f"{name} must be a float32 device tensor")
            if t.stride(2) != 1 or t.stride(1) != self.W or (t.shape[0] > 1 and t.stride(0) < self.V * self.W):
                raise ValueError(f"{name}: every frame must be contiguous")
        n = dirc.shape[0]
        if conj.shape[0] != n:
            raise ValueError("dirc and conj hold different frame counts")
        stride = dirc.stride(0) if n > 1 else self.V * self.W
        if n > 1 and conj.stride(0) != stride:
            raise ValueError("dirc and conj frames must have the same stride")
        if out is None:
            out = torch.empty((n, self.V, 2 * self.W), dtype=torch.float32, device=dirc.device)
        _require_cuda(out, "out", torch.float32)
        call("jh_composite_run", self._h, _ptr(dirc), _ptr(conj), ctypes.c_longlong(stride), _ptr(out),
             view0, n, _stream())
        return out

    def process(self, frames, out=None):
        """frames: all N views (N, V, W) resident on the device -> (N, V, 2W)."""
        _require_cuda(frames, "frames", torch.float32)
        N = self.n_views
        if frames.dim() != 3 or frames.shape != (N, self.V, self.W):
            raise ValueError(f"frames must be ({N}, {self.V}, {self.W})")
        if out is None:
            out = torch.empty((N, self.V, 2 * self.W), dtype=torch.float32, device=frames.device)
        half = N // 2
        # views [0, N/2) pair with [N/2, N), views [N/2, N) with [0, N/2)
        self.process_pairs(frames[:half], frames[half:], 0, out[:half])
        self.process_pairs(frames[half:], frames[:half], half, out[half:])
        return out

    def process_pairs_operand(self, dirc, conj, filt, view0=0, out=None, aview0=0):
        """process_pairs() written straight into `filt`'s pre-weighted MFMA
        operand (jh_composite_run_operand): views view0.. land at the rows of
        an operand whose filter launch starts at view aview0."""
        for t, name in ((dirc, "dirc"), (conj, "conj")):
            if t.dim() != 3 or t.shape[1:] != (self.V, self.W):
                raise ValueError(f"{name} must be (n, {self.V}, {self.W})")
            if not t.is_cuda or t.dtype != torch.float32:
                raise ValueError(f"{name} must be a float32 device tensor")
            if t.stride(2) != 1 or t.stride(1) != self.W or (t.shape[0] > 1 and t.stride(0) < self.V * self.W):
                raise ValueError(f"{name}: every frame must be contiguous")
        n = dirc.shape[0]
        if conj.shape[0] != n:
            raise ValueError("dirc and conj hold different frame counts")
        stride = dirc.stride(0) if n > 1 else self.V * self.W
        if n > 1 and conj.stride(0) != stride:
            raise ValueError("dirc and conj frames must have the same stride")
        if out is None:
            out = torch.empty(filt.operand_bytes(view0 + n - aview0) // 4, dtype=torch.float32, device=dirc.device)
        _require_cuda(out, "out", torch.float32)
        if out.numel() * 4 < filt.operand_bytes(view0 + n - aview0):
            raise ValueError("out is smaller than the operand of these views")
        call("jh_composite_run_operand", self._h, filt._h, _ptr(dirc), _ptr(conj), ctypes.c_longlong(stride),
             view0, n, _ptr(out), aview0, _stream())
        return out

    def process_operand(self, frames, filt, out=None):
        """process() written straight into `filt`'s pre-weighted operand of
        all N views: feed it to FilterCore.process_operand_staged(out, N, ...)."""
        _require_cuda(frames, "frames", torch.float32)
        N = self.n_views
        if frames.dim() != 3 or frames.shape != (N, self.V, self.W):
            raise ValueError(f"frames must be ({N}, {self.V}, {self.W})")
        if out is None:
            out = torch.empty(filt.operand_bytes(N) // 4, dtype=torch.float32, device=frames.device)
        half = N // 2
        self.process_pairs_operand(frames[:half], frames[half:], filt, 0, out)
        self.process_pairs_operand(frames[half:], frames[:half], filt, half, out)
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            try:
                _lib.lib().jh_composite_destroy(h)
            except Exception:  # interpreter shutdown: the library may be gone already
                pass


class BPCore:
    """Normalised voxel-driven backprojection (BPCore.cpp:15-205)."""

    def __init__(self, ptm, W, V, X, Y, Z, pxy, pz, device="cuda:0"):
        ptm = np.ascontiguousarray(ptm, np.float64).reshape(-1, 12)
        self.n_views = ptm.shape[0]
        self.W, self.V, self.X, self.Y, self.Z = W, V, X, Y, Z
        h = ctypes.c_void_p()
        call("jh_bp_create", ctypes.byref(h), _dev_index(device), ptm.ctypes.data_as(ctypes.c_void_p),
             self.n_views, W, V, X, Y, Z, ctypes.c_float(pxy), ctypes.c_float(pz))
        self._h = h

    def uweights(self):
        out = np.zeros((self.n_views, self.W), np.float32)
        call("jh_bp_uweights", self._h, out.ctypes.data_as(ctypes.c_void_p))
        return out

    def geometry_mode(self):
        m = ctypes.c_int()
        call("jh_bp_geometry_mode", self._h, ctypes.byref(m))
        return m.value

    def set_tile(self, zc=0, ty=0):
        call("jh_bp_set_tile", self._h, zc, ty)

    FETCH = {"gather": 0, "lds": 1}

    def autotune(self, q, dtype="bf16", n_views=None, reps=1):
        """Times every fetch form and tile shape on views 0..n_views-1 of q and
        keeps the fastest; returns {(fetch, zc, ty): ms} (forms that do not
        apply to this geometry are left out)."""
        n = self.n_views if n_views is None else n_views
        vol = torch.empty((self.Z, self.Y, self.X), dtype=torch.float32, device=q.device)
        ms = (ctypes.c_float * 18)()
        call("jh_bp_autotune", self._h, _ptr(q), DTYPES[dtype], n, _ptr(vol), reps, _stream(), ms)
        keys = [(f, z, t) for f in ("gather", "lds") for z in (8, 16, 32) for t in (1, 2, 4)]
        return {k: v for k, v in zip(keys, list(ms)) if v >= 0}

    def set_zfilter(self, enable=True):
        """Fuse ZFilterAgent's v smoothing into stage()."""
        call("jh_bp_set_zfilter", self._h, int(bool(enable)))

    def set_ctnum(self, enable=True, norm0=1.0, norm1=1.0, mu_water=0.02):
        """Fuse CTNumAgent's map into normalised outputs (run, finalize)."""
        call("jh_bp_set_ctnum", self._h, int(bool(enable)), ctypes.c_float(norm0), ctypes.c_float(norm1),
             ctypes.c_float(mu_water))

    def set_fetch(self, fetch):
        call("jh_bp_set_fetch", self._h, self.FETCH[fetch])

    def set_f16_exponent(self, exponent):
        """FP16 storage holds q * 2^-exponent; outputs are scaled back."""
        call("jh_bp_set_f16_exponent", self._h, int(exponent))

    def f16_exponent(self):
        k = ctypes.c_int()
        call("jh_bp_get_f16_exponent", self._h, ctypes.byref(k))
        return k.value

    WEIGHT = {"count": 0, "fdk": 1}

    def set_weight(self, weight="count", scale=0.0):
        """"count": the reference's normalisation by the back-projected uW;
        "fdk": the distance weight, vol = scale * sum_v bilinear(q_v) / w_v^2
        (scale 0 selects 2 / n_views)."""
        call("jh_bp_set_weight", self._h, self.WEIGHT[weight], ctypes.c_float(scale))

    def weight(self):
        m, sc = ctypes.c_int(), ctypes.c_float()
        call("jh_bp_get_weight", self._h, ctypes.byref(m), ctypes.byref(sc))
        return {v: k for k, v in self.WEIGHT.items()}[m.value], sc.value

    LAYOUT = {"quad": 0, "strip": 1}

    def set_layout(self, layout):
        """Staged BF16/FP16 layout ("quad" or "strip"); stage again after a change."""
        call("jh_bp_set_layout", self._h, self.LAYOUT[layout])

    def layout(self):
        v = ctypes.c_int()
        call("jh_bp_get_layout", self._h, ctypes.byref(v))
        return {v2: k for k, v2 in self.LAYOUT.items()}[v.value]

    def fetch(self):
        f = ctypes.c_int()
        call("jh_bp_get_fetch", self._h, ctypes.byref(f))
        return {v: k for k, v in self.FETCH.items()}[f.value]

    def fetch_in_use(self, dtype):
        """The fetch form a count-normalised launch on `dtype` takes now:
        "lds" only where the LDS-staged form applies at the current tile."""
        f = ctypes.c_int()
        call("jh_bp_fetch_in_use", self._h, DTYPES[dtype], _stream(), ctypes.byref(f))
        return {v: k for k, v in self.FETCH.items()}[f.value]

    def tile(self):
        zc, ty = ctypes.c_int(), ctypes.c_int()
        call("jh_bp_get_tile", self._h, ctypes.byref(zc), ctypes.byref(ty))
        return zc.value, ty.value

    def stage_bytes(self, dtype, n_views):
        b = ctypes.c_size_t()
        call("jh_bp_stage_bytes", self._h, DTYPES[dtype], n_views, ctypes.byref(b))
        return b.value

    def stage(self, filtered, dtype="f32", view0=0, out=None):
        """q = filtered * uW in the backprojector's layout (a byte tensor)."""
        _require_cuda(filtered, "filtered", torch.float32)
        n = filtered.shape[0]
        if out is None:
            out = torch.empty(self.stage_bytes(dtype, n), dtype=torch.uint8, device=filtered.device)
        call("jh_bp_stage", self._h, _ptr(filtered), view0, n, DTYPES[dtype], _ptr(out), _stream())
        return out

    def pack(self, frames, dtype="f32", view0=0, out=None):
        """frames in the backprojector's layout without uW weighting."""
        _require_cuda(frames, "frames", torch.float32)
        n = frames.shape[0]
        if out is None:
            out = torch.empty(self.stage_bytes(dtype, n), dtype=torch.uint8, device=frames.device)
        call("jh_bp_pack", self._h, _ptr(frames), view0, n, DTYPES[dtype], _ptr(out), _stream())
        return out

    def backproject(self, q, view0, n_views, out=None):
        """Un-normalised backprojection B of an FP32 packed/staged buffer."""
        _require_cuda(q, "q")
        if out is None:
            out = torch.empty((self.Z, self.Y, self.X), dtype=torch.float32, device=q.device)
        call("jh_bp_backproject", self._h, _ptr(q), view0, n_views, _ptr(out), _stream())
        return out

    def project(self, vol, view0=0, n_views=None, weighted=False, out=None):
        """Forward projection B^T vol (frames n_views x V x W)."""
        _require_cuda(vol, "vol", torch.float32)
        n = self.n_views - view0 if n_views is None else n_views
        if out is None:
            out = torch.empty((n, self.V, self.W), dtype=torch.float32, device=vol.device)
        _require_cuda(out, "out", torch.float32)
        call("jh_fp_project", self._h, _ptr(vol), _ptr(out), view0, n, int(weighted), _stream())
        return out

    def run(self, q, dtype="f32", out=None):
        _require_cuda(q, "q")
        if out is None:
            out = torch.empty((self.Z, self.Y, self.X), dtype=torch.float32, device=q.device)
        _require_cuda(out, "out", torch.float32)
        call("jh_bp_run", self._h, _ptr(q), DTYPES[dtype], _ptr(out), _stream())
        return out

    def reset(self):
        call("jh_bp_reset", self._h, _stream())

    def accumulate(self, q, dtype, view0, n_views, z0=None, nz=None):
        """Streaming accumulation of views view0.. into the whole volume, or
        (z0, nz given) into slices z0..z0+nz-1 only (jh_bp_accumulate_slab)."""
        _require_cuda(q, "q")
        if z0 is None:
            call("jh_bp_accumulate", self._h, _ptr(q), DTYPES[dtype], view0, n_views, _stream())
        else:
            call("jh_bp_accumulate_slab", self._h, _ptr(q), DTYPES[dtype], view0, n_views, z0, nz, _stream())

    def finalize(self, z0=0, nz=None, out=None):
        nz = self.Z - z0 if nz is None else nz
        if out is None:
            out = torch.empty((nz, self.Y, self.X), dtype=torch.float32, device="cuda")
        call("jh_bp_finalize", self._h, _ptr(out), z0, nz, _stream())
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            try:
                _lib.lib().jh_bp_destroy(h)
            except Exception:  # interpreter shutdown: the library may be gone already
                pass


def f16_exponent_for(bound):
    """The FP16 storage exponent k that maps values bounded by `bound` into
    [0, 2^15]: FP16's largest finite value is 65504."""
    if not bound > 0.0:
        return 0
    m, e = np.frexp(bound)  # bound = m 2^e, 0.5 <= m < 1: bound <= 2^e
    return int(min(64, max(-64, e - 15)))


def zfilter(frames, out=None):
    """ZFilterAgent (FilterAgent/ZFilterAgent.cpp:49-71): 1-2-1 smoothing
    along v with clamped edges, frames (N, V, W) f32 on the device."""
    _require_cuda(frames, "frames", torch.float32)
    if frames.dim() != 3:
        raise ValueError("frames must be (N, V, W)")
    n, V, W = frames.shape
    if out is None:
        out = torch.empty_like(frames)
    _require_cuda(out, "out", torch.float32)
    with torch.cuda.device(frames.device):
        call("jh_zfilter_run", _ptr(frames), _ptr(out), W, V, n, _stream())
    return out


def ctnum(x, norm0=1.0, norm1=1.0, mu_water=0.02):
    """CTNumCore (CTNumAgent/CTNumCore.cpp:22-36), in place on a device tensor."""
    _require_cuda(x, "x", torch.float32)
    with torch.cuda.device(x.device):
        call("jh_ctnum_run", _ptr(x), ctypes.c_size_t(x.numel()), ctypes.c_float(norm0), ctypes.c_float(norm1),
             ctypes.c_float(mu_water), _stream())
    return x


class FDK:
    """PreProcessingCore [-> CompositeCore] -> FilterCore [-> ZFilterAgent]
    -> BPCore [-> CTNumAgent] on one device.  zfilter and ct (norm0, norm1,
    mu_water) are fused into the backprojector's staging and normalisation.
    composite=True is JEngine's offset-detector chain (JEngine.cpp:152-238):
    every view composed with its conjugate into 2W columns, filtered back to
    W with yMirroring = detectorOnTheRight."""

    def __init__(self, scan, recon, ptm, filter_mode="fp32", storage="f32", device="cuda:0", zfilter=False,
                 ct=None, composite=False, weight="count", weight_scale=0.0):
        self.scan, self.recon, self.storage = scan, recon, storage
        W, V = scan.NumUsedDetsU, scan.NumUsedDetsV
        self.pre = PreProcessingCore(scan.NumDetsU, scan.NumDetsV,
                                     (scan.BorderSizeUp, scan.BorderSizeDown, scan.BorderSizeLeft,
                                      scan.BorderSizeRight), W, V, scan.BrightField,
                                     scan.BeamHardeningParams, device)
        self.composite = None
        win, mirror = W, False
        if composite:
            self.composite = CompositeCore(W, V, ptm, scan.DSO, None, device)
            win, mirror = 2 * W, self.composite.detector_on_right
        self.filter = FilterCore(win, V, ptm, scan.DSO, scan.HalfSampleRate, recon.FilterCutOffStart,
                                 recon.FilterCutOffEnd, recon.FilterAdjustPoints, recon.FilterAdjustLevelInDB,
                                 recon.FilterWindow, filter_mode, device, w_out=W, y_mirroring=mirror)
        self.bp = BPCore(ptm, W, V, recon.NumPixelsX, recon.NumPixelsY, recon.NumPixelsZ, recon.PitchXY,
                         recon.PitchZ, device)
        if zfilter:
            self.bp.set_zfilter(True)
        if ct is not None:
            self.bp.set_ctnum(True, *ct)
        if weight != "count":
            self.bp.set_weight(weight, weight_scale)
        self._f16_gain = ()  # the gain tensor the FP16 exponent was chosen for

    def prepare(self, gain=None):
        """FP16 storage: chooses the storage exponent from an a-priori bound
        of |q| (preprocessing bound for this gain map's range, times the
        filter's gain bound; uW <= 1), so no staged value can overflow.  Reads
        the gain map's range once per gain tensor."""
        if self.storage != "f16" or (self._f16_gain and self._f16_gain[0] is gain):
            return
        if gain is None:
            gmin = gmax = 1.0
        else:
            lo, hi = torch.aminmax(gain)
            gmin, gmax = float(lo), float(hi)
        bound = self.filter.output_bound(self.pre.output_bound(gmin, gmax))
        self.bp.set_f16_exponent(f16_exponent_for(bound))
        self._f16_gain = (gain,)

    def fused(self):
        """Whether the filter writes the staged strips directly (filter and
        staging in one kernel, FilterCore.process_staged)."""
        return self.filter.stage_fusable(self.bp, self.storage)

    def operand_path(self):
        """Whether preprocessing writes the filter's pre-weighted operand
        (fused chain without composite or resizing): pre -> filter+stage -> bp."""
        return (self.composite is None and not self.pre.resizes() and self.filter.mode != "fp32valu"
                and self.fused())

    def composite_operand_path(self):
        """Whether the composite writes the filter's pre-weighted operand
        (fused offset-detector chain): pre -> composite -> filter+stage -> bp."""
        return self.composite is not None and self.filter.mode != "fp32valu" and self.fused()

    def __call__(self, raw, dark=None, gain=None, return_stages=False):
        self.prepare(gain)
        if not return_stages and self.operand_path() and raw.shape[0] * self.filter.V <= 1048560:
            a = self.pre.process_operand(raw, self.filter, dark, gain)
            q = self.filter.process_operand_staged(a, raw.shape[0], self.bp, self.storage)
            return self.bp.run(q, self.storage)
        p = self.pre.process(raw, dark, gain)
        if not return_stages and self.composite_operand_path() and raw.shape[0] * self.filter.V <= 1048560:
            a = self.composite.process_operand(p, self.filter)
            q = self.filter.process_operand_staged(a, raw.shape[0], self.bp, self.storage)
            return self.bp.run(q, self.storage)
        c = self.composite.process(p) if self.composite is not None else p
        if not return_stages and self.fused():
            q = self.filter.process_staged(c, self.bp, self.storage)
            return self.bp.run(q, self.storage)
        f = self.filter.process(c)
        q = self.bp.stage(f, self.storage)
        vol = self.bp.run(q, self.storage)
        if return_stages:
            return vol, dict(pre=p, comp=c if self.composite is not None else None, filt=f, q=q)
        return vol


class BatchGraph:
    """Many scans of uneven sizes reconstructed by one hipGraph launch
    (the ``jh_graph_*`` entry points).  Each scan is an ``FDK`` (its plans)
    with bound device buffers; the graph reads the buffers at every launch.
    """

    def __init__(self, device="cuda:0"):
        h = ctypes.c_void_p()
        call("jh_graph_create", ctypes.byref(h), _dev_index(device))
        self._h = h
        self._keep = []  # plans and buffers the graph refers to

    def add(self, fdk, raw, vol, dark=None, gain=None, views_per_chunk=0):
        """Adds one scan: raw (N, inH, inW) 16-bit counts, vol (Z, Y, X) f32."""
        _require_cuda(raw, "raw")
        if raw.element_size() != 2:
            raise ValueError("raw must hold 16-bit counts")
        _require_cuda(vol, "vol", torch.float32)
        _require_cuda(dark, "dark")
        _require_cuda(gain, "gain", torch.float32)
        fdk.prepare(gain)
        idx = ctypes.c_int()
        call("jh_graph_add_scan", self._h, fdk.pre._h, fdk.filter._h, fdk.bp._h, _ptr(raw), _ptr(dark),
             _ptr(gain), DTYPES[fdk.storage], views_per_chunk, _ptr(vol), ctypes.byref(idx))
        self._keep.append((fdk, raw, vol, dark, gain))
        return idx.value

    def autotune(self, reps=1):
        call("jh_graph_autotune", self._h, reps, _stream())

    def tuning(self, index):
        """(views per chunk, zc, ty) of scan `index`."""
        c, z, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        call("jh_graph_scan_tuning", self._h, index, ctypes.byref(c), ctypes.byref(z), ctypes.byref(t))
        return c.value, z.value, t.value

    def instantiate(self, branches=4):
        call("jh_graph_instantiate", self._h, branches)

    def launch(self):
        call("jh_graph_launch", self._h, _stream())

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            try:
                _lib.lib().jh_graph_destroy(h)
            except Exception:  # interpreter shutdown: the library may be gone already
                pass
