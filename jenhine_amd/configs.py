"""The reconstruction workloads used by the tests and the benchmark.

* ``cbct128``  -- plain FDK at 128^3, 128 views of 256^2 (Ram-Lak, no window)
* ``dental512`` -- 512^3 at 0.15 mm, 600 views of 768^2 at 0.2 mm, Shepp-Logan,
  cubic beam hardening, 40 ellipsoids plus a 0.3/mm tooth
* ``adjoint256`` -- 256^3 at 0.3 mm, 360 views of 512^2 at 0.4 mm
* ``batch16``   -- 16 scans of uneven sizes drawn from the seed
* ``offset128``, ``offset_dental512`` -- half-fan (offset) detectors, the
  composite chain

Raw counts are I = Poisson(I0 * g * exp(-p)) + d with I0 = 4000, per-pixel gain
g ~ U(0.9, 1.1) and dark d ~ U{90..110}.
"""
from dataclasses import dataclass, field, replace
from typing import List, Optional

import numpy as np

from . import geometry, synth
from .geometry import ReconParams, ScanParams


@dataclass
class Workload:
    name: str
    sod: float
    sdd: float
    n_views: int
    n_u: int
    n_v: int
    du: float
    X: int
    Y: int
    Z: int
    pxy: float
    pz: float
    n_ellipsoids: int = 10
    tooth: bool = False
    window: int = 0          # 0 reference cut-off semantics, 1 Shepp-Logan
    ramlak: bool = True      # reference window: cut-offs f_N / 2 f_N (no taper)
    bh: Optional[List[float]] = None  # None = seeded cubic when tooth, else {1,0,0}
    I0: float = 4000.0
    seed: int = 1
    u_shift: float = 0.0     # principal point left of the detector centre (offset detector), pixels

    @property
    def gups_units(self):
        return self.X * self.Y * self.Z * self.n_views

    def scan_params(self):
        return ScanParams(NumViews=self.n_views, DSO=self.sod, DSD=self.sdd, NumDetsU=self.n_u,
                          NumDetsV=self.n_v, DetectorPixelSize=self.du, BrightField=self.I0,
                          BeamHardeningParams=list(self.bh_params()))

    def recon_params(self):
        scan = self.scan_params()
        hsr = scan.HalfSampleRate
        if self.window == 0 and self.ramlak:
            cs, ce = hsr, 2.0 * hsr
        else:
            cs, ce = 0.0, 0.0
        return ReconParams(NumPixelsX=self.X, NumPixelsY=self.Y, NumPixelsZ=self.Z, PitchXY=self.pxy,
                           PitchZ=self.pz, FilterCutOffStart=cs, FilterCutOffEnd=ce,
                           FilterWindow=self.window)

    def bh_params(self):
        if self.bh is not None:
            return tuple(self.bh)
        if not self.tooth:
            return (1.0, 0.0, 0.0)
        u = synth.uniform(self.seed, (2,), 0.0, 1.0, stream=7)
        return (1.0, float(np.float32(0.02 + 0.02 * u[0])), float(np.float32(0.002 + 0.003 * u[1])))

    def ptms(self):
        return geometry.circular_ptms(self.n_views, self.sod, self.sdd, self.n_u, self.n_v, self.du,
                                      u_shift=self.u_shift)

    def ellipsoids(self):
        half = (self.X * self.pxy / 2.0, self.Y * self.pxy / 2.0, self.Z * self.pz / 2.0)
        return synth.phantom(self.seed, self.n_ellipsoids, half, tooth=self.tooth)

    def data(self, noise=True):
        """ptm (N x 12 f64), line integrals, raw u16, gain, dark."""
        ptm = self.ptms()
        ell = self.ellipsoids()
        p = synth.line_integrals(ptm, self.n_u, self.n_v, ell)
        gain, dark = synth.gain_dark(self.seed, self.n_u, self.n_v)
        raw = synth.raw_counts(self.seed, p, self.I0, gain, dark, noise=noise)
        return dict(ptm=ptm, lineint=p, raw=raw, gain=gain, dark=dark, ellipsoids=ell)

    def scaled(self, **kw):
        return replace(self, **kw)


CBCT128 = Workload("cbct128", 300.0, 500.0, 128, 256, 256, 0.5, 128, 128, 128, 0.25, 0.25,
                   n_ellipsoids=10, window=0, ramlak=True)
DENTAL512 = Workload("dental512", 350.0, 600.0, 600, 768, 768, 0.2, 512, 512, 512, 0.15, 0.15,
                     n_ellipsoids=40, tooth=True, window=1, ramlak=False)
ADJOINT256 = Workload("adjoint256", 300.0, 500.0, 360, 512, 512, 0.4, 256, 256, 256, 0.3, 0.3,
                      n_ellipsoids=10, window=0, ramlak=True)

# Offset (half-fan) detector: the central ray hits column 31.5 of 160, so the
# detector extends to the right of it and the reconstruction needs the
# composite of every view with its conjugate (JEngine.cpp:152-238).
OFFSET128 = Workload("offset128", 300.0, 500.0, 128, 160, 128, 0.5, 128, 128, 64, 0.25, 0.25,
                     n_ellipsoids=10, window=0, ramlak=True, u_shift=48.0)
# Dental-scale half-fan: 768 x 768 detector, central ray at column 95.5.
OFFSET_DENTAL512 = Workload("offset_dental512", 350.0, 600.0, 600, 768, 768, 0.2, 512, 512, 512, 0.15, 0.15,
                            n_ellipsoids=40, tooth=True, window=1, ramlak=False, u_shift=288.0)

WORKLOADS = {w.name: w for w in (CBCT128, DENTAL512, ADJOINT256, OFFSET128, OFFSET_DENTAL512)}


def batch_scan(name, n, views, det, seed):
    """One scan of the batched form: n^3 at 0.25 mm, SOD 300, SDD 500,
    Ram-Lak, 10 ellipsoids; the detector pitch is chosen so the volume fits
    the field of view."""
    fov = n * 0.25 * 1.25
    du = fov * (500.0 / 300.0) / det
    return Workload(name, 300.0, 500.0, views, det, det, float(du), n, n, n, 0.25, 0.25,
                    n_ellipsoids=10, window=0, ramlak=True, seed=seed)


def batch16(seed=4):
    """16 scans of uneven sizes: volumes from {128..400}^3, views from
    {180, 360, 540, 720}, detectors from {256..1024}^2."""
    rng = np.random.default_rng(seed)
    vols = [128, 192, 256, 320, 400]
    views = [180, 360, 540, 720]
    dets = [256, 512, 768, 1024]
    out = []
    for s in range(16):
        n = int(rng.choice(vols))
        d = int(rng.choice(dets))
        v = int(rng.choice(views))
        out.append(batch_scan(f"batch{s}", n, v, d, seed * 100 + s))
    return out
