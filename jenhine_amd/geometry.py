"""Scan and reconstruction parameters and circular-orbit projection matrices.

``ScanParams`` and ``ReconParams`` keep the field names of the reference
structs (Config/ScanParams.h:13-40, Config/ReconParams.h:14-68) for the fields
the reconstruction path reads, with the derived fields computed as
Config/ConfigHelper.cpp:52-63 does.

Projection matrices follow the reference convention (row-major 3x4,
TransformMatrix/TransformMatrixBase.h:33-42): (u, v, w) = P (x, y, z, 1) in mm,
detector coordinates in pixels, and the backprojector samples pixel
(u/w + 0.5, v/w + 0.5) (BPAgent/BPCore.cpp:207-220).  They are built in FP64.
"""
from dataclasses import dataclass, field
from typing import List

import numpy as np


@dataclass
class ScanParams:
    NumViews: int
    DSO: float
    DSD: float
    NumDetsU: int
    NumDetsV: int
    DetectorPixelSize: float
    BorderSizeUp: int = 0
    BorderSizeDown: int = 0
    BorderSizeLeft: int = 0
    BorderSizeRight: int = 0
    BrightField: float = 4000.0
    BeamHardeningParams: List[float] = field(default_factory=lambda: [1.0, 0.0, 0.0])

    @property
    def HalfSampleRate(self):
        # ConfigHelper.cpp:52-53 (float arithmetic)
        det_iso = np.float32(self.DetectorPixelSize) / np.float32(self.DSD) * np.float32(self.DSO)
        return float(np.float32(0.5) * (np.float32(1.0) / det_iso))

    @property
    def NumUsedDetsU(self):
        return self.NumDetsU - self.BorderSizeLeft - self.BorderSizeRight

    @property
    def NumUsedDetsV(self):
        return self.NumDetsV - self.BorderSizeUp - self.BorderSizeDown


@dataclass
class ReconParams:
    NumPixelsX: int
    NumPixelsY: int
    NumPixelsZ: int
    PitchXY: float
    PitchZ: float
    FilterCutOffStart: float = 0.0
    FilterCutOffEnd: float = 0.0
    FilterAdjustPoints: List[float] = field(default_factory=list)
    FilterAdjustLevelInDB: List[float] = field(default_factory=list)
    FilterWindow: int = 0  # 0 = reference cut-off semantics, 1 = Shepp-Logan (extension)


def circular_ptms(n_views, sod, sdd, n_u, n_v, du, dv=None, start_deg=0.0, u_shift=0.0):
    """FP64 projection matrices of a circular orbit about z.

    Source at angle b_k = start + 360 k / n_views on a circle of radius SOD in
    the xy-plane, flat detector at distance SDD from the source, u along the
    in-plane tangent and v along +z.  w = 1 at the iso-centre (P[11] = 1) and
    the central ray samples the detector centre ((n-1)/2 in both axes), or
    u_shift pixels left of it (an offset, half-fan detector that extends to
    the right of the central ray when u_shift > 0)."""
    dv = du if dv is None else dv
    ku = sdd / (sod * du)
    kv = sdd / (sod * dv)
    cu = (n_u - 1) / 2.0 - 0.5 - u_shift
    cv = (n_v - 1) / 2.0 - 0.5
    P = np.zeros((n_views, 3, 4))
    for k in range(n_views):
        b = np.deg2rad(start_deg) + 2.0 * np.pi * k / n_views
        c, s = np.cos(b), np.sin(b)
        w = np.array([-c / sod, -s / sod, 0.0, 1.0])
        P[k, 2] = w
        P[k, 0] = np.array([-s * ku, c * ku, 0.0, 0.0]) + cu * w
        P[k, 1] = np.array([0.0, 0.0, kv, 0.0]) + cv * w
    return P.reshape(n_views, 12)


# ProjectionMatrixModifier restated in FP64 on the host
# (TransformMatrix/ProjectionMatrixModifier.cpp:5-43); the device form is
# jh_ptm_modify.
def move_u(ptm, u):
    P = np.array(ptm, np.float64).reshape(-1, 3, 4)
    P[:, 0, :] -= P[:, 2, :] * u
    return P.reshape(np.shape(ptm))


def move_v(ptm, v):
    P = np.array(ptm, np.float64).reshape(-1, 3, 4)
    P[:, 1, :] -= P[:, 2, :] * v
    return P.reshape(np.shape(ptm))


def reverse_x(ptm):
    P = np.array(ptm, np.float64).reshape(-1, 3, 4)
    P[:, :, 0] *= -1.0
    return P.reshape(np.shape(ptm))


def reverse_y(ptm):
    P = np.array(ptm, np.float64).reshape(-1, 3, 4)
    P[:, :, 1] *= -1.0
    return P.reshape(np.shape(ptm))


def move_z(ptm, z):
    P = np.array(ptm, np.float64).reshape(-1, 3, 4)
    P[:, :, 3] += P[:, :, 2] * z
    return P.reshape(np.shape(ptm))


def adjust_ptm(ptm, border_left=0, border_up=0, mirroring_y=False, center_z=0.0):
    """JEngine's adjustment of the matrices read from PTM.dat
    (JEngine/JEngine.cpp:139-150), in FP64: MoveUAxis(BorderSizeLeft),
    MoveVAxis(BorderSizeUp), ReverseXAxis, ReverseYAxis when MirroringY, then
    MoveZAxis(-CenterZ) (CenterZ = PitchZ * Volume_Translation_Z,
    Config/ConfigHelper.cpp:85-86)."""
    P = move_v(move_u(ptm, float(np.float32(border_left))), float(np.float32(border_up)))
    P = reverse_x(P)
    if mirroring_y:
        P = reverse_y(P)
    return move_z(P, -float(center_z))


def unadjust_ptm(ptm, border_left=0, border_up=0, mirroring_y=False, center_z=0.0):
    """The inverse of adjust_ptm (each step undone in reverse order): the
    scanner matrices whose adjustment gives ptm."""
    P = move_z(ptm, float(center_z))
    if mirroring_y:
        P = reverse_y(P)
    P = reverse_x(P)
    return move_u(move_v(P, -float(np.float32(border_up))), -float(np.float32(border_left)))


def project(ptm, xyz):
    """(u, v) of world points xyz (n x 3) through one 3x4 matrix."""
    P = np.asarray(ptm, np.float64).reshape(3, 4)
    X = np.c_[np.asarray(xyz, np.float64), np.ones(len(xyz))]
    uvw = X @ P.T
    return uvw[:, 0] / uvw[:, 2], uvw[:, 1] / uvw[:, 2]
