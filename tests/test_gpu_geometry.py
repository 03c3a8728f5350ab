"""ProjectionMatrixModifier on the device (jh_ptm_modify, FP64) against the
FP64 host restatement in jenhine_amd.geometry: every operation alone and
JEngine's adjustment sequence (JEngine/JEngine.cpp:139-150), to the bit."""
import ctypes

import numpy as np
import pytest
import torch

from jenhine_amd import _lib, geometry
from jenhine_amd._lib import call

pytestmark = pytest.mark.gpu

HOST = {
    _lib.JH_PTM_MOVE_U: geometry.move_u,
    _lib.JH_PTM_MOVE_V: geometry.move_v,
    _lib.JH_PTM_REVERSE_X: lambda P, _v: geometry.reverse_x(P),
    _lib.JH_PTM_REVERSE_Y: lambda P, _v: geometry.reverse_y(P),
    _lib.JH_PTM_MOVE_Z: geometry.move_z,
}


def _matrices(n=37, seed=3):
    # a circular orbit with an off-centre principal point, plus a seeded
    # perturbation so that no entry is zero
    P = geometry.circular_ptms(n, 350.0, 600.0, 768, 640, 0.2, u_shift=17.25)
    rng = np.random.default_rng(seed)
    return P * (1.0 + 1e-3 * rng.standard_normal(P.shape))


def _modify(d, n, op, value):
    call("jh_ptm_modify", ctypes.c_void_p(d.data_ptr()), n, op, ctypes.c_double(value),
         ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))


@pytest.mark.parametrize("op,value", [(_lib.JH_PTM_MOVE_U, 11.0), (_lib.JH_PTM_MOVE_U, -3.75),
                                      (_lib.JH_PTM_MOVE_V, 7.0), (_lib.JH_PTM_REVERSE_X, 0.0),
                                      (_lib.JH_PTM_REVERSE_Y, 0.0), (_lib.JH_PTM_MOVE_Z, -4.35),
                                      (_lib.JH_PTM_MOVE_Z, 0.11)])
def test_ptm_modify_each_op_matches_fp64_host(gpu, op, value):
    P = _matrices()
    d = torch.from_numpy(P.copy()).to(gpu)
    _modify(d, P.shape[0], op, value)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d.cpu().numpy(), HOST[op](P, value))


@pytest.mark.parametrize("mirroring_y", [False, True])
def test_ptm_adjust_sequence_matches_fp64_host(gpu, mirroring_y):
    """MoveU(borderLeft), MoveV(borderUp), ReverseX, [ReverseY], MoveZ(-CenterZ)."""
    P = _matrices()
    bl, bu, cz = 12, 5, 0.15 * 7.5
    d = torch.from_numpy(P.copy()).to(gpu)
    _modify(d, P.shape[0], _lib.JH_PTM_MOVE_U, float(bl))
    _modify(d, P.shape[0], _lib.JH_PTM_MOVE_V, float(bu))
    _modify(d, P.shape[0], _lib.JH_PTM_REVERSE_X, 0.0)
    if mirroring_y:
        _modify(d, P.shape[0], _lib.JH_PTM_REVERSE_Y, 0.0)
    _modify(d, P.shape[0], _lib.JH_PTM_MOVE_Z, -cz)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d.cpu().numpy(), geometry.adjust_ptm(P, bl, bu, mirroring_y, cz))


def test_ptm_modify_rejects_bad_op(gpu):
    d = torch.zeros(12, dtype=torch.float64, device=gpu)
    with pytest.raises(_lib.JhError):
        _modify(d, 1, 5, 0.0)
