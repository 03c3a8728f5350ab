"""The reference's on-disk formats, for Python hosts and tests.

* ``PTM.dat``: 12 whitespace-separated numbers per view, the row-major 3x4
  projection matrix (JEngine/PTMReader.h:17-23);
* raw projections: one file per view named by a template with one integer
  field (``Proj_%d.prj``), NumDetsU x NumDetsV little-endian uint16, u
  fastest (JEngine/InputAgent.cpp:50-63);
* slices: one file per z slice (``slice_%d.slc``), X x Y little-endian
  float32, x fastest (JEngine/OutputAgent.cpp:53-64).

The C++ host reads and writes the same formats (jh/FileAgents.h).
"""
import os

import numpy as np


def write_ptm_dat(path, ptm):
    """Writes N x 12 matrices, one view per line, with round-trip precision."""
    P = np.asarray(ptm, np.float64).reshape(-1, 12)
    with open(path, "w") as f:
        for row in P:
            f.write(" ".join(repr(float(x)) for x in row) + "\n")


def read_ptm_dat(path, n_views):
    with open(path) as f:
        vals = f.read().split()
    if len(vals) < 12 * n_views:
        raise ValueError(f"{path} holds {len(vals)} numbers, {12 * n_views} needed")
    return np.array([float(v) for v in vals[:12 * n_views]], np.float64).reshape(n_views, 12)


def write_projection_folder(folder, raw, template="Proj_%d.prj"):
    """raw: (N, NumDetsV, NumDetsU) uint16 -> one file per view."""
    os.makedirs(folder, exist_ok=True)
    raw = np.ascontiguousarray(raw, np.uint16)
    for v in range(raw.shape[0]):
        raw[v].astype("<u2").tofile(os.path.join(folder, template % v))


def read_projection_folder(folder, n_views, n_u, n_v, template="Proj_%d.prj"):
    out = np.empty((n_views, n_v, n_u), np.uint16)
    for v in range(n_views):
        a = np.fromfile(os.path.join(folder, template % v), "<u2")
        if a.size != n_u * n_v:
            raise ValueError(f"{template % v}: {a.size} values, {n_u * n_v} expected")
        out[v] = a.reshape(n_v, n_u)
    return out


def write_slices(folder, vol, template="slice_%d.slc"):
    """vol: (Z, Y, X) float32 -> one file per slice."""
    os.makedirs(folder, exist_ok=True)
    vol = np.ascontiguousarray(vol, np.float32)
    for z in range(vol.shape[0]):
        vol[z].astype("<f4").tofile(os.path.join(folder, template % z))


def read_slices(folder, n_slices, X, Y, template="slice_%d.slc"):
    out = np.empty((n_slices, Y, X), np.float32)
    for z in range(n_slices):
        a = np.fromfile(os.path.join(folder, template % z), "<f4")
        if a.size != X * Y:
            raise ValueError(f"{template % z}: {a.size} values, {X * Y} expected")
        out[z] = a.reshape(Y, X)
    return out
