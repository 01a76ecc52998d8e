"""Host set-up of a CT state's face field from its cell-centred B (float64 numpy FFT).

Restates ct_faces_from_cells! (lib/MHDKernels/src/ct.jl:246-308): with theta_d = k_d dx/2
the cell field is projected orthogonally onto t . b = 0 (t_d = tan theta_d, the smallest
L2 change of the cell field that admits divergence-free faces), components whose axis
sits on its Nyquist plane are dropped, and the face modes are b_d exp(-i theta_d)/cos theta_d,
which average back to the projected cell field exactly.
"""
from __future__ import annotations

import numpy as np


def _theta(n: int, m: np.ndarray):
    """Half-cell phase of 0-based FFT index m on an n-point axis and its Nyquist flag
    (ct.jl:240-244)."""
    nyq = (n % 2 == 0) & (m == n // 2)
    signed = np.where(m > n / 2, m - n, m)
    return np.where(nyq, 0.0, np.pi * signed / n), nyq


def faces_from_cells(bx, by, bz):
    """Cell fields of shape (nz, ny, nx) -> lower-face fields of the same shape."""
    nz, ny, nx = bx.shape
    tx, qx = _theta(nx, np.arange(nx // 2 + 1))
    ty, qy = _theta(ny, np.arange(ny))
    tz, qz = _theta(nz, np.arange(nz))
    th = (tx[None, None, :], ty[None, :, None], tz[:, None, None])
    nyq = (qx[None, None, :], qy[None, :, None], qz[:, None, None])
    spec = []
    for d, b in enumerate((bx, by, bz)):
        s = np.fft.rfftn(np.asarray(b, dtype=np.float64), axes=(0, 1, 2))
        spec.append(np.where(nyq[d], 0.0, s))
    t = [np.tan(a) for a in th]
    t2 = t[0] ** 2 + t[1] ** 2 + t[2] ** 2
    proj = (t[0] * spec[0] + t[1] * spec[1] + t[2] * spec[2]) / np.where(t2 > 0, t2, 1.0)
    proj = np.where(t2 > 0, proj, 0.0)
    faces = []
    for d in range(3):
        fd = (spec[d] - t[d] * proj) * np.exp(-1j * th[d]) / np.cos(th[d])
        faces.append(np.fft.irfftn(fd, s=(nz, ny, nx), axes=(0, 1, 2)))
    return faces
