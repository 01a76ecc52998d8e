"""Float64 restatement of the surrogate-assisted face flux and step — TEST
INFRASTRUCTURE ONLY.

The learned flux has no counterpart in the reference: it replaces the HLLD solve
(glm_physics.jl:317-348) on faces whose jumps are small. On every face, after the floors
and the GLM pair exactly as in the reference's riemann (glm_physics.jl:317-323), in the
face's normal frame (rho, un, ua, ub, p, bn*, ba, bb):
  gate    |dRho| < thr rho, |dp| < thr p, ||B_R| - |B_L|| <= thr (|B_L|+|B_R|)/2 (face means),
          and the low-density LLF switch off;
  inputs  x[2v] = mean_v / scale_v, x[2v+1] = (R_v - L_v) / (2 thr scale_v),
          scale = (rho, a, a, a, rho a^2, a sqrt(rho) x3), a^2 = (gamma p + |B|^2) / rho;
  flux    (F(L) + F(R))/2 + s * MLP(x), s = thr (rho a, rho a^2 x3, rho a^3, 0, a^2 sqrt(rho) x2, 0),
          with the GLM slots (normal B, psi) the exact glm_pair values;
and HLLD (or LLF under the switch) elsewhere. The network is mlp_np's (bfloat16 operands).
Pinned in tests/test_surrogate.py by: the target encoding inverting to HLLD exactly, the
central flux being HLLD's for equal states, thr = 0 reducing to the plain step of
face_np, and the network part by mlp_np's own finite-difference check.
"""
from __future__ import annotations

import numpy as np

import face_np as F
import mlp_np as R


def prepare(Lq, Rq, d, ch, smallr=1e-6, pfl=1e-7, llf_dmin=1e-4, llf_pmin=0.0):
    """Normal-frame states after floors and the GLM pair; exact GLM flux slots; LLF switch."""
    L = Lq[F.ROT_TO[d]].astype(np.float64).copy()
    Rr = Rq[F.ROT_TO[d]].astype(np.float64).copy()
    L[0] = np.maximum(L[0], smallr)
    L[4] = np.maximum(L[4], pfl)
    Rr[0] = np.maximum(Rr[0], smallr)
    Rr[4] = np.maximum(Rr[4], pfl)
    bns, psis = F.glm_pair(L[5], Rr[5], L[8], Rr[8], ch)
    L[5] = bns
    Rr[5] = bns
    use = np.zeros(L[0].shape, dtype=bool)
    if llf_dmin > 0:
        use |= np.minimum(L[0], Rr[0]) < llf_dmin
    if llf_pmin > 0:
        use |= np.minimum(L[4], Rr[4]) < llf_pmin
    return L, Rr, psis, ch * ch * bns, use


def encode(L, Rr, g, thr):
    """(x[16, ...], s[9, ...], gate, margin): margin is the smallest relative distance of the
    three jump tests from their thresholds (negative when a test fails)."""
    rho = 0.5 * (L[0] + Rr[0])
    p = 0.5 * (L[4] + Rr[4])
    bn = L[5]
    by, bz = 0.5 * (L[6] + Rr[6]), 0.5 * (L[7] + Rr[7])
    a2 = (g * p + bn * bn + by * by + bz * bz) / rho
    a = np.sqrt(a2)
    sr = np.sqrt(rho)
    scale = [rho, a, a, a, rho * a2, a * sr, a * sr, a * sr]
    x = np.empty((16,) + rho.shape)
    with np.errstate(divide="ignore", invalid="ignore"):
        for v in range(8):
            x[2 * v] = 0.5 * (L[v] + Rr[v]) / scale[v]
            x[2 * v + 1] = (Rr[v] - L[v]) / scale[v] / (2.0 * thr)
    s = np.zeros((9,) + rho.shape)
    s[0] = thr * rho * a
    s[1] = s[2] = s[3] = thr * rho * a2
    s[4] = thr * rho * a2 * a
    s[6] = s[7] = thr * a2 * sr
    bL = np.sqrt(bn * bn + L[6] ** 2 + L[7] ** 2)
    bR = np.sqrt(bn * bn + Rr[6] ** 2 + Rr[7] ** 2)
    lim = [thr * rho, thr * p, thr * 0.5 * (bL + bR)]
    jump = [np.abs(Rr[0] - L[0]), np.abs(Rr[4] - L[4]), np.abs(bR - bL)]
    gate = (jump[0] < lim[0]) & (jump[1] < lim[1]) & (jump[2] <= lim[2])
    with np.errstate(divide="ignore", invalid="ignore"):
        mg = [(lm - j) / np.maximum(lm, 1e-300) for lm, j in zip(lim, jump)]
    mg[2] = np.where((lim[2] == 0) & (jump[2] == 0), 1.0, mg[2])  # unmagnetised: passes exactly
    margin = np.min(mg, axis=0)
    return x, s, gate, margin


def central(L, Rr, g):
    """(F(L) + F(R)) / 2 in the normal frame (glm_physics.jl:28-34), psi slot zero."""
    qL = np.concatenate([L[:8], np.zeros((1,) + L.shape[1:])])
    qR = np.concatenate([Rr[:8], np.zeros((1,) + Rr.shape[1:])])
    return 0.5 * (F.phys_flux_x(qL, g) + F.phys_flux_x(qR, g))


def face_flux(Lq, Rq, d, ch, theta, thr, g=5.0 / 3.0, smallr=1e-6, pfl=1e-7, llf_dmin=1e-4, llf_pmin=0.0,
              round_bf16=True):
    """Lab-frame gated flux of lab-frame face states (9, n): (flux, gate, margin)."""
    L, Rr, fbn, fpsi, use = prepare(Lq, Rq, d, ch, smallr, pfl, llf_dmin, llf_pmin)
    shape = L.shape[1:]
    L2, R2 = L.reshape(9, -1), Rr.reshape(9, -1)
    x, s, gate, margin = encode(L2, R2, g, thr)
    gate &= ~use.reshape(-1)
    with np.errstate(all="ignore"):
        fh = F.hlld_x(L2, R2, g, L2[5], fbn.reshape(-1), fpsi.reshape(-1))
        fl = F.llf_x(L2, R2, g, ch, fbn.reshape(-1), fpsi.reshape(-1))
    f = np.where(use.reshape(-1), fl, fh)
    if gate.any():
        y, _ = R.forward(theta, x[:, gate].T, round_bf16=round_bf16)
        fs = central(L2[:, gate], R2[:, gate], g) + s[:, gate] * y.T
        fs[5] = fbn.reshape(-1)[gate]
        fs[8] = fpsi.reshape(-1)[gate]
        f[:, gate] = fs
    f = f.reshape((9,) + shape)[F.ROT_FROM[d]]
    return f, gate.reshape(shape), margin.reshape(shape)


def target(Lq, Rq, d, ch, thr, g=5.0 / 3.0, smallr=1e-6, pfl=1e-7):
    """The training target y = (F_HLLD - central) / s on the learned slots (normal frame)."""
    L, Rr, fbn, fpsi, _ = prepare(Lq, Rq, d, ch, smallr, pfl, 0.0, 0.0)
    x, s, gate, _ = encode(L, Rr, g, thr)
    fh = F.hlld_x(L, Rr, g, L[5], fbn, fpsi)
    y = np.zeros_like(fh)
    keep = s != 0
    y[keep] = (fh[keep] - central(L, Rr, g)[keep]) / s[keep]
    return x, y, s, gate


def step_periodic(U, dims, dtdx, ch, decay, theta, thr, g=5.0 / 3.0, recon="plm_minmod", round_bf16=True):
    """One surrogate-assisted GLM step (periodic): (U_new, gates, margins), gates and
    margins per direction on the lower face of every cell, (nz, ny, nx)."""
    gates, margins = [], []

    def flux(L, Rq, d):
        f, gt, mg = face_flux(L, Rq, d, ch, theta, thr, g, round_bf16=round_bf16)
        gates.append(gt)
        margins.append(mg)
        return f

    out = F.step_periodic(U, dims, dtdx, ch, decay, g, recon=recon, flux=flux)
    return out, gates, margins
