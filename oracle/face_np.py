"""Second, independent restatement of the GLM-HLLD step in vectorised numpy — TEST
INFRASTRUCTURE ONLY.

It is written array-at-a-time (whole grid per operation, neighbours by ``np.roll``),
unlike the per-cell C++ oracle, so a transcription slip in either shows up as a
disagreement between the two. Follows lib/MHDKernels/src:
  glm_physics.jl:11-61 (cons2prim, prim2cons, fast_speed, phys_flux_x, rotations, limiters),
  glm_physics.jl:188-231 (PPM/PLM offsets, Hancock), 234-348 (GLM pair, LLF, HLL, HLLD, riemann),
  integrator_ref.jl:69-160 (face states and the update grouping) for periodic grids.
Also holds ``fork_ct_step``: a restatement of the independent plain-Julia port of the
GLMMHDTurb fork CT step that the reference uses to check its CT solver
(test/inherited/test_ct.jl:10-144).

Arrays are (9, ...) with components first.
"""
from __future__ import annotations

import numpy as np


def cons2prim(c, g, smallr, pfl):
    rho = np.maximum(c[0], smallr)
    ir = 1.0 / rho
    ekin = 0.5 * (c[1] * c[1] + c[2] * c[2] + c[3] * c[3]) * ir
    emag = 0.5 * (c[5] * c[5] + c[6] * c[6] + c[7] * c[7])
    p = np.maximum((g - 1.0) * (c[4] - ekin - emag), pfl)
    return np.stack([rho, c[1] * ir, c[2] * ir, c[3] * ir, p, c[5], c[6], c[7], c[8]])


def prim2cons(q, g):
    rho, vx, vy, vz, p, bx, by, bz, psi = q
    E = p / (g - 1.0) + 0.5 * rho * (vx * vx + vy * vy + vz * vz) + 0.5 * (bx * bx + by * by + bz * bz)
    return np.stack([rho, rho * vx, rho * vy, rho * vz, E, bx, by, bz, psi])


def fast_speed(q, g, bn):
    rho = q[0]
    c2 = g * q[4] / rho
    b2 = (q[5] * q[5] + q[6] * q[6] + q[7] * q[7]) / rho
    d2 = 0.5 * (b2 + c2)
    return np.sqrt(d2 + np.sqrt(np.maximum(d2 * d2 - c2 * bn * bn / rho, 0.0)))


def phys_flux_x(q, g):
    rho, vx, vy, vz, p, bx, by, bz, _ = q
    b2 = bx * bx + by * by + bz * bz
    ptot = p + 0.5 * b2
    E = p / (g - 1.0) + 0.5 * rho * (vx * vx + vy * vy + vz * vz) + 0.5 * b2
    vb = vx * bx + vy * by + vz * bz
    z = np.zeros_like(rho)
    return np.stack([rho * vx, rho * vx * vx + ptot - bx * bx, rho * vx * vy - bx * by, rho * vx * vz - bx * bz,
                     (E + ptot) * vx - bx * vb, z, vx * by - vy * bx, vx * bz - vz * bx, z])


# cyclic permutations (glm_physics.jl:36-45) as index lists
ROT_TO = ([0, 1, 2, 3, 4, 5, 6, 7, 8], [0, 2, 3, 1, 4, 6, 7, 5, 8], [0, 3, 1, 2, 4, 7, 5, 6, 8])
ROT_FROM = ([0, 1, 2, 3, 4, 5, 6, 7, 8], [0, 3, 1, 2, 4, 7, 5, 6, 8], [0, 2, 3, 1, 4, 6, 7, 5, 8])


def dir_flux(q, d, g):
    return phys_flux_x(q[ROT_TO[d]], g)[ROT_FROM[d]]


def moncen(dl, dr):
    dc = 0.5 * (dl + dr)
    sgn = np.where(dc >= 0, 1.0, -1.0)
    val = sgn * np.minimum(2.0 * np.minimum(np.abs(dl), np.abs(dr)), np.abs(dc))
    return np.where(dl * dr <= 0, 0.0, val)


def minmod(dl, dr):
    sgn = np.where(dl >= 0, 1.0, -1.0)
    return np.where(dl * dr <= 0, 0.0, sgn * np.minimum(np.abs(dl), np.abs(dr)))


def recon_offsets(mm, m0, mp, recon):
    if recon == "ppm":
        slope = (mp - mm) * 0.25
        curve = (mm - 2.0 * m0 + mp) / 12.0
        qL = m0 - slope + curve
        qR = m0 + slope + curve
        qa = m0
        dq = qR - qL
        qmid = 0.5 * (qL + qR)
        dq2_6 = dq * dq / 6.0
        diff = (qa - qmid) * dq
        flat = (qR - qa) * (qa - qL) <= 0
        hi = (~flat) & (diff > dq2_6)
        lo = (~flat) & (~hi) & (diff < -dq2_6)
        eL = np.where(flat, qa, np.where(hi, 3.0 * qa - 2.0 * qR, qL))
        eR = np.where(flat, qa, np.where(lo, 3.0 * qa - 2.0 * qL, qR))
        return eL - m0, eR - m0
    lim = minmod if recon == "plm_minmod" else moncen
    s = 0.5 * lim(m0 - mm, mp - m0)
    return -s, s


def hancock(m0, dL, dR, dtdx, g):
    u0 = prim2cons(m0, g)
    h = 0.5 * dtdx
    f = [dir_flux(m0 + dR[d], d, g) - dir_flux(m0 + dL[d], d, g) for d in range(3)]
    return u0 - h * ((f[0] + f[1]) + f[2])


def glm_pair(bnL, bnR, psiL, psiR, ch):
    return 0.5 * (bnL + bnR) - 0.5 * (psiR - psiL) / ch, 0.5 * (psiL + psiR) - 0.5 * ch * (bnR - bnL)


def llf_x(L, R, g, ch, fbn, fpsi):
    fL, fR = phys_flux_x(L, g), phys_flux_x(R, g)
    uL, uR = prim2cons(L, g), prim2cons(R, g)
    smax = np.maximum(np.maximum(np.abs(L[1]) + fast_speed(L, g, L[5]), np.abs(R[1]) + fast_speed(R, g, R[5])), ch)
    f = 0.5 * (fL + fR) - 0.5 * smax * (uR - uL)
    f[5] = fbn
    f[8] = fpsi
    return f


def hll_x(L, R, g, bn, fbn, fpsi):
    cfL, cfR = fast_speed(L, g, bn), fast_speed(R, g, bn)
    SL = np.minimum(np.minimum(L[1] - cfL, R[1] - cfR), 0.0)
    SR = np.maximum(np.maximum(L[1] + cfL, R[1] + cfR), 0.0)
    fL, fR = phys_flux_x(L, g), phys_flux_x(R, g)
    uL, uR = prim2cons(L, g), prim2cons(R, g)
    f = (SR * fL - SL * fR + SL * SR * (uR - uL)) * (1.0 / (SR - SL))
    f[5] = fbn
    f[8] = fpsi
    return f


def _star(d, u, v, w, by, bz, S, bn, SM):
    denK = d * (S - u) * (S - SM) - bn * bn
    deg = np.abs(denK) < 1e-12
    safe = np.where(deg, 1.0, denK)
    fac = bn * (SM - u) / safe
    num = d * (S - u) * (S - u) - bn * bn
    return (np.where(deg, v, v - by * fac), np.where(deg, w, w - bz * fac),
            np.where(deg, by, by * num / safe), np.where(deg, bz, bz * num / safe))


def hlld_x(L, R, g, bn, fbn, fpsi):
    dL, uL, vL, wL, pL, _, byL, bzL, _ = L
    dR, uR, vR, wR, pR, _, byR, bzR, _ = R
    z = np.zeros_like(dL)
    b2L = bn * bn + byL * byL + bzL * bzL
    b2R = bn * bn + byR * byR + bzR * bzR
    ptL, ptR = pL + 0.5 * b2L, pR + 0.5 * b2R
    EL = pL / (g - 1.0) + 0.5 * dL * (uL * uL + vL * vL + wL * wL) + 0.5 * b2L
    ER = pR / (g - 1.0) + 0.5 * dR * (uR * uR + vR * vR + wR * wR) + 0.5 * b2R
    cfL, cfR = fast_speed(L, g, bn), fast_speed(R, g, bn)
    SL = np.minimum(np.minimum(uL, uR) - np.maximum(cfL, cfR), 0.0)
    SR = np.maximum(np.maximum(uL, uR) + np.maximum(cfL, cfR), 0.0)
    UL = np.stack([dL, dL * uL, dL * vL, dL * wL, EL, bn, byL, bzL, z])
    UR = np.stack([dR, dR * uR, dR * vR, dR * wR, ER, bn, byR, bzR, z])
    FL = phys_flux_x(np.stack([dL, uL, vL, wL, pL, bn, byL, bzL, z]), g)
    FR = phys_flux_x(np.stack([dR, uR, vR, wR, pR, bn, byR, bzR, z]), g)
    den = (SR - uR) * dR - (SL - uL) * dL
    SM = ((SR - uR) * dR * uR - (SL - uL) * dL * uL - ptR + ptL) / den
    pts = ((SR - uR) * dR * ptL - (SL - uL) * dL * ptR + dL * dR * (SR - uR) * (SL - uL) * (uR - uL)) / den
    dLs = np.maximum(dL * (SL - uL) / (SL - SM), 1e-12)
    dRs = np.maximum(dR * (SR - uR) / (SR - SM), 1e-12)
    sL, sR = np.sqrt(dLs), np.sqrt(dRs)
    SLs, SRs = SM - np.abs(bn) / sL, SM + np.abs(bn) / sR
    vyLs, vzLs, byLs, bzLs = _star(dL, uL, vL, wL, byL, bzL, SL, bn, SM)
    vyRs, vzRs, byRs, bzRs = _star(dR, uR, vR, wR, byR, bzR, SR, bn, SM)
    vbL, vbLs = uL * bn + vL * byL + wL * bzL, SM * bn + vyLs * byLs + vzLs * bzLs
    ELs = ((SL - uL) * EL - ptL * uL + pts * SM + bn * (vbL - vbLs)) / (SL - SM)
    vbR, vbRs = uR * bn + vR * byR + wR * bzR, SM * bn + vyRs * byRs + vzRs * bzRs
    ERs = ((SR - uR) * ER - ptR * uR + pts * SM + bn * (vbR - vbRs)) / (SR - SM)
    ULs = np.stack([dLs, dLs * SM, dLs * vyLs, dLs * vzLs, ELs, bn, byLs, bzLs, z])
    URs = np.stack([dRs, dRs * SM, dRs * vyRs, dRs * vzRs, ERs, bn, byRs, bzRs, z])
    sg = np.where(bn >= 0, 1.0, -1.0)
    inv = 1.0 / (sL + sR)
    vyss = (sL * vyLs + sR * vyRs + (byRs - byLs) * sg) * inv
    vzss = (sL * vzLs + sR * vzRs + (bzRs - bzLs) * sg) * inv
    byss = (sL * byRs + sR * byLs + sL * sR * (vyRs - vyLs) * sg) * inv
    bzss = (sL * bzRs + sR * bzLs + sL * sR * (vzRs - vzLs) * sg) * inv
    vbss = SM * bn + vyss * byss + vzss * bzss
    ELss = ELs - sL * (vbLs - vbss) * sg
    ERss = ERs + sR * (vbRs - vbss) * sg
    ULss = np.stack([dLs, dLs * SM, dLs * vyss, dLs * vzss, ELss, bn, byss, bzss, z])
    URss = np.stack([dRs, dRs * SM, dRs * vyss, dRs * vzss, ERss, bn, byss, bzss, z])
    f1 = FL + SL * (ULs - UL)
    f2 = FL + SLs * ULss - (SLs - SL) * ULs - SL * UL
    f3 = FR + SRs * URss - (SRs - SR) * URs - SR * UR
    f4 = FR + SR * (URs - UR)
    f = np.where(SLs >= 0, f1, np.where(SM >= 0, f2, np.where(SRs >= 0, f3, f4)))
    f[5] = fbn
    f[8] = fpsi
    return f


def riemann(Lq, Rq, d, ch, g, smallr=1e-6, pfl=1e-7, llf_dmin=1e-4, llf_pmin=0.0, solver="hlld"):
    L = Lq[ROT_TO[d]].copy()
    R = Rq[ROT_TO[d]].copy()
    L[0] = np.maximum(L[0], smallr)
    L[4] = np.maximum(L[4], pfl)
    R[0] = np.maximum(R[0], smallr)
    R[4] = np.maximum(R[4], pfl)
    bns, psis = glm_pair(L[5], R[5], L[8], R[8], ch)
    fbn, fpsi = psis, ch * ch * bns
    L[5] = bns
    R[5] = bns
    if solver == "llf":
        f = llf_x(L, R, g, ch, fbn, fpsi)
    elif solver == "hll":
        f = hll_x(L, R, g, bns, fbn, fpsi)
    else:
        use = np.zeros(L[0].shape, dtype=bool)
        if llf_dmin > 0:
            use |= np.minimum(L[0], R[0]) < llf_dmin
        if llf_pmin > 0:
            use |= np.minimum(L[4], R[4]) < llf_pmin
        with np.errstate(all="ignore"):
            fh = hlld_x(L, R, g, bns, fbn, fpsi)
        f = np.where(use, llf_x(L, R, g, ch, fbn, fpsi), fh) if use.any() else fh
    return f[ROT_FROM[d]]


def hlld_branches(Lq, Rq, d, ch, g, smallr=1e-6, pfl=1e-7, llf_dmin=1e-4, llf_pmin=0.0):
    """Which part of the HLLD dispatch each face takes (glm_physics.jl:255-307, 320-327):
    -1 where the LLF switch takes the face, else 0..3 for the S_L*>=0 / S_M>=0 / S_R*>=0 /
    S_R* < 0 flux branch; and the mask of faces with a degenerate single-star state
    (|denK| < 1e-12 on either side). Test instrumentation: no flux is computed."""
    L = Lq[ROT_TO[d]].copy()
    R = Rq[ROT_TO[d]].copy()
    for q in (L, R):
        q[0] = np.maximum(q[0], smallr)
        q[4] = np.maximum(q[4], pfl)
    bn, _ = glm_pair(L[5], R[5], L[8], R[8], ch)
    dL, uL, pL, byL, bzL = L[0], L[1], L[4], L[6], L[7]
    dR, uR, pR, byR, bzR = R[0], R[1], R[4], R[6], R[7]
    ptL = pL + 0.5 * (bn * bn + byL * byL + bzL * bzL)
    ptR = pR + 0.5 * (bn * bn + byR * byR + bzR * bzR)
    cf = np.maximum(fast_speed(L, g, bn), fast_speed(R, g, bn))
    SL = np.minimum(np.minimum(uL, uR) - cf, 0.0)
    SR = np.maximum(np.maximum(uL, uR) + cf, 0.0)
    with np.errstate(all="ignore"):
        den = (SR - uR) * dR - (SL - uL) * dL
        SM = ((SR - uR) * dR * uR - (SL - uL) * dL * uL - ptR + ptL) / den
        dLs = np.maximum(dL * (SL - uL) / (SL - SM), 1e-12)
        dRs = np.maximum(dR * (SR - uR) / (SR - SM), 1e-12)
        SLs, SRs = SM - np.abs(bn) / np.sqrt(dLs), SM + np.abs(bn) / np.sqrt(dRs)
        degen = (np.abs(dL * (SL - uL) * (SL - SM) - bn * bn) < 1e-12) | \
                (np.abs(dR * (SR - uR) * (SR - SM) - bn * bn) < 1e-12)
    br = np.where(SLs >= 0, 0, np.where(SM >= 0, 1, np.where(SRs >= 0, 2, 3)))
    use = np.zeros(dL.shape, dtype=bool)
    if llf_dmin > 0:
        use |= np.minimum(dL, dR) < llf_dmin
    if llf_pmin > 0:
        use |= np.minimum(pL, pR) < llf_pmin
    return np.where(use, -1, br), degen & ~use


def step_periodic(U, dims, dtdx, ch, decay, g=5.0 / 3.0, smallr=1e-6, pfl=1e-7, llf_dmin=1e-4, llf_pmin=0.0,
                  recon="plm", solver="hlld", flux=None):
    """One GLM step of a periodic grid; U is (9, nc) conserved, i fastest. ``flux(L, R, d)``
    replaces the Riemann solver when given (lab-frame face states in, lab-frame flux out)."""
    nx, ny, nz = dims
    u = U.reshape(9, nz, ny, nx)
    ax = (3, 2, 1)  # numpy axis of x, y, z in the (9, nz, ny, nx) layout
    P = cons2prim(u, g, smallr, pfl)
    dL, dR = [], []
    for d in range(3):
        a, b = recon_offsets(np.roll(P, 1, axis=ax[d]), P, np.roll(P, -1, axis=ax[d]), recon)
        dL.append(a)
        dR.append(b)
    mh = cons2prim(hancock(P, dL, dR, dtdx, g), g, smallr, pfl)
    dU = []
    for d in range(3):
        # lower face of every cell: L = mh(i-1) + dR(i-1), R = mh(i) + dL(i)
        L = np.roll(mh + dR[d], 1, axis=ax[d])
        R = mh + dL[d]
        if flux is None:
            F = riemann(L, R, d, ch, g, smallr, pfl, llf_dmin, llf_pmin, solver)
        else:
            F = flux(L, R, d)
        dU.append(F - np.roll(F, -1, axis=ax[d]))
    out = u + dtdx * ((dU[0] + dU[1]) + dU[2])
    out[8] = out[8] * decay
    return out.reshape(9, -1)


# ---------------------------------------------------------------------------
# Restatement of the fork-port CT step (test/inherited/test_ct.jl:10-144): full MonCen
# slopes, 3-direction Hancock, HLL per face with the face normal field, GS05 edge EMFs,
# hydro update of 5 variables and the face curl. U: (5, nz, ny, nx); faces (nz, ny, nx).
FK_SMALLR = 1e-6
FK_SMALLP = 1e-6


def _fk_prim(U, bcc, g):
    rho = np.maximum(U[0], FK_SMALLR)
    ir = 1.0 / rho
    vx, vy, vz = U[1] * ir, U[2] * ir, U[3] * ir
    bx, by, bz = bcc
    p = np.maximum((g - 1.0) * (U[4] - 0.5 * rho * (vx * vx + vy * vy + vz * vz) - 0.5 * (bx * bx + by * by + bz * bz)),
                   FK_SMALLP)
    return np.stack([rho, vx, vy, vz, p, bx, by, bz])


def _fk_dflux(q, d, g):
    rho, vx, vy, vz, p, bx, by, bz = q
    b2 = bx * bx + by * by + bz * bz
    ptot = p + 0.5 * b2
    E = p / (g - 1.0) + 0.5 * rho * (vx * vx + vy * vy + vz * vz) + 0.5 * b2
    vb = vx * bx + vy * by + vz * bz
    v = (vx, vy, vz)
    b = (bx, by, bz)
    un, bn = v[d], b[d]
    Fm = [rho * un * v[c] - bn * b[c] + (ptot if c == d else 0.0) for c in range(3)]
    Fb = [un * b[c] - (un * b[c] if c == d else v[c] * bn) for c in range(3)]
    return np.stack([rho * un, Fm[0], Fm[1], Fm[2], (E + ptot) * un - bn * vb, Fb[0], Fb[1], Fb[2]])


def _fk_fast(q, d, g):
    rho = q[0]
    bn = q[5 + d]
    c2 = g * q[4] / rho
    b2 = (q[5] * q[5] + q[6] * q[6] + q[7] * q[7]) / rho
    dd = 0.5 * (b2 + c2)
    return np.sqrt(dd + np.sqrt(np.maximum(dd * dd - c2 * bn * bn / rho, 0.0)))


def _fk_toC(q, g):
    rho, vx, vy, vz, p, bx, by, bz = q
    return np.stack([rho, rho * vx, rho * vy, rho * vz,
                     p / (g - 1.0) + 0.5 * rho * (vx * vx + vy * vy + vz * vz) + 0.5 * (bx * bx + by * by + bz * bz),
                     bx, by, bz])


def _fk_hll(L, R, d, g):
    cfL, cfR = _fk_fast(L, d, g), _fk_fast(R, d, g)
    SL = np.minimum(np.minimum(L[1 + d] - cfL, R[1 + d] - cfR), 0.0)
    SR = np.maximum(np.maximum(L[1 + d] + cfL, R[1 + d] + cfR), 0.0)
    FL, FR = _fk_dflux(L, d, g), _fk_dflux(R, d, g)
    UL, UR = _fk_toC(L, g), _fk_toC(R, g)
    return (SR * FL - SL * FR + SL * SR * (UR - UL)) * (1.0 / (SR - SL))


def _fk_gs05(Ea0, Eam, Eb0, Ebm, mA0, mAm, mB0, mBm, Er00, Erm0, Er0m, Ermm):
    def up(m, lo, hi):
        return np.where(m > 0, lo, np.where(m < 0, hi, 0.5 * (lo + hi)))
    A1 = up(mA0, Erm0 - Ebm, Er00 - Eb0)
    A3 = up(mAm, Ebm - Ermm, Eb0 - Er0m)
    B1 = up(mB0, Er0m - Eam, Er00 - Ea0)
    B3 = up(mBm, Eam - Ermm, Ea0 - Erm0)
    return 0.25 * ((Ea0 + Eam + Eb0 + Ebm) + (A3 - A1) + (B3 - B1))


def fork_ct_step(U, bxf, byf, bzf, dt, dx, g):
    """Advance (U[5], faces) one step in place; arrays are (nz, ny, nx) with x last."""
    dtdx = dt / dx
    ax = (2, 1, 0)  # x, y, z axes of an (nz, ny, nx) array

    def sh(a, d, s):  # value at index i+s along direction d
        return np.roll(a, -s, axis=ax[d] + (a.ndim - 3))

    bf = (bxf, byf, bzf)
    bcc = [0.5 * (bf[d] + sh(bf[d], d, 1)) for d in range(3)]
    q = _fk_prim(U, bcc, g)
    s = [np.stack([moncen(q[t] - sh(q[t], d, -1), sh(q[t], d, 1) - q[t]) for t in range(8)]) for d in range(3)]
    h = 0.5 * dtdx
    fl = [_fk_dflux(q + 0.5 * s[d], d, g) - _fk_dflux(q - 0.5 * s[d], d, g) for d in range(3)]
    Uh = _fk_toC(q, g) - h * ((fl[0] + fl[1]) + fl[2])
    rho = np.maximum(Uh[0], FK_SMALLR)
    ir = 1.0 / rho
    vx, vy, vz = Uh[1] * ir, Uh[2] * ir, Uh[3] * ir
    p = np.maximum((g - 1.0) * (Uh[4] - 0.5 * rho * (vx * vx + vy * vy + vz * vz)
                                - 0.5 * (Uh[5] * Uh[5] + Uh[6] * Uh[6] + Uh[7] * Uh[7])), FK_SMALLP)
    mh = np.stack([rho, vx, vy, vz, p, Uh[5], Uh[6], Uh[7]])
    F = []
    for d in range(3):
        L = sh(mh + 0.5 * s[d], d, -1)
        R = mh - 0.5 * s[d]
        L = L.copy()
        R = R.copy()
        L[0] = np.maximum(L[0], FK_SMALLR)
        L[4] = np.maximum(L[4], FK_SMALLP)
        R[0] = np.maximum(R[0], FK_SMALLR)
        R[4] = np.maximum(R[4], FK_SMALLP)
        L[5 + d] = bf[d]
        R[5 + d] = bf[d]
        F.append(_fk_hll(L, R, d, g))
    Fx, Fy, Fz = F
    ec = [mh[3] * bcc[1] - mh[2] * bcc[2], mh[1] * bcc[2] - mh[3] * bcc[0], mh[2] * bcc[0] - mh[1] * bcc[1]]

    def m(a, d):  # value at i-1 along d
        return sh(a, d, -1)

    ez = _fk_gs05(-Fx[6], -m(Fx[6], 1), Fy[5], m(Fy[5], 0), Fx[0], m(Fx[0], 1), Fy[0], m(Fy[0], 0),
                  ec[2], m(ec[2], 0), m(ec[2], 1), m(m(ec[2], 0), 1))
    ex = _fk_gs05(-Fy[7], -m(Fy[7], 2), Fz[6], m(Fz[6], 1), Fy[0], m(Fy[0], 2), Fz[0], m(Fz[0], 1),
                  ec[0], m(ec[0], 1), m(ec[0], 2), m(m(ec[0], 1), 2))
    ey = _fk_gs05(-Fz[5], -m(Fz[5], 0), Fx[7], m(Fx[7], 2), Fz[0], m(Fz[0], 0), Fx[0], m(Fx[0], 2),
                  ec[1], m(ec[1], 2), m(ec[1], 0), m(m(ec[1], 2), 0))
    for t in range(5):
        U[t] += dtdx * ((Fx[t] - sh(Fx[t], 0, 1)) + (Fy[t] - sh(Fy[t], 1, 1)) + (Fz[t] - sh(Fz[t], 2, 1)))
    bxf -= dtdx * ((sh(ez, 1, 1) - ez) - (sh(ey, 2, 1) - ey))
    byf -= dtdx * ((sh(ex, 2, 1) - ex) - (sh(ez, 0, 1) - ez))
    bzf -= dtdx * ((sh(ey, 0, 1) - ey) - (sh(ex, 1, 1) - ex))
