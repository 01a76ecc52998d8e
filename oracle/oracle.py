"""CPU oracle driver — TEST INFRASTRUCTURE ONLY.

Python front end over ``liboracle_mhd.so`` (the C++ restatement in ``mhd_oracle.cpp``)
with the reference's host semantics:

* ``OracleState`` mirrors ``MHDState``/``allocate_state`` (state.jl:63-79,123-141):
  9 SoA conserved arrays ``U`` in order (rho, mx, my, mz, E, Bx, By, Bz, psi), i fastest,
  plus the 21-slot CT buffer for ``mhd_solver="ct"``.
* ``step`` mirrors ``step!`` (driver.jl:20-37) -> ``step_ref!`` (integrator_ref.jl:210-220),
  including the decay factor computed in the state's precision and the U/scratch swap.
* ``compute_dt`` mirrors ``compute_dt!`` (timestep.jl:37-41); ``evolve`` mirrors
  ``evolve!`` (driver.jl:48-71).
* ``ct_faces_from_cells`` mirrors ct.jl:271-308 with a float64 numpy FFT.

Only the test suite, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may
import this module, and only as the checker.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_mhd.so")

GLM_CR = 0.18  # driver.jl:8
RECON = {"plm": 0, "ppm": 1, "plm_minmod": 2}  # glm_physics.jl:202-207
RIEMANN = {"hlld": 0, "hll": 1, "llf": 2}  # glm_physics.jl:309-314
BCS = {"periodic": 0, "outflow": 1, "reflecting": 2}  # boundary.jl:7-13

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++, OpenMP)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE] + (["-B"] if force else []), check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i64 = ctypes.c_int64
        d = ctypes.c_double
        ci = ctypes.c_int
        for suf in ("f64", "f32"):
            getattr(L, f"orc_step_{suf}").argtypes = [P, P, i64, i64, i64, P, ci, ci, d, d, d, d, d, d, d, d, P]
            getattr(L, f"orc_step_src_{suf}").argtypes = [P, P, i64, i64, i64, P, ci, ci, d, d, d, d, d, d, d, d, P,
                                                          ci, d, d, d, d, P]
            getattr(L, f"orc_max_wavespeed_{suf}").argtypes = [P, i64, i64, i64, d, d, d]
            getattr(L, f"orc_max_wavespeed_{suf}").restype = d
            getattr(L, f"orc_max_divb_{suf}").argtypes = [P, i64, i64, i64, P, d]
            getattr(L, f"orc_max_divb_{suf}").restype = d
            getattr(L, f"orc_ct_update_{suf}").argtypes = [P, P, i64, i64, i64, d, ci]
            getattr(L, f"orc_max_divb_face_{suf}").argtypes = [P, i64, i64, i64, d]
            getattr(L, f"orc_max_divb_face_{suf}").restype = d
            getattr(L, f"orc_ct_install_faces_{suf}").argtypes = [P, P, i64, i64, i64, P, P, P]
            getattr(L, f"orc_ct_curl_{suf}").argtypes = [P, P, P, P, P, P, i64, i64, i64, d]
            getattr(L, f"orc_riemann_batch_{suf}").argtypes = [P, P, P, i64, ci, ci, d, d, d, d, d, d, ci]
            getattr(L, f"orc_predict_batch_{suf}").argtypes = [P, P, P, P, i64, ci, d, d, d, d]
        L.orc_num_threads.restype = ci
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _suf(dtype) -> str:
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


class OracleError(RuntimeError):
    pass


class OracleState:
    """Host mirror of ``MHDState`` (state.jl:63-79)."""

    def __init__(self, dims, dx, dtype=np.float64, gamma=5.0 / 3.0, smallr=1e-6, pfl=1e-7,
                 llf_dmin=1e-4, llf_pmin=0.0, riemann="hlld", recon="plm", bcs=("periodic",) * 3,
                 mhd_solver="glm"):
        if mhd_solver not in ("glm", "ct"):
            raise OracleError(f"unknown mhd_solver :{mhd_solver} (have :glm, :ct)")
        if mhd_solver == "ct" and any(b != "periodic" for b in bcs):
            raise OracleError("mhd_solver=:ct requires all-periodic boundaries")
        self.dims = tuple(int(n) for n in dims)
        self.dtype = np.dtype(dtype)
        T = self.dtype.type
        self.dx = T(dx)
        self.gamma = T(gamma)
        self.smallr, self.pfl = T(smallr), T(pfl)
        self.llf_dmin, self.llf_pmin = T(llf_dmin), T(llf_pmin)
        self.riemann, self.recon, self.bcs = riemann, recon, tuple(bcs)
        self.nc = int(np.prod(self.dims))
        self.U = np.zeros((9, self.nc), dtype=self.dtype)
        self.scratch = np.zeros((9, self.nc), dtype=self.dtype)
        self.ct = np.zeros((21, self.nc), dtype=self.dtype) if mhd_solver == "ct" else None

    @property
    def uses_ct(self) -> bool:
        return self.ct is not None

    def set_conserved(self, U):
        self.U[...] = np.asarray(U, dtype=np.float64).reshape(9, self.nc).astype(self.dtype)

    def copy(self) -> "OracleState":
        o = OracleState.__new__(OracleState)
        o.__dict__.update(self.__dict__)
        o.U = self.U.copy()
        o.scratch = self.scratch.copy()
        o.ct = None if self.ct is None else self.ct.copy()
        return o

    # --- reference host API -------------------------------------------------
    def _bc(self):
        return (ctypes.c_int * 3)(*[BCS[b] for b in self.bcs])

    def step_ref(self, dt, ch, decay, compensate=False):
        """integrator_ref.jl:210-220: kernel -> swap -> ct_update!."""
        T = self.dtype.type
        dtdx = T(dt) / self.dx
        L = lib()
        nx, ny, nz = self.dims
        getattr(L, f"orc_step_{_suf(self.dtype)}")(
            _p(self.U), _p(self.scratch), nx, ny, nz, self._bc(), RECON[self.recon], RIEMANN[self.riemann],
            float(dtdx), float(T(ch)), float(T(decay)), float(self.gamma), float(self.smallr), float(self.pfl),
            float(self.llf_dmin), float(self.llf_pmin), _p(self.ct) if self.ct is not None else None)
        self.U, self.scratch = self.scratch, self.U
        if self.ct is not None:
            getattr(L, f"orc_ct_update_{_suf(self.dtype)}")(_p(self.U), _p(self.ct), nx, ny, nz, float(dtdx),
                                                             1 if compensate else 0)

    def step_sources(self, dt, ch, decay, drag_impulse=0.0, target_velocity=(0.0, 0.0, 0.0), correction=None,
                     track_density=False, track_magnetic=False, track_energy=False, check=0, drag=True,
                     recon=None, riemann=None):
        """step_ref_drag! (correction None) / step_ref_radiation! (integrator_ref.jl:224-258):
        kernel -> swap -> ct_update!(compensate=track_magnetic). `correction` is a (9, nc)
        array in the state's precision, updated in place (slots 4-8 when tracking). `check`
        follows the cube kernel: 1 flags inadmissible cells with a negative density, 2
        zeroes the slopes of cells whose slot 4 is positive and solves their faces with LLF,
        3 zeroes every slope (the caller passes riemann="llf", as step_cube_radiation! does).
        recon / riemann override the state's codes for this call."""
        T = self.dtype.type
        dtdx = T(dt) / self.dx
        L = lib()
        nx, ny, nz = self.dims
        rad = correction is not None
        if rad:
            assert correction.dtype == self.dtype and correction.shape == (9, self.nc)
            assert correction.flags["C_CONTIGUOUS"]
        flags = (1 if drag else 0) | (2 if rad else 0) | (4 if track_density else 0) | \
            (8 if track_magnetic else 0) | (16 if track_energy else 0) | (int(check) << 8)
        getattr(L, f"orc_step_src_{_suf(self.dtype)}")(
            _p(self.U), _p(self.scratch), nx, ny, nz, self._bc(), RECON[recon or self.recon],
            RIEMANN[riemann or self.riemann], float(dtdx), float(T(ch)), float(T(decay)), float(self.gamma),
            float(self.smallr), float(self.pfl), float(self.llf_dmin), float(self.llf_pmin),
            _p(self.ct) if self.ct is not None else None, flags, float(T(drag_impulse)),
            float(T(target_velocity[0])), float(T(target_velocity[1])), float(T(target_velocity[2])),
            _p(correction) if rad else None)
        self.U, self.scratch = self.scratch, self.U
        if self.ct is not None:
            getattr(L, f"orc_ct_update_{_suf(self.dtype)}")(_p(self.U), _p(self.ct), nx, ny, nz, float(dtdx),
                                                             1 if (rad and track_magnetic) else 0)

    def step(self, dt, ch=None, glm_cr=GLM_CR):
        """driver.jl:20-37 (step! with integrator=:ref)."""
        T = self.dtype.type
        if self.uses_ct:
            return self.step_ref(dt, T(0), T(1))
        if ch is None:
            raise OracleError("step! needs the GLM cleaning speed `ch` (GLM states)")
        decay = np.exp(-T(glm_cr) * T(ch) * T(dt) / self.dx)
        self.step_ref(dt, ch, decay)

    def max_wavespeed(self):
        nx, ny, nz = self.dims
        v = getattr(lib(), f"orc_max_wavespeed_{_suf(self.dtype)}")(
            _p(self.U), nx, ny, nz, float(self.gamma), float(self.smallr), float(self.pfl))
        return self.dtype.type(v)

    def compute_dt(self, cfl=0.4):
        """timestep.jl:37-41: dt = cfl*dx/max(smax, eps(T)) in the state's precision."""
        T = self.dtype.type
        smax = self.max_wavespeed()
        dt = T(cfl) * self.dx / max(smax, np.finfo(self.dtype).eps)
        return dt, smax

    def evolve(self, tfinal, cfl=0.4, glm_ch_fac=1.0, glm_cr=GLM_CR, callback=None):
        """driver.jl:48-71."""
        if not glm_ch_fac > 0:
            raise OracleError("glm_ch_fac must be positive")
        if not glm_cr >= 0:
            raise OracleError("glm_cr must be non-negative")
        T = self.dtype.type
        t, n, tf = 0.0, 0, float(tfinal)
        while t < tf * (1 - 1e-9):
            dt, smax = self.compute_dt(cfl)
            if self.uses_ct:
                ch = T(0)
                dtf = float(dt)
            else:
                ch = T(glm_ch_fac) * T(smax)
                dtf = float(dt) * float(smax) / float(max(T(smax), ch))
            if t + dtf > tf:
                dtf = tf - t
            self.step(dtf, ch=ch, glm_cr=glm_cr)
            t += dtf
            n += 1
            if callback is not None:
                callback(self, t, n)
        return t, n

    def max_divb(self):
        nx, ny, nz = self.dims
        return getattr(lib(), f"orc_max_divb_{_suf(self.dtype)}")(_p(self.U), nx, ny, nz, self._bc(),
                                                                  float(self.dx))

    def conserved_totals(self):
        """diagnostics.jl:44-55 (float64 sums x dV, psi excluded)."""
        dV = float(self.dx) ** 3
        names = ("mass", "momx", "momy", "momz", "energy", "Bx", "By", "Bz")
        return {nm: float(np.sum(self.U[v].astype(np.float64))) * dV for v, nm in enumerate(names)}

    # --- constrained transport ---------------------------------------------
    def max_divb_face(self):
        if not self.uses_ct:
            raise OracleError("max_divb_face needs a CT state (mhd_solver=:ct)")
        nx, ny, nz = self.dims
        return getattr(lib(), f"orc_max_divb_face_{_suf(self.dtype)}")(_p(self.ct), nx, ny, nz, float(self.dx))

    def ct_install_faces(self, bx, by, bz):
        nx, ny, nz = self.dims
        f = [np.ascontiguousarray(np.asarray(a).reshape(-1), dtype=self.dtype) for a in (bx, by, bz)]
        getattr(lib(), f"orc_ct_install_faces_{_suf(self.dtype)}")(_p(self.U), _p(self.ct), nx, ny, nz,
                                                                   _p(f[0]), _p(f[1]), _p(f[2]))

    def ct_faces_from_potential(self, Ax, Ay, Az):
        """ct.jl:332-345."""
        nx, ny, nz = self.dims
        A = [np.ascontiguousarray(np.asarray(a).reshape(-1), dtype=self.dtype) for a in (Ax, Ay, Az)]
        b = [np.zeros(self.nc, dtype=self.dtype) for _ in range(3)]
        getattr(lib(), f"orc_ct_curl_{_suf(self.dtype)}")(_p(A[0]), _p(A[1]), _p(A[2]), _p(b[0]), _p(b[1]),
                                                          _p(b[2]), nx, ny, nz, float(self.dx))
        self.ct_install_faces(*b)

    def ct_faces_from_cells(self):
        """ct.jl:271-308 (float64 rfft projection with the tan(k dx/2) symbol)."""
        nx, ny, nz = self.dims
        Bold = [self.U[5 + d].astype(np.float64).reshape(nz, ny, nx) for d in range(3)]
        faces, _ = faces_from_cells_fft(Bold)
        self.ct_install_faces(*[f.reshape(-1) for f in faces])
        Bnew = [self.U[5 + d].astype(np.float64) for d in range(3)]
        num = sum(np.sum((Bnew[d] - Bold[d].reshape(-1)) ** 2) for d in range(3))
        den = sum(np.sum(Bold[d] ** 2) for d in range(3))
        enew = sum(np.sum(Bnew[d] ** 2) for d in range(3))
        brms = np.sqrt(enew / self.nc)
        return dict(cell_b_rel_l2_change=np.sqrt(num / den) if den > 0 else 0.0,
                    magnetic_energy_ratio=enew / den if den > 0 else 1.0,
                    max_divb_face_dx_over_brms=self.max_divb_face() * float(self.dx) / brms if brms > 0 else 0.0)


def _half_phase(n_axis: int, m: np.ndarray):
    """ct.jl:240-244: theta = k dx / 2 of 0-based fft index m, and the Nyquist flag."""
    nyq = (n_axis % 2 == 0) & (m == n_axis // 2)
    mm = np.where(m > n_axis / 2, m - n_axis, m)
    th = np.where(nyq, 0.0, np.pi * mm / n_axis)
    return th, nyq


def faces_from_cells_fft(B):
    """Arrays B[d] of shape (nz, ny, nx) -> divergence-free lower-face fields (ct.jl:271-297)."""
    nz, ny, nx = B[0].shape
    Bh = [np.fft.rfftn(b, axes=(0, 1, 2)) for b in B]  # last axis (x) is the half axis
    tx, nyx = _half_phase(nx, np.arange(nx // 2 + 1))
    ty, nyy = _half_phase(ny, np.arange(ny))
    tz, nyz = _half_phase(nz, np.arange(nz))
    TX = tx[None, None, :]
    TY = ty[None, :, None]
    TZ = tz[:, None, None]
    bx = np.where(nyx[None, None, :], 0, Bh[0])
    by = np.where(nyy[None, :, None], 0, Bh[1])
    bz = np.where(nyz[:, None, None], 0, Bh[2])
    tnx, tny, tnz = np.tan(TX), np.tan(TY), np.tan(TZ)
    t2 = tnx * tnx + tny * tny + tnz * tnz
    with np.errstate(invalid="ignore", divide="ignore"):
        dot = np.where(t2 > 0, (tnx * bx + tny * by + tnz * bz) / np.where(t2 > 0, t2, 1), 0)
    bx = bx - tnx * dot
    by = by - tny * dot
    bz = bz - tnz * dot
    Fx = bx * np.exp(-1j * TX) / np.cos(TX)
    Fy = by * np.exp(-1j * TY) / np.cos(TY)
    Fz = bz * np.exp(-1j * TZ) / np.cos(TZ)
    faces = [np.fft.irfftn(F, s=(nz, ny, nx), axes=(0, 1, 2)) for F in (Fx, Fy, Fz)]
    return faces, [bx, by, bz]


def riemann_batch(L, R, dir, ch, dtype=np.float64, gamma=5.0 / 3.0, smallr=1e-6, pfl=1e-7, llf_dmin=1e-4,
                  llf_pmin=0.0, riemann="hlld", ct=False):
    """One flux per face (glm_physics.jl:317-348). L, R: (n, 9) primitive states; dir 0,1,2."""
    L = np.ascontiguousarray(L, dtype=dtype)
    R = np.ascontiguousarray(R, dtype=dtype)
    F = np.zeros_like(L)
    getattr(lib(), f"orc_riemann_batch_{_suf(dtype)}")(_p(L), _p(R), _p(F), L.shape[0], int(dir), int(ct),
                                                       float(ch), float(gamma), float(smallr), float(pfl),
                                                       float(llf_dmin), float(llf_pmin), RIEMANN[riemann])
    return F


def predict_batch(M, dtdx, dtype=np.float64, recon="plm", gamma=5.0 / 3.0, smallr=1e-6, pfl=1e-7):
    """Hancock predictor (glm_physics.jl:225-231 + integrator_ref.jl:24) of n cells.

    M: (n, 7, 9) primitive means (self, -x, +x, -y, +y, -z, +z). Returns mh (n, 9), dL, dR (n, 3, 9).
    """
    M = np.ascontiguousarray(M, dtype=dtype)
    n = M.shape[0]
    MH = np.zeros((n, 9), dtype=dtype)
    DL = np.zeros((n, 3, 9), dtype=dtype)
    DR = np.zeros((n, 3, 9), dtype=dtype)
    getattr(lib(), f"orc_predict_batch_{_suf(dtype)}")(_p(M), _p(MH), _p(DL), _p(DR), n, RECON[recon],
                                                       float(dtdx), float(gamma), float(smallr), float(pfl))
    return MH, DL, DR


def drag_coefficients(q, dtype=np.float64):
    """(exp(-q), phi1(q), A, C, D) of the drag propagator (glm_physics.jl:63-103)."""
    out = np.zeros(5)
    fn = getattr(lib(), f"orc_drag_coeffs_{_suf(dtype)}")
    fn.argtypes = [ctypes.c_double, ctypes.c_void_p]
    fn(float(q), out.ctypes.data_as(ctypes.c_void_p))
    return out


def num_threads() -> int:
    return lib().orc_num_threads()
