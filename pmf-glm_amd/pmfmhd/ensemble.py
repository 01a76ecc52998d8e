"""Ensembles of independent periodic float32 GLM cubes (over the pmf_ensemble_* C ABI).

Every cube advances with its own CFL time step (evolve!'s rule, driver.jl:48-71); one
launch per stage covers all cubes and the step pair replays as a captured HIP graph.
The result for each cube is the one a separate state stepped with ``compute_dt`` and
``step(..., integrator="cube")`` would give.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _RECON, _RIEMANN, _code
from ._lib import check, lib, pmf_params


class Ensemble:
    def __init__(self, edges, recon="plm_minmod", riemann="hlld", gamma=5.0 / 3.0, smallr=1e-6, pfl=1e-7,
                 llf_dmin=1e-4, llf_pmin=0.0):
        self.edges = [int(n) for n in edges]
        p = pmf_params()
        check(lib().pmf_params_default(ctypes.byref(p), 16, 16, 16, 1.0 / 16))
        p.gamma, p.smallr, p.pfl = float(gamma), float(smallr), float(pfl)
        p.llf_dmin, p.llf_pmin = float(llf_dmin), float(llf_pmin)
        p.riemann = _code(_RIEMANN, "MHD riemann", riemann, ":hlld, :hll, :llf")
        p.recon = _code(_RECON, "recon", recon, ":plm, :plm_minmod, :ppm")
        p.precision = 0
        p.solver = 0
        e = np.asarray(self.edges, dtype=np.int32)
        h = ctypes.c_void_p()
        check(lib().pmf_ensemble_create(ctypes.byref(p), len(e), e.ctypes.data_as(ctypes.c_void_p), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().pmf_ensemble_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def ncubes(self):
        return len(self.edges)

    def upload(self, cube, U):
        """Conserved state (9, edge^3) of one cube (cast to float32)."""
        n = self.edges[cube] ** 3
        for v in range(9):
            a = np.ascontiguousarray(U[v], dtype=np.float32).reshape(-1)
            check(lib().pmf_ensemble_upload(self._h, cube, v, a.ctypes.data_as(ctypes.c_void_p), n))

    def download(self, cube):
        n = self.edges[cube] ** 3
        out = np.empty((9, n), dtype=np.float32)
        for v in range(9):
            check(lib().pmf_ensemble_download(self._h, cube, v, out[v].ctypes.data_as(ctypes.c_void_p), n))
        return out

    def run(self, nsteps, cfl=0.2, glm_ch_fac=1.0, glm_cr=0.18, graph=True):
        """Advance every cube nsteps steps; returns the GPU seconds of the steps."""
        sec = ctypes.c_double()
        check(lib().pmf_ensemble_run(self._h, int(nsteps), float(cfl), float(glm_ch_fac), float(glm_cr),
                                     1 if graph else 0, ctypes.byref(sec)))
        return sec.value

    def clock(self):
        """(t, last dt, smax behind the last dt), one entry per cube."""
        t, dt, sm = (np.empty(self.ncubes) for _ in range(3))
        P = ctypes.c_void_p
        check(lib().pmf_ensemble_clock(self._h, t.ctypes.data_as(P), dt.ctypes.data_as(P), sm.ctypes.data_as(P)))
        return t, dt, sm

    def tune(self, reps=3):
        ms = ctypes.c_double()
        check(lib().pmf_ensemble_tune(self._h, int(reps), ctypes.byref(ms)))
        return ms.value

    def config(self):
        zc = np.empty(self.ncubes, dtype=np.int32)
        nwg, nit = ctypes.c_int32(), ctypes.c_int32()
        check(lib().pmf_ensemble_config(self._h, zc.ctypes.data_as(ctypes.c_void_p), ctypes.byref(nwg),
                                        ctypes.byref(nit)))
        return zc, nwg.value, nit.value

    def set_config(self, zc=None, nwg=512):
        z = None if zc is None else np.ascontiguousarray(zc, dtype=np.int32)
        check(lib().pmf_ensemble_set_config(self._h, None if z is None else z.ctypes.data_as(ctypes.c_void_p),
                                            int(nwg)))
