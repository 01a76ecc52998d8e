"""CPU checks of the C ABI: the library loads, exports every entry point that
include/pmfmhd.h declares, fills the reference defaults and rejects misuse with the
reference's text before touching the GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

import pmfmhd
from pmfmhd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pmfmhd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pmf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_params_struct_matches_header_layout():
    # int32 x 10 (nx, ny, nz, riemann, recon, bc[3], solver, precision) then 6 doubles
    assert ctypes.sizeof(_lib.pmf_params) == 10 * 4 + 6 * 8
    assert _lib.pmf_params.dx.offset == 40


def test_params_default_is_reference_default():
    p = _lib.pmf_params()
    assert _lib.lib().pmf_params_default(ctypes.byref(p), 8, 9, 10, 0.125) == 0
    # state.jl:123-129 and glm_physics.jl codes
    assert (p.nx, p.ny, p.nz, p.dx) == (8, 9, 10, 0.125)
    assert abs(p.gamma - 5 / 3) < 1e-15 and p.smallr == 1e-6 and p.pfl == 1e-7
    assert p.llf_dmin == 1e-4 and p.llf_pmin == 0.0
    assert p.riemann == 0 and p.recon == 0 and list(p.bc) == [0, 0, 0] and p.solver == 0 and p.precision == 0


def _create(**kw):
    p = _lib.pmf_params()
    _lib.lib().pmf_params_default(ctypes.byref(p), 8, 8, 8, 0.125)
    for k, v in kw.items():
        if k == "bc":
            for d in range(3):
                p.bc[d] = v[d]
        else:
            setattr(p, k, v)
    h = ctypes.c_void_p()
    rc = _lib.lib().pmf_create(ctypes.byref(p), ctypes.byref(h))
    return rc, _lib.lib().pmf_last_error().decode()


def test_create_rejects_ct_with_non_periodic_bc():
    rc, msg = _create(solver=1, bc=(1, 0, 0))
    assert rc != 0 and "requires all-periodic boundaries" in msg


@pytest.mark.parametrize("kw,text", [(dict(riemann=7), "unknown MHD riemann"), (dict(recon=9), "unknown recon"),
                                     (dict(bc=(0, 5, 0)), "unknown BC"), (dict(nx=0), "must be positive"),
                                     (dict(precision=3), "precision")])
def test_create_rejects_bad_codes(kw, text):
    rc, msg = _create(**kw)
    assert rc != 0 and text in msg


def test_null_state_is_an_error_not_a_crash():
    d = ctypes.c_double()
    assert _lib.lib().pmf_max_wavespeed(None, ctypes.byref(d)) != 0
    assert "null pmf_state" in _lib.lib().pmf_last_error().decode()
    assert _lib.lib().pmf_destroy(None) == 0


def test_python_mirror_validates_like_the_reference():
    with pytest.raises(pmfmhd.MHDError, match="unknown mhd_solver :bogus"):
        pmfmhd.allocate_state((8, 8, 8), 1 / 8, mhd_solver="bogus")
    with pytest.raises(pmfmhd.MHDError, match="periodic"):
        pmfmhd.allocate_state((8, 8, 8), 1 / 8, mhd_solver="ct", bc="reflecting")
    with pytest.raises(pmfmhd.MHDError, match="unknown recon :weno"):
        pmfmhd.allocate_state((8, 8, 8), 1 / 8, recon="weno")
    with pytest.raises(pmfmhd.MHDError, match="unknown MHD riemann :roe"):
        pmfmhd.allocate_state((8, 8, 8), 1 / 8, riemann="roe")


def test_version_string():
    assert b"gfx950" in _lib.lib().pmf_version()


def _compile_capi_check(tmp_path):
    """tests/capi_check.c as strict C99 against include/pmfmhd.h, linked to the library."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    root = ROOT
    libdir = os.path.join(root, "pmf-glm_amd", "pmfmhd")
    exe = str(tmp_path / "capi_check")
    subprocess.run([cc, "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", "-I", os.path.join(root, "include"),
                    os.path.join(root, "tests", "capi_check.c"), "-L", libdir, "-lpmfmhd", f"-Wl,-rpath,{libdir}",
                    "-o", exe], check=True)
    return exe


def test_header_is_c99_and_a_c_program_gets_codes_and_messages(tmp_path):
    import subprocess
    exe = _compile_capi_check(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "ok (no GPU calls)" in r.stdout


@pytest.mark.gpu
def test_a_c_program_steps_a_uniform_state_exactly(tmp_path):
    import subprocess
    exe = _compile_capi_check(tmp_path)
    r = subprocess.run([exe, "gpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "uniform state unchanged" in r.stdout
