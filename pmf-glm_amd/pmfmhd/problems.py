"""Initial conditions, computed on the host in float64 (numpy only).

Each function returns the conserved state as a ``(9, nz*ny*nx)`` float64 array in the
order (rho, mx, my, mz, E, Bx, By, Bz, psi) with i fastest (the layout of state.jl:1-13).
The caller uploads it (``MHDState.set_conserved``), cast to the state's precision, so
every consumer of one seed sees the same values.

Restated from lib/MHDKernels/src/problems.jl (test problems) and, for the random
solenoidal fields of the benchmark configurations, a counter-based splitmix64 hash in
the manner of pmf.jl:109-141 with a Box-Muller normal and a projection that makes the
field solenoidal under the centred-difference divergence of diagnostics.jl:6-21.
"""
from __future__ import annotations

import numpy as np

U64 = np.uint64


def prim2cons(q, gamma, out=None):
    """glm_physics.jl:19-23 on arrays: q is a sequence of 9 primitive arrays
    (rho, vx, vy, vz, p, Bx, By, Bz, psi); returns (9, n) float64 conserved."""
    rho, vx, vy, vz, p, bx, by, bz, psi = q
    n = np.size(rho)
    U = np.empty((9, n)) if out is None else out
    U[0] = rho
    U[1] = rho * vx
    U[2] = rho * vy
    U[3] = rho * vz
    U[4] = p / (gamma - 1.0) + 0.5 * rho * (vx * vx + vy * vy + vz * vz) + 0.5 * (bx * bx + by * by + bz * bz)
    U[5] = bx
    U[6] = by
    U[7] = bz
    U[8] = psi
    return U


def _grid(dims):
    nx, ny, nz = dims
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return i.reshape(-1), j.reshape(-1), k.reshape(-1)


def alfven_wave(dims, dx, gamma=5.0 / 3.0, amp=1e-3, B0=1.0, p0=0.1):
    """problems.jl:31-55: right-going circularly polarised Alfven wave along x (vA=1)."""
    nx = dims[0]
    i, _, _ = _grid(dims)
    x = (i + 0.5) * dx
    L = nx * dx
    ph = 2 * np.pi / L * x
    By = amp * np.sin(ph)
    Bz = amp * np.cos(ph)
    one = np.ones_like(x)
    z = np.zeros_like(x)
    return prim2cons(np.stack([one, z, -By, -Bz, p0 * one, B0 * one, By, Bz, z]), gamma)


def alfven_By_exact(x, t, L, amp=1e-3):
    """problems.jl:57-59."""
    return amp * np.sin(2 * np.pi / L * np.mod(x - t, L))


def orszag_tang(dims, dx, gamma=5.0 / 3.0):
    """problems.jl:61-80."""
    nx = dims[0]
    i, j, _ = _grid(dims)
    x = (i + 0.5) * dx
    y = (j + 0.5) * dx
    tp = 2 * np.pi / (nx * dx)
    B0 = 1.0 / np.sqrt(4 * np.pi)
    one = np.ones_like(x)
    z = np.zeros_like(x)
    q = np.stack([gamma * gamma * one, -np.sin(tp * y), np.sin(tp * x), z, gamma * one,
                  -B0 * np.sin(tp * y), B0 * np.sin(2 * tp * x), z, z])
    return prim2cons(q, gamma)


def brio_wu(dims, gamma=2.0):
    """problems.jl:82-98 (left state for 1-based i <= Nx/2)."""
    nx = dims[0]
    i, _, _ = _grid(dims)
    left = i < nx // 2
    one = np.ones(i.shape)
    z = np.zeros(i.shape)
    q = np.stack([np.where(left, 1.0, 0.125), z, z, z, np.where(left, 1.0, 0.1), 0.75 * one,
                  np.where(left, 1.0, -1.0), z, z])
    return prim2cons(q, gamma)


def turb_field(dims, dx, gamma=5.0 / 3.0):
    """problems.jl:9-29: smooth state varying along all three axes."""
    nx, ny, nz = dims
    i, j, k = _grid(dims)
    x = 2 * np.pi * (i + 0.5) * dx / (nx * dx)
    y = 2 * np.pi * (j + 0.5) * dx / (ny * dx)
    z = 2 * np.pi * (k + 0.5) * dx / (nz * dx)
    one = np.ones_like(x)
    q = np.stack([one, 0.2 * np.sin(z), 0.2 * np.sin(x), 0.2 * np.sin(y), one,
                  0.3 + 0.1 * np.sin(2 * y), 0.1 * np.sin(2 * z), 0.1 * np.sin(2 * x), np.zeros_like(x)])
    return prim2cons(q, gamma)


def field_loop(dims, dx, gamma=5.0 / 3.0, radius=0.3, amplitude=1e-3, rho0=1.0, p0=1.0, velocity=(1.0, 1.0, 0.0)):
    """problems.jl:100-132,160-170: centred curl of a cell-centred A_z cone."""
    nx, ny, nz = dims
    i, j, k = _grid(dims)
    x = (i + 0.5) * dx - 0.5 * nx * dx
    y = (j + 0.5) * dx - 0.5 * ny * dx
    A = (amplitude * np.maximum(radius - np.sqrt(x * x + y * y), 0.0)).reshape(nz, ny, nx)
    inv2dx = 1.0 / (2 * dx)
    bx = ((np.roll(A, -1, axis=1) - np.roll(A, 1, axis=1)) * inv2dx).reshape(-1)
    by = (-(np.roll(A, -1, axis=2) - np.roll(A, 1, axis=2)) * inv2dx).reshape(-1)
    one = np.ones_like(x)
    z = np.zeros_like(x)
    vx0, vy0, vz0 = velocity
    return prim2cons(np.stack([rho0 * one, vx0 * one, vy0 * one, vz0 * one, p0 * one, bx, by, z, z]), gamma)


def field_loop_Az_edge(dims, dx, radius=0.3, amplitude=1e-3):
    """problems.jl:134-146: A_z at the edges (i-1/2, j-1/2) for a CT state."""
    nx, ny, nz = dims
    i, j, _ = _grid(dims)
    x = i * dx - 0.5 * nx * dx
    y = j * dx - 0.5 * ny * dx
    return amplitude * np.maximum(radius - np.sqrt(x * x + y * y), 0.0)


# ---------------------------------------------------------------------------
# Counter-based random numbers (splitmix64 finaliser, pmf.jl:109-126 as the model).

_C1, _C2, _C3 = U64(0x9E3779B97F4A7C15), U64(0xBF58476D1CE4E5B9), U64(0x94D049BB133111EB)
_S30, _S27, _S31 = U64(30), U64(27), U64(31)


def _mix64(x):
    """splitmix64 finaliser on a uint64 array (wrapping arithmetic)."""
    with np.errstate(all="ignore"):
        z = np.asarray(x, dtype=U64) + _C1
        z ^= z >> _S30
        z *= _C2
        z ^= z >> _S27
        z *= _C3
        z ^= z >> _S31
    return z


def hash_words(seed, *keys):
    h = _mix64(np.asarray(np.int64(seed)).view(U64))
    for v in keys:
        h = _mix64(h ^ np.asarray(v, dtype=np.int64).view(U64))
    return h


def hash_uniform(seed, *keys):
    """Uniform in (0, 1): the top 53 bits of the hash word, centred in their bin."""
    w = hash_words(seed, *keys)
    return ((w >> U64(11)).astype(np.float64) + 0.5) * 2.0 ** -53


def hash_normal(seed, *keys):
    """Standard normal from two hashed uniforms (Box-Muller), keyed by (keys..., lane)."""
    u1 = hash_uniform(seed, *keys, 1)
    u2 = hash_uniform(seed, *keys, 2)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2 * np.pi * u2)


def solenoidal_field(dims, seed, stream, kmin, kmax, spectral_index, rms):
    """Gaussian random vector field, solenoidal under the centred divergence.

    Every integer mode m (|m| in [kmin, kmax]) gets complex Gaussian components keyed by
    (seed, stream, mx, my, mz, component); the per-mode amplitude is |m|^(n/2), i.e. the
    per-mode power P(k) ~ k^n. Each mode is projected orthogonally to the centred
    difference symbol s_d = sin(2 pi m_d / N_d), the field is built by a real inverse
    FFT and scaled to the requested rms |B|. Returns three (nz, ny, nx) float64 arrays.
    """
    nx, ny, nz = dims
    K = int(np.floor(kmax))
    # mode lists restricted to the cube |m_d| <= kmax, then to the shell [kmin, kmax]
    rz = np.arange(-min(K, nz // 2), min(K, (nz - 1) // 2) + 1, dtype=np.int64)
    ry = np.arange(-min(K, ny // 2), min(K, (ny - 1) // 2) + 1, dtype=np.int64)
    rx = np.arange(0, min(K, nx // 2) + 1, dtype=np.int64)
    MZ, MY, MX = np.meshgrid(rz, ry, rx, indexing="ij")
    kk = np.sqrt(MX.astype(np.float64) ** 2 + MY ** 2 + MZ ** 2)
    sel = (kk >= kmin) & (kk <= kmax)
    mx, my, mz, k = MX[sel], MY[sel], MZ[sel], kk[sel]
    out = []
    canon = (mx > 0) | ((mx == 0) & ((my > 0) | ((my == 0) & (mz > 0))))
    if nx % 2 == 0:
        canon |= (mx == nx // 2)
    sgn = np.where(canon | (mx > 0), 1, -1)
    kx_, ky_, kz_ = mx * sgn, my * sgn, mz * sgn
    amp = k ** (0.5 * spectral_index)
    coef = []
    for comp in range(3):
        re = hash_normal(seed, stream, kx_, ky_, kz_, comp, 0)
        im = hash_normal(seed, stream, kx_, ky_, kz_, comp, 1)
        c = (re + 1j * np.where(sgn > 0, im, -im)) * amp / np.sqrt(2.0)
        coef.append(c)
    sx = np.sin(2 * np.pi * mx / nx)
    sy = np.sin(2 * np.pi * my / ny)
    sz = np.sin(2 * np.pi * mz / nz)
    s2 = sx * sx + sy * sy + sz * sz
    with np.errstate(invalid="ignore", divide="ignore"):
        dot = np.where(s2 > 0, (sx * coef[0] + sy * coef[1] + sz * coef[2]) / np.where(s2 > 0, s2, 1.0), 0.0)
    coef = [coef[0] - sx * dot, coef[1] - sy * dot, coef[2] - sz * dot]
    iz = np.where(mz < 0, mz + nz, mz)
    iy = np.where(my < 0, my + ny, my)
    for comp in range(3):
        spec = np.zeros((nz, ny, nx // 2 + 1), dtype=np.complex128)
        spec[iz, iy, mx] = coef[comp]
        out.append(np.fft.irfftn(spec, s=(nz, ny, nx), axes=(0, 1, 2)))
        del spec
    cur = np.sqrt(np.mean(out[0] ** 2 + out[1] ** 2 + out[2] ** 2))
    scale = rms / cur if cur > 0 else 0.0
    return [o * scale for o in out]


def cell_normal(dims, seed, stream):
    """One standard normal per cell, keyed by (seed, stream, linear cell index)."""
    c = np.arange(int(np.prod(dims)), dtype=np.int64)
    return hash_normal(seed, stream, c)


def random_mhd(n, seed=1, kmax=None, brms=0.3, b_index=2.0, vrms_mach=0.0, v_index=-2.0,
               rho_mode="gauss", rho_sigma=0.01, gamma=5.0 / 3.0, p0=1.0):
    """The benchmark initial conditions (cubic N, unit box, dx = 1/N).

    rho_mode "gauss": rho = 1 + sigma*N(0,1) per cell; "lognormal": rho = exp(sigma*g - sigma^2/2).
    B: solenoidal Gaussian, P(k) ~ k^b_index for 1 <= |m| <= kmax, rms |B| = brms.
    v: solenoidal Gaussian, P(k) ~ k^v_index, same band, rms |v| = vrms_mach * sqrt(gamma p0 / 1).
    p = p0, psi = 0, E = p/(gamma-1) + rho v^2/2 + B^2/2.
    """
    dims = (n, n, n)
    if kmax is None:
        kmax = n // 4
    g = cell_normal(dims, seed, 7)
    if rho_mode == "gauss":
        rho = 1.0 + rho_sigma * g
    elif rho_mode == "lognormal":
        rho = np.exp(rho_sigma * g - 0.5 * rho_sigma ** 2)
    else:
        raise ValueError(f"unknown rho_mode {rho_mode}")
    B = solenoidal_field(dims, seed, 1, 1.0, kmax, b_index, brms)
    if vrms_mach > 0:
        cs = np.sqrt(gamma * p0)
        V = solenoidal_field(dims, seed, 2, 1.0, kmax, v_index, vrms_mach * cs)
    else:
        V = [np.zeros(n ** 3)] * 3
    p = np.full(n ** 3, p0)
    q = [rho, V[0].reshape(-1), V[1].reshape(-1), V[2].reshape(-1), p, B[0].reshape(-1), B[1].reshape(-1),
         B[2].reshape(-1), np.zeros(n ** 3)]
    return prim2cons(q, gamma)


# The benchmark configurations (unit box, gamma = 5/3, PLM minmod, HLLD, CFL 0.2).
CONFIGS = {
    "bstatic_32": dict(n=32, kmax=8, brms=0.3, b_index=2.0, vrms_mach=0.0, rho_mode="gauss", rho_sigma=0.01),
    "turb_256": dict(n=256, kmax=32, brms=1.0, b_index=2.0, vrms_mach=0.5, v_index=-2.0, rho_mode="gauss",
                   rho_sigma=0.01),
    "mach15_512": dict(n=512, kmax=64, brms=1.0, b_index=2.0, vrms_mach=1.5, v_index=-2.0, rho_mode="lognormal",
                   rho_sigma=0.3),
    "bstatic_256": dict(n=256, kmax=8, brms=0.3, b_index=2.0, vrms_mach=0.0, rho_mode="gauss", rho_sigma=0.01),
}


def config_state(name, seed=1, n=None):
    """Conserved state of a named benchmark configuration (optionally at another N)."""
    kw = dict(CONFIGS[name])
    if n is not None:
        scale = n / kw["n"]
        kw["kmax"] = max(1, int(round(kw["kmax"] * scale)))
        kw["n"] = n
    return random_mhd(seed=seed, **kw)



# The ensemble workload: 32 cubes of 32^3, 24 of 64^3 and 8 of 128^3; cube c has seed c,
# rms |B| uniform in [0.1, 2.0] and spectral index n_B uniform in [-2.5, 2.0] (both drawn
# from the splitmix stream of the cube id), cutoff |m| <= N/4, and otherwise the
# 32^3 benchmark recipe (rho = 1 + 0.01 N(0,1), p = 1, v = 0).
ENSEMBLE_LAYOUT = ((32, 32), (24, 64), (8, 128))


def ensemble_members(layout=ENSEMBLE_LAYOUT):
    out = []
    c = 0
    for count, n in layout:
        for _ in range(count):
            u = hash_uniform(c, 21, np.arange(2, dtype=np.int64))
            out.append(dict(cube=c, n=n, seed=c, brms=0.1 + 1.9 * float(u[0]), b_index=-2.5 + 4.5 * float(u[1])))
            c += 1
    return out


def ensemble_state(member):
    m = member
    return random_mhd(m["n"], seed=m["seed"], kmax=m["n"] // 4, brms=m["brms"], b_index=m["b_index"],
                      vrms_mach=0.0, rho_mode="gauss", rho_sigma=0.01)
