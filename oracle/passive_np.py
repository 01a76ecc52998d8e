"""Float64/float32 numpy restatement of the passive species transport — TEST
INFRASTRUCTURE ONLY.

lib/MHDKernels/src/passive_transport.jl: conserved species densities to bounded mass
fractions (5-25), and the semi-Lagrangian backtrace with the post-step cell velocity on
the periodic cubic grid, sampled trilinearly (27-54), for one fraction, two coupled
fractions, or an unconstrained scalar (78-190). Arrays are flat, i fastest.
Pinned in tests/test_passive.py by: a uniform composition preserved exactly, zero
velocity = identity, a one-cell displacement = a periodic shift, linear profiles moved
exactly by a fractional displacement, and the clamps.
"""
import numpy as np


def species_to_fractions(s1, rho, maximum_total=1.0, s2=None):
    T = s1.dtype.type
    d = np.maximum(rho, np.finfo(s1.dtype).eps)
    if s2 is None:
        return np.clip(s1 / d, T(0), T(maximum_total))
    q2 = np.clip(s2 / d, T(0), T(maximum_total))
    q1 = np.minimum(np.maximum(s1 / d, T(0)), T(maximum_total) - q2)
    return q1, q2


def _trilinear(a, n, i0, j0, k0, fx, fy, fz):
    def at(di, dj, dk):
        return a[((i0 + di) % n) + n * (((j0 + dj) % n) + n * ((k0 + dk) % n))]

    def lerp(l, r, f):
        return f * (r - l) + l

    x00 = lerp(at(0, 0, 0), at(1, 0, 0), fx)
    x10 = lerp(at(0, 1, 0), at(1, 1, 0), fx)
    x01 = lerp(at(0, 0, 1), at(1, 0, 1), fx)
    x11 = lerp(at(0, 1, 1), at(1, 1, 1), fx)
    return lerp(lerp(x00, x10, fy), lerp(x01, x11, fy), fz)


def backtrace(a1, U, n, dtdx, maximum_total=1.0, mode=0, a2=None):
    """mode 0: one fraction -> density * clamp; 1: two fractions (returns both); 2: scalar."""
    T = a1.dtype.type
    c = np.arange(n ** 3)
    i, j, k = c % n, (c // n) % n, c // (n * n)
    d = np.maximum(U[0], np.finfo(a1.dtype).eps)
    x = i.astype(a1.dtype) - T(dtdx) * U[1] / d
    y = j.astype(a1.dtype) - T(dtdx) * U[2] / d
    z = k.astype(a1.dtype) - T(dtdx) * U[3] / d
    i0, j0, k0 = np.floor(x).astype(np.int64), np.floor(y).astype(np.int64), np.floor(z).astype(np.int64)
    fx, fy, fz = x - i0.astype(a1.dtype), y - j0.astype(a1.dtype), z - k0.astype(a1.dtype)
    s1 = _trilinear(a1, n, i0, j0, k0, fx, fy, fz)
    if mode == 2:
        return s1
    mt = T(maximum_total)
    if mode == 1:
        s2 = _trilinear(a2, n, i0, j0, k0, fx, fy, fz)
        q2 = np.clip(s2, T(0), mt)
        q1 = np.minimum(np.maximum(s1, T(0)), mt - q2)
        return d * q1, d * q2
    return d * np.clip(s1, T(0), mt)
