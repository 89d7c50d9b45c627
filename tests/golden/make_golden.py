"""Golden vectors for the interpolation primitives, by brute force.

A literal FP64 NumPy transcription of the reference's scalar loops, one point
at a time, on tiny grids: no vectorisation, no separable contraction, no
shared code with the oracle or the HIP kernels.  Run from the repository root
to regenerate:

    python tests/golden/make_golden.py

Writes tests/golden/interp_golden.npz (a few hundred KB).  Each output follows:
  * sample            util/taichi_utils.py:61-67   index truncated, clamped
  * N_2, dN_2         util/taichi_utils.py:229-250
  * interp_2 / interp_grad_2
                      util/taichi_utils.py:253-319 64 taps iu-1..iu+2
  * interp_u_MAC(_grad) util/taichi_utils.py:322-341
  * hermite_f/df/g/dg sim/hermite.py:4-41
  * interp_hermite / interp_grad_hermite
                      sim/hermite.py:459-607        8 corners, T*dx, dT*dx^2, d2T*dx^3
  * interp_(grad_)hermite_garm_fd_unroll
                      sim/hermite.py:298-456        d2T from central differences of dT
  * interp_(grad_)hermite_fd_unroll
                      sim/hermite.py:64-295         dT and d2T from central differences of T
with eps = 1e-6 (util/taichi_utils.py:3).
"""
import math
import os

import numpy as np

EPS = 1.e-6


def sample(f, u, v, w):
    n = f.shape
    i = max(0, min(int(u), n[0] - 1))
    j = max(0, min(int(v), n[1] - 1))
    k = max(0, min(int(w), n[2] - 1))
    return f[i, j, k]


def N_2(x):
    a = abs(x)
    if a < 0.5:
        return 3.0 / 4.0 - a ** 2
    if a < 1.5:
        return 0.5 * (3.0 / 2.0 - a) ** 2
    return 0.0


def dN_2(x):
    a = abs(x)
    r = 0.0
    if a < 0.5:
        r = -2 * a
    elif a < 1.5:
        r = 0.5 * (2 * a - 3)
    return -r if x < 0.0 else r


def _clamped(f, p, dx, BL):
    n = f.shape
    out = []
    for d in range(3):
        u = p[d] / dx - BL[d]
        out.append(max(1.0, min(u, n[d] - 2 - EPS)))
    return out


def interp_grad_2(f, p, dx, BL):
    s, t, l = _clamped(f, p, dx, BL)
    iu, iv, iw = math.floor(s), math.floor(t), math.floor(l)
    val = gx = gy = gz = 0.0
    for i in range(-1, 3):
        for j in range(-1, 3):
            for k in range(-1, 3):
                x, y, z = s - (iu + i), t - (iv + j), l - (iw + k)
                v = sample(f, iu + i, iv + j, iw + k)
                gx += 1. / dx * (v * dN_2(x) * N_2(y) * N_2(z))
                gy += 1. / dx * (v * N_2(x) * dN_2(y) * N_2(z))
                gz += 1. / dx * (v * N_2(x) * N_2(y) * dN_2(z))
                val += v * N_2(x) * N_2(y) * N_2(z)
    return val, (gx, gy, gz)


def hermite_f(x):
    a = abs(x)
    return 2 * a ** 3 - 3 * a ** 2 + 1 if a < 1.0 else 0.0


def hermite_df(x):
    a = abs(x)
    r = 6 * a ** 2 - 6 * a if a < 1.0 else 0.0
    return -r if x < 0 else r


def hermite_g(x):
    a = abs(x)
    r = a ** 3 - 2 * a ** 2 + a if a < 1.0 else 0.0
    return -r if x < 0 else r


def hermite_dg(x):
    a = abs(x)
    return 3 * a ** 2 - 4 * a + 1 if a < 1.0 else 0.0


def coef_full(T_c, dT_c, d2T_c, ii, jj, kk, dx):
    """sim/hermite.py:459-512: the stored T, dT, d2T of corner (ii, jj, kk)."""
    v1 = sample(T_c, ii, jj, kk) * dx
    v2 = sample(dT_c, ii, jj, kk) * dx * dx
    return v1, v2, sample(d2T_c, ii, jj, kk) * dx * dx * dx


def coef_garm(T_c, dT_c, d2T_c, ii, jj, kk, dx):
    """sim/hermite.py:298-355: T, dT stored; d2T from central differences of dT."""
    g = lambda a, b, c, m: sample(dT_c, a, b, c)[m]   # noqa: E731
    v3 = ((g(ii + 1, jj, kk, 0) - g(ii - 1, jj, kk, 0)) + (g(ii, jj + 1, kk, 1) - g(ii, jj - 1, kk, 1)) +
          (g(ii, jj, kk + 1, 2) - g(ii, jj, kk - 1, 2))) / 6 * dx * dx
    return sample(T_c, ii, jj, kk) * dx, sample(dT_c, ii, jj, kk) * dx * dx, v3


def coef_fd(T_c, dT_c, d2T_c, ii, jj, kk, dx):
    """sim/hermite.py:64-158: only T stored; dT and d2T from central differences of T."""
    g = lambda a, b, c, m: sample(T_c, a, b, c)[m]   # noqa: E731
    v2x = ((g(ii, jj, kk + 1, 1) - g(ii, jj, kk - 1, 1)) + (g(ii, jj + 1, kk, 2) - g(ii, jj - 1, kk, 2))) / 4 * dx
    v2y = ((g(ii, jj, kk + 1, 0) - g(ii, jj, kk - 1, 0)) + (g(ii + 1, jj, kk, 2) - g(ii - 1, jj, kk, 2))) / 4 * dx
    v2z = ((g(ii, jj + 1, kk, 0) - g(ii, jj - 1, kk, 0)) + (g(ii + 1, jj, kk, 1) - g(ii - 1, jj, kk, 1))) / 4 * dx
    v30 = ((g(ii + 1, jj, kk + 1, 1) - g(ii + 1, jj, kk - 1, 1)) + (g(ii + 1, jj + 1, kk, 2) - g(ii + 1, jj - 1, kk, 2))) / 4
    v31 = ((g(ii - 1, jj, kk + 1, 1) - g(ii - 1, jj, kk - 1, 1)) + (g(ii - 1, jj + 1, kk, 2) - g(ii - 1, jj - 1, kk, 2))) / 4
    v32 = ((g(ii, jj + 1, kk + 1, 0) - g(ii, jj + 1, kk - 1, 0)) + (g(ii + 1, jj + 1, kk, 2) - g(ii - 1, jj + 1, kk, 2))) / 4
    v33 = ((g(ii, jj - 1, kk + 1, 0) - g(ii, jj - 1, kk - 1, 0)) + (g(ii + 1, jj - 1, kk, 2) - g(ii - 1, jj - 1, kk, 2))) / 4
    v34 = ((g(ii, jj + 1, kk + 1, 0) - g(ii, jj - 1, kk + 1, 0)) + (g(ii + 1, jj, kk + 1, 1) - g(ii - 1, jj, kk + 1, 1))) / 4
    v35 = ((g(ii, jj + 1, kk - 1, 0) - g(ii, jj - 1, kk - 1, 0)) + (g(ii + 1, jj, kk - 1, 1) - g(ii - 1, jj, kk - 1, 1))) / 4
    v3 = (v30 - v31 + v32 - v33 + v34 - v35) / 6 * dx
    return sample(T_c, ii, jj, kk) * dx, np.array([v2x, v2y, v2z]), v3


def interp_grad_hermite(psi_c, T_c, dT_c, d2T_c, p, dx, coef=coef_full):
    """Component c: psi_c (n), T_c (n, 3), dT_c (n, 3), d2T_c (n); BL = 0.5.
    `coef` gives the corner's T*dx, dT*dx^2, d2T*dx^3 (stored or derived)."""
    s, t, l = _clamped(psi_c, p, dx, (0.5, 0.5, 0.5))
    iu, iv, iw = math.floor(s), math.floor(t), math.floor(l)
    val = px = py = pz = 0.0
    for i in range(2):
        for j in range(2):
            for k in range(2):
                x, y, z = s - (iu + i), t - (iv + j), l - (iw + k)
                hx = (hermite_f(x), hermite_g(x), hermite_df(x), hermite_dg(x))
                hy = (hermite_f(y), hermite_g(y), hermite_df(y), hermite_dg(y))
                hz = (hermite_f(z), hermite_g(z), hermite_df(z), hermite_dg(z))
                ii, jj, kk = iu + i, iv + j, iw + k
                # (coefficient, basis index per axis: 0 = f, 1 = g)
                terms = [(sample(psi_c, ii, jj, kk), (0, 0, 0))]
                v1, v2, v3 = coef(T_c, dT_c, d2T_c, ii, jj, kk, dx)
                terms += [(v1[0], (1, 0, 0)), (v1[1], (0, 1, 0)), (v1[2], (0, 0, 1))]
                terms += [(v2[0], (0, 1, 1)), (v2[1], (1, 0, 1)), (v2[2], (1, 1, 0))]
                terms += [(v3, (1, 1, 1))]
                for cf, (a, b, c) in terms:
                    val += cf * hx[a] * hy[b] * hz[c]
                    px += cf * hx[2 + a] * hy[b] * hz[c]
                    py += cf * hx[a] * hy[2 + b] * hz[c]
                    pz += cf * hx[a] * hy[b] * hz[2 + c]
    return val, (px / dx, py / dx, pz / dx)


def main():
    rng = np.random.default_rng(2024)
    n, dx = (6, 5, 7), 0.125
    hi = np.array(n) * dx
    pts = rng.uniform(-0.7 * dx, hi + 0.7 * dx, size=(60, 3))
    pts[:4] = [[0.0, 0.0, 0.0], hi, [dx, 2 * dx, 3 * dx], [2.5 * dx, 1.5 * dx, 0.5 * dx]]   # nodes, corners
    out = {"dx": np.float64(dx), "pts": pts}
    # one scalar field per B-spline offset
    f = rng.normal(size=(n[0] + 1, n[1], n[2]))
    out["f"] = f
    for name, BL in (("own_x", (0.0, 0.5, 0.5)), ("cell", (0.5, 0.5, 0.5))):
        vg = [interp_grad_2(f, p, dx, BL) for p in pts]
        out[f"bspline_{name}_val"] = np.array([v for v, _ in vg])
        out[f"bspline_{name}_grad"] = np.array([g for _, g in vg])
    # MAC velocity
    u = [rng.normal(size=(n[0] + 1, n[1], n[2])), rng.normal(size=(n[0], n[1] + 1, n[2])),
         rng.normal(size=(n[0], n[1], n[2] + 1))]
    for a in range(3):
        out[f"u{a}"] = u[a]
    BLs = [(0.0, 0.5, 0.5), (0.5, 0.0, 0.5), (0.5, 0.5, 0.0)]
    val = np.zeros((len(pts), 3))
    J = np.zeros((len(pts), 3, 3))
    for q, p in enumerate(pts):
        for a in range(3):
            v, g = interp_grad_2(u[a], p, dx, BLs[a])
            val[q, a] = v
            J[q, a] = g
    out["umac_val"], out["umac_jac"] = val, J
    # full Hermite map with arbitrary (inconsistent) derivative data
    psi = rng.normal(size=(3,) + n)
    T = rng.normal(size=(3, 3) + n)
    dT = rng.normal(size=(3, 3) + n)
    d2T = rng.normal(size=(3,) + n)
    out.update(psi=psi, T=T, dT=dT, d2T=d2T)
    for kind, coef in (("full", coef_full), ("garm", coef_garm), ("fd", coef_fd)):
        hv = np.zeros((len(pts), 3))
        hg = np.zeros((len(pts), 3, 3))
        for c in range(3):
            Tc = np.moveaxis(T[c], 0, -1)
            dTc = np.moveaxis(dT[c], 0, -1)
            for q, p in enumerate(pts):
                v, g = interp_grad_hermite(psi[c], Tc, dTc, d2T[c], p, dx, coef)
                hv[q, c] = v
                hg[q, c] = g
        key = "hermite" if kind == "full" else f"hermite_{kind}"
        out[f"{key}_val"], out[f"{key}_grad"] = hv, hg
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "interp_golden.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
