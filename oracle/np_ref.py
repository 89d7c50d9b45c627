"""Independent NumPy FP64 restatement of the interpolation primitives.

TEST INFRASTRUCTURE ONLY.  Written from the mathematics, not from the C++
oracle, so that the two cross-check each other:
  * quadratic B-spline by the closed-form 3-tap weights (the standard MPM
    form) instead of the reference's piecewise N_2 over 64 taps
    (util/taichi_utils.py:230-319);
  * tricubic Hermite by the unit-interval basis h00, h10, h01, h11 instead of
    the reference's |x|-based hermite_f/g (sim/hermite.py:4-41).
Clamping follows the reference's sample() and interpolation clamps
(util/taichi_utils.py:61-67, 262-264).
"""
import numpy as np


def _clampf(u, dim):
    return np.clip(u, 1.0, dim - 2 - 1e-6)


def bspline(f, BL, pts, dx, grad=False):
    """Value (and gradient) of the quadratic B-spline of f at pts (n, 3)."""
    f = np.asarray(f, np.float64)
    pts = np.asarray(pts, np.float64)
    idx, w, dw = [], [], []
    for d in range(3):
        s = _clampf(pts[:, d] / dx - BL[d], f.shape[d])
        base = np.floor(s - 0.5).astype(int)          # first of the 3 taps
        fx = s - base                                   # in [0.5, 1.5)
        ww = np.stack([0.5 * (1.5 - fx) ** 2, 0.75 - (fx - 1) ** 2, 0.5 * (fx - 0.5) ** 2], 1)
        dd = np.stack([fx - 1.5, -2 * (fx - 1), fx - 0.5], 1)
        ii = np.clip(base[:, None] + np.arange(3)[None], 0, f.shape[d] - 1)
        idx.append(ii)
        w.append(ww)
        dw.append(dd)
    val = np.zeros(len(pts))
    g = np.zeros((len(pts), 3))
    for a in range(3):
        for b in range(3):
            for c in range(3):
                v = f[idx[0][:, a], idx[1][:, b], idx[2][:, c]]
                val += v * w[0][:, a] * w[1][:, b] * w[2][:, c]
                g[:, 0] += v * dw[0][:, a] * w[1][:, b] * w[2][:, c]
                g[:, 1] += v * w[0][:, a] * dw[1][:, b] * w[2][:, c]
                g[:, 2] += v * w[0][:, a] * w[1][:, b] * dw[2][:, c]
    return (val, g / dx) if grad else val


def umac(u, pts, dx, grad=False):
    BLs = [(0, .5, .5), (.5, 0, .5), (.5, .5, 0)]
    outs = [bspline(u[a], BLs[a], pts, dx, grad=True) for a in range(3)]
    val = np.stack([o[0] for o in outs], 1)
    J = np.stack([o[1] for o in outs], 1)
    return (val, J) if grad else val


def _h(t):
    """Cubic Hermite basis on [0,1]: h00, h10, h01, h11 and derivatives."""
    h00 = 2 * t ** 3 - 3 * t ** 2 + 1
    h10 = t ** 3 - 2 * t ** 2 + t
    h01 = -2 * t ** 3 + 3 * t ** 2
    h11 = t ** 3 - t ** 2
    d00 = 6 * t ** 2 - 6 * t
    d10 = 3 * t ** 2 - 4 * t + 1
    d01 = -6 * t ** 2 + 6 * t
    d11 = 3 * t ** 2 - 2 * t
    return (h00, h10, h01, h11), (d00, d10, d01, d11)


def hermite_full(psi, T, dT, d2T, pts, dx, grad=False):
    """Tricubic Hermite of each component c from corner data psi_c, T_c, dT_c, d2T_c."""
    pts = np.asarray(pts, np.float64)
    n = len(pts)
    shp = psi.shape[1:]
    base, t = [], []
    for d in range(3):
        s = _clampf(pts[:, d] / dx - 0.5, shp[d])
        b = np.floor(s).astype(int)
        base.append(b)
        t.append(s - b)
    H = [_h(t[d]) for d in range(3)]
    vals = np.zeros((n, 3))
    G = np.zeros((n, 3, 3))
    for c in range(3):
        for i in range(2):
            for j in range(2):
                for k in range(2):
                    I = (np.clip(base[0] + i, 0, shp[0] - 1), np.clip(base[1] + j, 0, shp[1] - 1),
                         np.clip(base[2] + k, 0, shp[2] - 1))
                    # (value basis index, derivative basis index) per corner side
                    sel = [(0, 1) if i == 0 else (2, 3), (0, 1) if j == 0 else (2, 3), (0, 1) if k == 0 else (2, 3)]
                    coef = {
                        (0, 0, 0): psi[c][I],
                        (1, 0, 0): T[c][0][I] * dx, (0, 1, 0): T[c][1][I] * dx, (0, 0, 1): T[c][2][I] * dx,
                        (0, 1, 1): dT[c][0][I] * dx ** 2, (1, 0, 1): dT[c][1][I] * dx ** 2,
                        (1, 1, 0): dT[c][2][I] * dx ** 2, (1, 1, 1): d2T[c][I] * dx ** 3,
                    }
                    for (ex, ey, ez), cf in coef.items():
                        bx = H[0][0][sel[0][ex]]
                        by = H[1][0][sel[1][ey]]
                        bz = H[2][0][sel[2][ez]]
                        vals[:, c] += cf * bx * by * bz
                        G[:, c, 0] += cf * H[0][1][sel[0][ex]] * by * bz
                        G[:, c, 1] += cf * bx * H[1][1][sel[1][ey]] * bz
                        G[:, c, 2] += cf * bx * by * H[2][1][sel[2][ez]]
    return (vals, G / dx) if grad else vals


# ---------------------------------------------------------------------------
# Multigrid V-cycle of sim/mgpcg.py:177-195, vectorised (all walls, no solids)
# ---------------------------------------------------------------------------
def _nsum_diag(z):
    n = z.shape
    ns = np.zeros_like(z)
    dg = np.full(n, 6.0)
    for a in range(3):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[a], hi[a] = slice(0, -1), slice(1, None)
        ns[tuple(lo)] += z[tuple(hi)]
        ns[tuple(hi)] += z[tuple(lo)]
        first = [slice(None)] * 3
        last = [slice(None)] * 3
        first[a], last[a] = 0, -1
        dg[tuple(first)] -= 1
        dg[tuple(last)] -= 1
    return ns, dg


def _colour(n, phase):
    i, j, k = np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij")
    return ((i + j + k) & 1) == phase


def _smooth(r, z, phase):
    ns, dg = _nsum_diag(z)
    m = _colour(z.shape, phase)
    z[m] = ((r + ns) / dg)[m]


def vcycle(r0, levels, pre_post=2, bottom=20):
    r = [np.asarray(r0, np.float64)]
    z = [np.zeros_like(r[0])]
    for l in range(levels - 1):
        for _ in range(pre_post):
            _smooth(r[l], z[l], 0)
            _smooth(r[l], z[l], 1)
        ns, dg = _nsum_diag(z[l])
        res = r[l] - (dg * z[l] - ns)
        n = res.shape
        rc = res.reshape(n[0] // 2, 2, n[1] // 2, 2, n[2] // 2, 2).sum(axis=(1, 3, 5)) / 2.0
        r.append(rc)
        z.append(np.zeros_like(rc))
    for _ in range(bottom):
        _smooth(r[-1], z[-1], 0)
        _smooth(r[-1], z[-1], 1)
    for l in reversed(range(levels - 1)):
        z[l] += np.repeat(np.repeat(np.repeat(z[l + 1], 2, 0), 2, 1), 2, 2)
        for _ in range(pre_post):
            _smooth(r[l], z[l], 1)
            _smooth(r[l], z[l], 0)
    return z[0]
