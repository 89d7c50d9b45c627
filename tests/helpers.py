"""Input generators with closed forms for the parity and pinning tests."""
import numpy as np


def face_coords(a, nx, ny, nz, dx):
    """Coordinates of face grid a (util/taichi_utils.py:29-44), shape (3, d0, d1, d2)."""
    d = [nx + (a == 0), ny + (a == 1), nz + (a == 2)]
    i, j, k = np.meshgrid(np.arange(d[0]), np.arange(d[1]), np.arange(d[2]), indexing="ij")
    o = [0.0 if a == 0 else 0.5, 0.0 if a == 1 else 0.5, 0.0 if a == 2 else 0.5]
    return np.stack([(i + o[0]) * dx, (j + o[1]) * dx, (k + o[2]) * dx]).astype(np.float64)


def cell_coords(nx, ny, nz, dx):
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    return np.stack([(i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx]).astype(np.float64)


def mac_from_function(fn, nx, ny, nz, dx, dtype=np.float64):
    """MAC velocity sampling fn(X) -> (3, ...) at each face grid's own component."""
    out = []
    for a in range(3):
        X = face_coords(a, nx, ny, nz, dx)
        out.append(np.ascontiguousarray(fn(X)[a], dtype=dtype))
    return out


def linear_velocity(A, b, nx, ny, nz, dx, dtype=np.float64):
    A = np.asarray(A, np.float64)
    b = np.asarray(b, np.float64)

    def fn(X):
        return np.einsum("ab,b...->a...", A, X) + b.reshape(3, 1, 1, 1)
    return mac_from_function(fn, nx, ny, nz, dx, dtype)


def smooth_velocity(nx, ny, nz, dx, seed=0, amp=1.0, modes=4, dtype=np.float64):
    """A smooth random velocity (sum of a few Fourier modes), sampled on the MAC grid."""
    rng = np.random.default_rng(seed)
    L = np.array([nx, ny, nz]) * dx
    ks = rng.integers(1, 3, size=(modes, 3)) * (np.pi / L)
    ph = rng.uniform(0, 2 * np.pi, size=(modes, 3))
    co = rng.normal(size=(modes, 3)) * amp / modes

    def fn(X):
        out = np.zeros_like(X)
        for m in range(modes):
            arg = sum(ks[m, d] * X[d] for d in range(3))
            for a in range(3):
                out[a] += co[m, a] * np.sin(arg + ph[m, a])
        return out
    return mac_from_function(fn, nx, ny, nz, dx, dtype)


def taylor_P(M):
    """I + M + M^2/2 + M^3/6 + M^4/24: one RK4 step of the linear ODE x' = A x with M = A dt."""
    I = np.eye(3)
    M2 = M @ M
    M3 = M2 @ M
    return I + M + M2 / 2 + M3 / 6 + (M3 @ M) / 24


# ---------------------------------------------------------------------------
# Polynomial maps with exact Hermite data.  psi_c(x) = sum coef * x^a y^b z^c.
# ---------------------------------------------------------------------------
class PolyMap:
    """psi_c(x) = sum_t C[c, t] * x^e0 y^e1 z^e2 for exponent triples E[t]."""

    def __init__(self, E, C):
        self.E = np.asarray(E, int)
        self.C = np.asarray(C, float)

    def _d(self, X, dx_, dy_, dz_):
        out = np.zeros((3,) + X.shape[1:])
        for t, (e0, e1, e2) in enumerate(self.E):
            if e0 < dx_ or e1 < dy_ or e2 < dz_:
                continue
            f = 1.0
            for e, d in ((e0, dx_), (e1, dy_), (e2, dz_)):
                for q in range(d):
                    f *= (e - q)
            term = f * X[0] ** (e0 - dx_) * X[1] ** (e1 - dy_) * X[2] ** (e2 - dz_)
            for c in range(3):
                out[c] += self.C[c, t] * term
        return out

    def value(self, X):
        return self._d(X, 0, 0, 0)

    def grad_rows(self, X):
        """T[c][b] = d psi_c / d x_b, shape (3, 3, ...)."""
        return np.stack([self._d(X, 1, 0, 0), self._d(X, 0, 1, 0), self._d(X, 0, 0, 1)], axis=1)

    def cross(self, X):
        """dT[c] = [yz, xz, xy] second derivatives, shape (3, 3, ...)."""
        return np.stack([self._d(X, 0, 1, 1), self._d(X, 1, 0, 1), self._d(X, 1, 1, 0)], axis=1)

    def xyz(self, X):
        return self._d(X, 1, 1, 1)

    def tables(self, nx, ny, nz, dx, dtype=np.float64):
        X = cell_coords(nx, ny, nz, dx)
        return (self.value(X).astype(dtype), self.grad_rows(X).astype(dtype), self.cross(X).astype(dtype),
                self.xyz(X).astype(dtype))


def identity_plus(nx, ny, nz, dx, amp, seed=0):
    """A smooth perturbed identity map with exact data: psi = x + amp * poly(x)."""
    rng = np.random.default_rng(seed)
    E = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (0, 2, 0), (1, 1, 0), (0, 1, 1), (1, 0, 1), (1, 1, 1),
         (2, 1, 0), (0, 0, 2)]
    C = rng.normal(size=(3, len(E))) * amp
    C[:, :3] += np.eye(3)
    return PolyMap(E, C)


def interior_points(nx, ny, nz, dx, n, margin=2.5, seed=0):
    """Random points at least `margin` cells inside the box (no interpolation clamp active)."""
    rng = np.random.default_rng(seed)
    lo = margin * dx
    hi = np.array([nx, ny, nz]) * dx - margin * dx
    return rng.uniform(lo, hi, size=(n, 3))


def any_points(nx, ny, nz, dx, n, seed=0):
    """Random points over the box and slightly beyond (clamps active)."""
    rng = np.random.default_rng(seed)
    hi = np.array([nx, ny, nz]) * dx
    return rng.uniform(-0.7 * dx, hi + 0.7 * dx, size=(n, 3))


def rel_err(a, b, floor=1e-30):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), floor))
