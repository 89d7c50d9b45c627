"""The oracle against the golden vectors of tests/golden/ (a literal FP64
brute force of the reference's interpolation loops, written by
tests/golden/make_golden.py): quadratic B-spline value and gradient at both
staggerings, the MAC velocity and its Jacobian, and the tricubic Hermite (full,
garm and fd variants) with its gradient, at 60 points including grid nodes, corners and points
beyond the clamp."""
import os

import numpy as np
import pytest

from oracle import np_ref
from oracle.oracle import HK_FD, HK_FULL, HK_GARM_FD

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "interp_golden.npz")


@pytest.fixture(scope="module")
def g():
    with np.load(G, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("name,BL", [("own_x", (0.0, 0.5, 0.5)), ("cell", (0.5, 0.5, 0.5))])
def test_bspline_golden(orc64, g, name, BL):
    v, gr = orc64.eval_bspline(g["f"], BL, g["pts"], float(g["dx"]), grad=True)
    assert rel(v, g[f"bspline_{name}_val"]) < 1e-13
    assert rel(gr, g[f"bspline_{name}_grad"]) < 1e-13
    nv, ng = np_ref.bspline(g["f"], BL, g["pts"], float(g["dx"]), grad=True)
    assert rel(nv, g[f"bspline_{name}_val"]) < 1e-13
    assert rel(ng, g[f"bspline_{name}_grad"]) < 1e-13


def test_umac_golden(orc64, g):
    v, J = orc64.eval_umac([g["u0"], g["u1"], g["u2"]], g["pts"], float(g["dx"]), grad=True)
    assert rel(v, g["umac_val"]) < 1e-13
    assert rel(J, g["umac_jac"]) < 1e-13


def test_hermite_golden(orc64, g):
    v, Gr = orc64.eval_hermite(HK_FULL, g["psi"], g["T"], g["dT"], g["d2T"], g["pts"], float(g["dx"]), grad=True)
    assert rel(v, g["hermite_val"]) < 1e-13
    assert rel(Gr, g["hermite_grad"]) < 1e-13
    nv, ng = np_ref.hermite_full(g["psi"], g["T"], g["dT"], g["d2T"], g["pts"], float(g["dx"]), grad=True)
    assert rel(nv, g["hermite_val"]) < 1e-12
    assert rel(ng, g["hermite_grad"]) < 1e-12


@pytest.mark.parametrize("key,hk", [("hermite_garm", HK_GARM_FD), ("hermite_fd", HK_FD)])
def test_hermite_derived_variants_golden(orc64, g, key, hk):
    """The garm (d2T from differences of dT) and fd (dT, d2T from differences of
    T) variants the EDGE and ED4 marches use."""
    v, Gr = orc64.eval_hermite(hk, g["psi"], g["T"], g["dT"], g["d2T"], g["pts"], float(g["dx"]), grad=True)
    assert rel(v, g[f"{key}_val"]) < 1e-13
    assert rel(Gr, g[f"{key}_grad"]) < 1e-13
