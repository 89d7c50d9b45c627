"""Drop-in for the reference's sim/mgpcg_solid.py: MGPCG with solid cells.

MGPCG_Solid_3(boundary_types, boundary_mask, boundary_vel, N, base_level)
keeps the reference interface (sim/mgpcg_solid.py:313-432): solid cells
(boundary_mask > 0) are excluded from the operator, smoother, dots and
restriction; their faces take boundary_vel; the fluid divergence is corrected
one-sidedly when its sum is positive; the stop test is the absolute rTr < 1e-12
with at most 100 iterations (sim/mgpcg_solid.py:413-416).  The same keyword
extensions as sim/mgpcg.py's MGPCG_3 apply (mode, stop, check_every, x0).
"""
from edge_hip import _lib

from .mgpcg import _MGBase


class MGPCG_Solid_3(_MGBase):
    def __init__(self, boundary_types, boundary_mask, boundary_vel, N, base_level=3, real=float, **kw):
        kw.setdefault("abs_tol", 1e-12)
        super().__init__(boundary_types, N, base_level=base_level, **kw)
        self.boundary_mask = boundary_mask
        self.boundary_vel = boundary_vel
        _lib.call("edge_mg_set_solid", self._h, boundary_mask.ptr, boundary_vel.ptr)

    def Poisson(self, u_x, u_y, u_z, verbose=False, max_iters=100, x0=None):
        # the mask fields may have been re-allocated (swap_storage); re-bind them
        _lib.call("edge_mg_set_solid", self._h, self.boundary_mask.ptr, self.boundary_vel.ptr)
        self._bm = self.boundary_mask
        st = self._run_poisson(u_x, u_y, u_z, max_iters=max_iters, x0=x0, verbose=verbose)
        if verbose:
            print("[Poisson] Sum of fluid div: ", self.stats["div_sum"])
        return st

    @property
    def u_div_sum(self):
        return self.stats.get("div_sum", 0.0)
