"""Synthetic faulted problems of the benchmark configurations (host set-up over the C ABI).

* ``faulted_box``: a Q1 box with a through-going fault on the vertex plane i = n/2, optionally
  sheared so the fault dips, isotropic elastic material (lognormal Vs per cell when seeded), both
  x faces fixed (Dirichlet on all components; the fault never touches them).
* ``patch_slip``: slip on a grid of fault patches, each fault vertex taking its patch's value --
  the patch basis of the Green's-function and serving configurations.
"""
import math

import numpy as np

from . import meshgen as mg

F1V_ISOLINEAR, JF3VU_ISOLINEAR = 1, 2


class FaultedProblem:
    def __init__(self, mesh, fault, aux, sizes, dofs, dm):
        self.mesh, self.fault, self.aux, self.sizes, self.dofs, self.dm = mesh, fault, aux, sizes, dofs, dm

    @property
    def local_size(self):
        return self.dm.sizes()[0]


def faulted_box(n, length=6000.0, dip_deg=None, seed=None, vs_sigma=None, vp_vs=1.732, brick=8, vs=3000.0):
    """n^3 Q1 cells in a cube of side `length` (z from -length to 0), fault plane x = length/2
    (i = n/2); `dip_deg` shears x by cot(dip) z so the fault dips at that angle."""
    from .dm import DM
    shear = None if dip_deg is None else 1.0 / math.tan(math.radians(dip_deg))
    m, f = mg.hex_box_with_fault(n, (0.0, 0.0, -length), (length, length, 0.0), n // 2, brick=brick, shear=shear)
    if vs_sigma is None:  # uniform: Vp / Vs of the 6 km benchmark cube (5292 / 3000 m/s)
        aux, sizes = mg.elastic_aux(m.num_cells, vs=vs, vp=vs * 5292.0 / 3000.0)
    else:
        aux, sizes = mg.elastic_aux(m.num_cells, vs=vs, seed=seed, vs_sigma=vs_sigma, vp_vs=vp_vs)
    x0 = m.node_coords[:, 0] - (shear or 0.0) * m.node_coords[:, 2]
    fixed = np.nonzero(np.isclose(x0, 0.0, atol=1e-9 * length) | np.isclose(x0, length, atol=1e-9 * length))[0]
    dofs = np.sort((3 * fixed[:, None] + np.arange(3)).ravel()).astype(np.int32)
    dm = DM(m)
    dm.add_residual(0, 0, 0, F1V_ISOLINEAR)
    dm.add_jacobian(0, 0, 0, 0, 0, JF3VU_ISOLINEAR)
    dm.set_auxiliary(0, 0, aux, sizes)
    dm.add_fault(f)
    dm.set_constraints(dofs)
    return FaultedProblem(m, f, aux, sizes, dofs, dm)


def patch_index(fault, n_strike, n_dip, j_range=None, k_range=None):
    """Patch id of each fault vertex on an n_strike x n_dip grid over the fault's (j, k) range
    (-1 outside the range)."""
    jk = fault.lag_jk
    j0, j1 = j_range if j_range else (0, jk[:, 0].max() + 1)
    k0, k1 = k_range if k_range else (0, jk[:, 1].max() + 1)
    inside = (jk[:, 0] >= j0) & (jk[:, 0] < j1) & (jk[:, 1] >= k0) & (jk[:, 1] < k1)
    ps = np.minimum(((jk[:, 0] - j0) * n_strike) // max(j1 - j0, 1), n_strike - 1)
    pd = np.minimum(((jk[:, 1] - k0) * n_dip) // max(k1 - k0, 1), n_dip - 1)
    idx = ps + n_strike * pd
    return np.where(inside, idx, -1)


def patch_slip(fault, patch_of_vertex, values, component=1):
    """Slip [..., num_lag, 3]: vertex v gets values[..., patch_of_vertex[v]] on `component`."""
    values = np.asarray(values, float)
    slip = np.zeros(values.shape[:-1] + (fault.num_lag, 3))
    inside = patch_of_vertex >= 0
    slip[..., inside, component] = values[..., patch_of_vertex[inside]]
    return slip
