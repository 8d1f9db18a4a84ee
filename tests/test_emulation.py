"""CPU checks of the gfx950 kernels' indexing and element math: the host emulation of the tile
kernels (same plan, LDS staging, colour phases, halo reduction and element code, every index
bounds-checked) against the oracle.  No GPU."""
import ctypes
import os

import numpy as np
import pytest

from pylith_amd import _lib as L
from pylith_amd import meshgen as mg
from pylith_amd import meshio_ascii

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def emulate(m, aux, u, label=0, scalar=L.FP64, mode=L.INSERT_VALUES, r0=None, algo=L.ALGO_AUTO):
    lib = L.load()
    mu, K = aux[:, 1], aux[:, 2]
    lam_mu = np.ascontiguousarray(np.stack([K - 2 * mu / 3, mu], 1))
    dt = np.float64 if scalar == L.FP64 else np.float32
    uu = np.ascontiguousarray(u, dt)
    r = np.zeros_like(uu) if r0 is None else np.ascontiguousarray(r0, dt).copy()
    i32 = lambda a: np.ascontiguousarray(a, np.int32)
    cl, co, lb = i32(m.closure), i32(m.cone), i32(m.labels)
    xy = np.ascontiguousarray(m.coords, np.float64)
    err = lib.plx_debug_emulate_apply(m.cell_type, m.num_cells, m.num_vertices, m.num_nodes,
                                      cl.ctypes.data_as(L._i32p), co.ctypes.data_as(L._i32p),
                                      xy.ctypes.data_as(L._dblp), lb.ctypes.data_as(L._i32p), label,
                                      lam_mu.ctypes.data_as(L._dblp), scalar,
                                      ctypes.c_void_p(uu.ctypes.data), ctypes.c_void_p(r.ctypes.data), mode, algo)
    assert err == 0, f"emulation error {err}"
    return r


def relerr(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def cases():
    ref, _ = meshio_ascii.read(os.path.join(GOLDEN, "hex.mesh"))
    pert = mg.hex_box((9, 7, 5), (0, 0, 0), (900.0, 700.0, 500.0), ordered=False)
    rng = np.random.default_rng(11)
    x = pert.coords.copy()
    inner = np.all((x > 1) & (x < np.array([899, 699, 499])), axis=1)
    x[inner] += rng.uniform(-25, 25, (inner.sum(), 3))
    pert.coords = pert.node_coords = x
    return {
        "hex16": (mg.hex_box(16, (0, 0, 0), (6000.0, 6000.0, 6000.0)), 0),
        "hex_ragged": (mg.hex_box((13, 3, 2), (0, 0, 0), (13.0, 3.0, 2.0)), 0),
        "hex_unordered_nonaffine": (pert, 0),
        "hex_ref_fixture": (ref, 24),
        "tet8": (mg.tet_p2_box(8, (0, 0, 0), (6000.0, 6000.0, 6000.0)), 0),
        "tet_unordered": (mg.tet_p2_box((3, 4, 2), (0, 0, 0), (3.0, 4.0, 2.0), ordered=False), 0),
        # numbering without locality: the planner tiles these in Morton order
        "hex_renumbered": (mg.renumbered(mg.hex_box(12, (0, 0, 0), (1200.0, 1200.0, 1200.0)), 7), 0),
        "tet_renumbered": (mg.renumbered(mg.tet_p2_box(6, (0, 0, 0), (600.0, 600.0, 600.0)), 8), 0),
    }


@pytest.mark.parametrize("algo", [L.ALGO_QUADRATURE_POINTS, L.ALGO_FACTORED])
@pytest.mark.parametrize("name", list(cases()))
def test_emulated_kernels_match_oracle(oracle, name, algo):
    m, label = cases()[name]
    aux, sizes = mg.elastic_aux(m.num_cells, seed=5, vs_sigma=0.2, vp_vs=1.8)
    u = mg.random_field(m.num_nodes, 3)
    ref = oracle.Problem(m, aux, sizes).residual(u, 1)
    r = emulate(m, aux, u, label, algo=algo)
    assert relerr(r, ref) <= 1e-12
    r32 = emulate(m, aux, u, label, scalar=L.FP32, algo=algo)
    assert relerr(r32.astype(np.float64), ref) <= 2e-5
    radd = emulate(m, aux, u, label, mode=L.ADD_VALUES, r0=np.ones_like(u), algo=algo)
    assert relerr(radd - 1.0, ref) <= 1e-11
