"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module,
and only as the checker or the timed CPU baseline.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_i32p = ctypes.POINTER(ctypes.c_int32)
_dblp = ctypes.POINTER(ctypes.c_double)


class OracleMesh(ctypes.Structure):
    _fields_ = [("cell_type", ctypes.c_int), ("num_cells", ctypes.c_int),
                ("closure", _i32p), ("cone", _i32p), ("coords", _dblp),
                ("num_aux", ctypes.c_int), ("aux_sizes", _i32p), ("aux_per_qp", ctypes.c_int),
                ("aux", _dblp), ("num_constants", ctypes.c_int), ("constants", _dblp)]


class OracleCohesive(ctypes.Structure):
    _fields_ = [("num_cells", ctypes.c_int), ("closure_neg", _i32p), ("closure_pos", _i32p),
                ("closure_lag", _i32p), ("face_cone", _i32p), ("coords", _dblp),
                ("lag_offset", ctypes.c_int), ("slip", _dblp),
                ("ref_dir_1", ctypes.c_double * 3), ("ref_dir_2", ctypes.c_double * 3)]


def build():
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def plant_maxwell_error(kind):
    """Test hook: 1 drops tau/dt from the viscous-strain coefficient, 2 uses dev(e) in place of
    dev(e) - dev(e_t) in the update, 0 restores the reference update."""
    lib().oracle_plant_maxwell_error(ctypes.c_int(kind))


def _p(a, kind):
    return a.ctypes.data_as(_i32p if kind == "i" else _dblp)


class Problem:
    """One integrator key on the CPU: mesh arrays + aux + constants (arrays kept alive)."""

    def __init__(self, mesh, aux=None, aux_sizes=None, per_qp=False, constants=()):
        self.mesh = mesh
        self.closure = np.ascontiguousarray(mesh.closure, np.int32)
        self.cone = np.ascontiguousarray(mesh.cone, np.int32)
        self.coords = np.ascontiguousarray(mesh.coords, np.float64)
        self.aux = np.ascontiguousarray(aux, np.float64) if aux is not None else None
        self.aux_sizes = np.ascontiguousarray(aux_sizes, np.int32) if aux_sizes is not None else None
        self.constants = np.ascontiguousarray(constants, np.float64)
        self.s = OracleMesh()
        self.s.cell_type = mesh.cell_type
        self.s.num_cells = self.closure.shape[0]
        self.s.closure = _p(self.closure, "i")
        self.s.cone = _p(self.cone, "i")
        self.s.coords = _p(self.coords, "d")
        self.s.num_aux = 0 if self.aux_sizes is None else self.aux_sizes.size
        self.s.aux_sizes = _p(self.aux_sizes, "i") if self.aux_sizes is not None else None
        self.s.aux_per_qp = int(per_qp)
        self.s.aux = _p(self.aux, "d") if self.aux is not None else None
        self.s.num_constants = self.constants.size
        self.s.constants = _p(self.constants, "d") if self.constants.size else None

    def residual(self, locX, f1, f0=0, t=0.0, locX_t=None, cells=None, out=None):
        locX = np.ascontiguousarray(locX, np.float64)
        out = np.zeros_like(locX) if out is None else out
        xt = None if locX_t is None else np.ascontiguousarray(locX_t, np.float64)
        cp, n = (None, 0) if cells is None else (_p(np.ascontiguousarray(cells, np.int32), "i"), len(cells))
        err = lib().oracle_compute_residual(ctypes.byref(self.s), cp, n, f0, f1, ctypes.c_double(t),
                                            _p(locX, "d"), None if xt is None else _p(xt, "d"),
                                            _p(out, "d"))
        if err:
            raise RuntimeError(f"oracle residual error {err}")
        return out

    def residual_colored(self, locX, f1, colors, nthreads=0, out=None):
        locX = np.ascontiguousarray(locX, np.float64)
        out = np.zeros_like(locX) if out is None else out
        order = np.argsort(colors, kind="stable").astype(np.int32)
        ncol = int(colors.max()) + 1
        ptr = np.searchsorted(colors[order], np.arange(ncol + 1)).astype(np.int32)
        err = lib().oracle_compute_residual_colored(ctypes.byref(self.s), _p(ptr, "i"), _p(order, "i"),
                                                    ncol, f1, ctypes.c_double(0.0), _p(locX, "d"),
                                                    _p(out, "d"), nthreads)
        if err:
            raise RuntimeError(f"oracle residual error {err}")
        return out

    def jacobian_action(self, locY, jf3, locX=None, jf0=0, jf1=0, jf2=0, t=0.0, s_tshift=0.0,
                        out=None):
        locY = np.ascontiguousarray(locY, np.float64)
        locX = locY if locX is None else np.ascontiguousarray(locX, np.float64)
        out = np.zeros_like(locY) if out is None else out
        err = lib().oracle_compute_jacobian_action(ctypes.byref(self.s), None, 0, jf0, jf1, jf2, jf3,
                                                   ctypes.c_double(t), ctypes.c_double(s_tshift),
                                                   _p(locX, "d"), None, _p(locY, "d"), _p(out, "d"))
        if err:
            raise RuntimeError(f"oracle jacobian error {err}")
        return out

    def pbdiag(self, jf3, num_nodes, locX=None):
        blocks = np.zeros((num_nodes, 3, 3))
        x = np.zeros(3 * num_nodes) if locX is None else np.ascontiguousarray(locX, np.float64)
        err = lib().oracle_compute_jacobian_pbdiag(ctypes.byref(self.s), None, 0, jf3, ctypes.c_double(0),
                                                   ctypes.c_double(0), _p(x, "d"), _p(blocks, "d"))
        if err:
            raise RuntimeError(f"oracle pbdiag error {err}")
        return blocks

    def element_matrix(self, cell, jf3, locX=None):
        nb = self.closure.shape[1]
        Ke = np.zeros((3 * nb, 3 * nb))
        x = np.zeros(3 * (int(self.closure.max()) + 1)) if locX is None else np.ascontiguousarray(locX, np.float64)
        err = lib().oracle_element_matrix(ctypes.byref(self.s), cell, jf3, ctypes.c_double(0),
                                          ctypes.c_double(0), _p(x, "d"), _p(Ke, "d"))
        if err:
            raise RuntimeError(f"oracle element matrix error {err}")
        return Ke

    def update_state(self, locX, kernel_id, subfield, t=0.0):
        out = self.aux.copy()
        locX = np.ascontiguousarray(locX, np.float64)
        err = lib().oracle_update_state(ctypes.byref(self.s), kernel_id, subfield, ctypes.c_double(t),
                                        _p(locX, "d"), _p(out, "d"))
        if err:
            raise RuntimeError(f"oracle state update error {err}")
        return out


def quadrature(cell_type):
    nq = lib().oracle_num_quad(cell_type)
    p = np.zeros((nq, 3))
    w = np.zeros(nq)
    lib().oracle_quadrature(cell_type, _p(p, "d"), _p(w, "d"))
    return p, w


def tabulate(cell_type, xi):
    nb = lib().oracle_nodes_per_cell(cell_type)
    phi = np.zeros(nb)
    dphi = np.zeros((nb, 3))
    xi = np.ascontiguousarray(xi, np.float64)
    lib().oracle_tabulate(cell_type, _p(xi, "d"), _p(phi, "d"), _p(dphi, "d"))
    return phi, dphi


def point_f1(kernel_id, s_x, aux, aux_sizes, constants=()):
    """Call one PetscPointFn-signature kernel at a single point (f accumulates into zeros)."""
    fn = lib().oracle_point_kernel
    fn.restype = ctypes.c_void_p
    addr = fn(kernel_id)
    FN = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i32p, _dblp, _dblp,
                          _dblp, _i32p, _i32p, _dblp, _dblp, _dblp, ctypes.c_double, _dblp,
                          ctypes.c_int, _dblp, _dblp)
    f = FN(addr)
    aux = np.ascontiguousarray(aux, np.float64)
    sizes = np.asarray(aux_sizes)
    aOff = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int32)
    zero = np.zeros(1, np.int32)
    sx = np.ascontiguousarray(s_x, np.float64)
    s = np.zeros(3)
    x = np.zeros(3)
    c = np.ascontiguousarray(constants if len(constants) else [0.0], np.float64)
    out = np.zeros(64)
    f(3, 1, len(sizes), _p(zero, "i"), _p(zero, "i"), _p(s, "d"), None, _p(sx, "d"), _p(aOff, "i"),
      _p(aOff, "i"), _p(aux, "d"), None, None, 0.0, _p(x, "d"), len(constants), _p(c, "d"), _p(out, "d"))
    return out


def point_jf3(kernel_id, aux, aux_sizes, constants=()):
    fn = lib().oracle_point_jac_kernel
    fn.restype = ctypes.c_void_p
    addr = fn(kernel_id)
    FN = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i32p, _dblp, _dblp,
                          _dblp, _i32p, _i32p, _dblp, _dblp, _dblp, ctypes.c_double, ctypes.c_double,
                          _dblp, ctypes.c_int, _dblp, _dblp)
    f = FN(addr)
    aux = np.ascontiguousarray(aux, np.float64)
    sizes = np.asarray(aux_sizes)
    aOff = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int32)
    zero = np.zeros(1, np.int32)
    s = np.zeros(3)
    sx = np.zeros(9)
    x = np.zeros(3)
    c = np.ascontiguousarray(constants if len(constants) else [0.0], np.float64)
    out = np.zeros(81)
    f(3, 1, len(sizes), _p(zero, "i"), _p(zero, "i"), _p(s, "d"), None, _p(sx, "d"), _p(aOff, "i"),
      _p(aOff, "i"), _p(aux, "d"), None, None, 0.0, 0.0, _p(x, "d"), len(constants), _p(c, "d"),
      _p(out, "d"))
    return out


def local_to_global(loc, constrained):
    mask = np.ones(loc.size, bool)
    mask[np.asarray(constrained)] = False
    l2g = np.full(loc.size, -1, np.int32)
    l2g[mask] = np.arange(mask.sum(), dtype=np.int32)
    g = np.zeros(int(mask.sum()))
    lib().oracle_local_to_global(loc.size, _p(l2g, "i"), _p(np.ascontiguousarray(loc), "d"), _p(g, "d"))
    return g


def tangential_directions(r1, r2, n):
    t1, t2 = np.zeros(3), np.zeros(3)
    r1, r2, n = (np.ascontiguousarray(a, np.float64) for a in (r1, r2, n))
    lib().oracle_tangential_directions(_p(t1, "d"), _p(t2, "d"), _p(r1, "d"), _p(r2, "d"), _p(n, "d"))
    return t1, t2


class Cohesive:
    """Cohesive cells of one fault on the CPU (prescribed slip, FaultCohesiveKin kernels)."""

    def __init__(self, mesh, fault, slip, lag_offset):
        self.closure_neg = np.ascontiguousarray(fault.closure_neg, np.int32)
        self.closure_pos = np.ascontiguousarray(fault.closure_pos, np.int32)
        self.closure_lag = np.ascontiguousarray(fault.closure_lag, np.int32)
        self.face_cone = np.ascontiguousarray(fault.face_cone, np.int32)
        self.coords = np.ascontiguousarray(mesh.coords, np.float64)
        self.slip = np.ascontiguousarray(slip, np.float64)
        s = OracleCohesive()
        s.num_cells = self.closure_neg.shape[0]
        s.closure_neg = _p(self.closure_neg, "i")
        s.closure_pos = _p(self.closure_pos, "i")
        s.closure_lag = _p(self.closure_lag, "i")
        s.face_cone = _p(self.face_cone, "i")
        s.coords = _p(self.coords, "d")
        s.lag_offset = lag_offset
        s.slip = _p(self.slip, "d")
        for i in range(3):
            s.ref_dir_1[i] = fault.ref_dir_1[i]
            s.ref_dir_2[i] = fault.ref_dir_2[i]
        self.s = s

    def residual(self, locX, out=None, t=0.0):
        locX = np.ascontiguousarray(locX, np.float64)
        out = np.zeros_like(locX) if out is None else out
        err = lib().oracle_cohesive_residual(ctypes.byref(self.s), ctypes.c_double(t), _p(locX, "d"),
                                             _p(out, "d"))
        if err:
            raise RuntimeError(f"oracle cohesive residual error {err}")
        return out

    def jacobian_action(self, locY, out=None):
        locY = np.ascontiguousarray(locY, np.float64)
        out = np.zeros_like(locY) if out is None else out
        err = lib().oracle_cohesive_jacobian_action(ctypes.byref(self.s), _p(locY, "d"), _p(out, "d"))
        if err:
            raise RuntimeError(f"oracle cohesive jacobian error {err}")
        return out
