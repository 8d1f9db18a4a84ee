"""Python handle over the C ABI: a device-resident mesh + section + weak form.

Mirrors the objects PyLith's assembly layer drives (a PETSc DM with its section, PetscDS
weak form and auxiliary vectors; feassemble/IntegratorDomain.cc) with the device twins of
the pointwise kernels.  Vectors are torch tensors on the GPU (device memory plumbing only).
"""
import ctypes

import numpy as np
import torch

from . import _lib as L


class PlxError(RuntimeError):
    pass


def _i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(L._i32p)


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(L._dblp)


def _ptr(t):
    if t is None:
        return None
    if not isinstance(t, torch.Tensor) or not t.is_cuda or not t.is_contiguous():
        raise PlxError("vectors must be contiguous CUDA (HIP) tensors")
    return ctypes.c_void_p(t.data_ptr())


def _stype(t):
    if t.dtype == torch.float64:
        return L.FP64
    if t.dtype == torch.float32:
        return L.FP32
    raise PlxError(f"unsupported dtype {t.dtype}")


class DM:
    """Mesh + section + weak form on the device (one stream, one host thread)."""

    def __init__(self, mesh, labels=None, stream=None, vertex_perm=None):
        """vertex_perm: the caller's vertex order of the Q1 cells (plx_dm_create_ordered)."""
        self.lib = L.load()
        self.mesh = mesh
        h = ctypes.c_void_p()
        cl, clp = _i32(mesh.closure)
        co, cop = _i32(mesh.cone)
        xy, xyp = _f64(mesh.coords)
        lab = labels if labels is not None else mesh.labels
        if lab is not None:
            lb, lbp = _i32(lab)
        else:
            lbp = None
        if vertex_perm is None:
            err = self.lib.plx_dm_create(ctypes.byref(h), mesh.cell_type, mesh.num_cells,
                                         mesh.num_vertices, mesh.num_nodes, clp, cop, xyp, lbp)
        else:
            vp, vpp = _i32(vertex_perm)
            err = self.lib.plx_dm_create_ordered(ctypes.byref(h), mesh.cell_type, mesh.num_cells,
                                                 mesh.num_vertices, mesh.num_nodes, clp, cop, xyp, lbp, vpp)
        if err:
            raise PlxError(f"plx_dm_create failed with code {err}")
        self.h = h
        if stream is None:
            stream = torch.cuda.current_stream()
        self._check(self.lib.plx_dm_set_stream(self.h, ctypes.c_void_p(stream.cuda_stream)))

    def set_stream(self, stream):
        """Launch on `stream` (a torch.cuda.Stream) from now on, e.g. the capture stream of a
        HIP graph."""
        self._check(self.lib.plx_dm_set_stream(self.h, ctypes.c_void_p(stream.cuda_stream)))

    def _check(self, err):
        if err:
            msg = self.lib.plx_dm_last_error(self.h)
            raise PlxError(f"code {err}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.plx_dm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- section / constraints -----------------------------------------------------------
    def sizes(self):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.plx_dm_get_sizes(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_point_order(self, quad_perm=None):
        """The caller's quadrature-point order of per-point aux (plx_dm_set_point_order)."""
        qp = None if quad_perm is None else _i32(quad_perm)
        self._check(self.lib.plx_dm_set_point_order(self.h, 0 if qp is None else qp[0].size,
                                                    None if qp is None else qp[1]))

    def set_kernel_options(self, algorithm=L.ALGO_AUTO, accumulate=L.ACCUMULATE_AUTO, min_waves=0):
        self._check(self.lib.plx_dm_set_kernel_options(self.h, algorithm, accumulate, min_waves))

    def set_constraints(self, dofs):
        d, dp = _i32(dofs)
        self._check(self.lib.plx_dm_set_constraints(self.h, d.size, dp))

    # --- weak form ---------------------------------------------------------------------------
    def add_residual(self, label_value, part, f0, f1):
        self._check(self.lib.plx_dm_add_residual(self.h, label_value, part, f0, f1))

    def add_jacobian(self, label_value, part, jf0, jf1, jf2, jf3):
        self._check(self.lib.plx_dm_add_jacobian(self.h, label_value, part, jf0, jf1, jf2, jf3))

    def set_auxiliary(self, label_value, part, aux, aux_sizes, per_qp=False):
        a, ap = _f64(aux)
        s, sp = _i32(aux_sizes)
        self._check(self.lib.plx_dm_set_auxiliary(self.h, label_value, part, s.size, sp,
                                                  int(per_qp), ap))

    def set_constants(self, label_value, part, constants):
        c, cp = _f64(constants)
        self._check(self.lib.plx_dm_set_constants(self.h, label_value, part, c.size, cp))

    def update_state_vars(self, label_value, part, locX, t=0.0):
        if locX.dtype != torch.float64:
            raise PlxError("state variables are updated from a float64 solution")
        self._check(self.lib.plx_update_state_vars(self.h, label_value, part, t, _ptr(locX)))

    def get_auxiliary(self, label_value, part, like):
        """Current auxiliary field (with state variables) in the layout of `like` (host array)."""
        out = np.zeros_like(np.ascontiguousarray(like, np.float64))
        self._check(self.lib.plx_dm_get_auxiliary(self.h, label_value, part, out.ctypes.data_as(L._dblp)))
        return out

    # --- assembly ----------------------------------------------------------------------------
    def compute_residual(self, label_value, part, locX, locF, mode=L.ADD_VALUES, t=0.0,
                         locX_t=None):
        self._check(self.lib.plx_compute_residual(self.h, label_value, part, _stype(locX), t,
                                                  _ptr(locX), _ptr(locX_t), _ptr(locF), mode))

    def compute_residual_noalloc(self, label_value, part, locX, locF, mode=L.ADD_VALUES, t=0.0,
                                 locX_t=None):
        """plx_compute_residual_noalloc: the same residual with the library's static scratch
        array for the shared-node partial sums (allocates nothing)."""
        self._check(self.lib.plx_compute_residual_noalloc(self.h, label_value, part, _stype(locX), t,
                                                          _ptr(locX), _ptr(locX_t), _ptr(locF), mode))

    def compute_jacobian_action(self, label_value, part, locY, locF, mode=L.ADD_VALUES, t=0.0,
                                s_tshift=0.0, locX=None, locX_t=None):
        self._check(self.lib.plx_compute_jacobian_action(self.h, label_value, part, _stype(locY),
                                                         t, s_tshift, _ptr(locX), _ptr(locX_t),
                                                         _ptr(locY), _ptr(locF), mode))

    def compute_jacobian_pbdiag(self, label_value, part, blocks, t=0.0, s_tshift=0.0, locX=None):
        self._check(self.lib.plx_compute_jacobian_pbdiag(self.h, label_value, part, t, s_tshift,
                                                         _ptr(locX), _ptr(blocks)))

    # --- assembled Jacobian ------------------------------------------------------------------
    def jacobian_pattern(self):
        """(row_ptr [num_nodes + 1], col [nnzb]): 3x3 blocks over the mesh nodes (host arrays)."""
        nnzb = ctypes.c_int64()
        self._check(self.lib.plx_jacobian_pattern(self.h, ctypes.byref(nnzb), None, None))
        row_ptr = np.zeros(self.mesh.num_nodes + 1, np.int32)
        col = np.zeros(max(nnzb.value, 1), np.int32)
        self._check(self.lib.plx_jacobian_pattern(self.h, ctypes.byref(nnzb), row_ptr.ctypes.data_as(L._i32p),
                                                  col.ctypes.data_as(L._i32p)))
        return row_ptr, col[:nnzb.value]

    def compute_jacobian(self, label_value, part, values, mode=L.ADD_VALUES, t=0.0, s_tshift=0.0, locX=None,
                         locX_t=None):
        """values: float64 CUDA tensor [nnzb, 3, 3] (the pattern's blocks)."""
        if values.dtype != torch.float64:
            raise PlxError("the assembled Jacobian is float64")
        self._check(self.lib.plx_compute_jacobian(self.h, label_value, part, t, s_tshift, _ptr(locX), _ptr(locX_t),
                                                  _ptr(values), mode))

    def add_residual_pointfn(self, label_value, part, launcher):
        """Register a PetscPointFn-form residual launcher (an address from a hipcc-built library)."""
        self._check(self.lib.plx_dm_add_residual_pointfn(self.h, label_value, part, ctypes.c_void_p(launcher)))

    def add_jacobian_pointfn(self, label_value, part, action, assemble):
        self._check(self.lib.plx_dm_add_jacobian_pointfn(self.h, label_value, part, ctypes.c_void_p(action),
                                                         ctypes.c_void_p(assemble)))

    # --- fault (cohesive cells, prescribed slip) ----------------------------------------------
    def add_fault(self, fault):
        """Cohesive cells of `fault` (meshgen.Fault layout); the Lagrange-multiplier dofs follow
        the displacement dofs in the local vector."""
        arrs = [_i32(getattr(fault, k)) for k in ("closure_neg", "closure_pos", "closure_lag", "face_cone")]
        r1, r1p = _f64(fault.ref_dir_1)
        r2, r2p = _f64(fault.ref_dir_2)
        self._check(self.lib.plx_dm_add_fault(self.h, int(arrs[0][0].shape[0]), *[a[1] for a in arrs],
                                              int(fault.num_lag), r1p, r2p))

    def fault_info(self):
        off, n = ctypes.c_int64(), ctypes.c_int()
        self._check(self.lib.plx_dm_get_fault_info(self.h, ctypes.byref(off), ctypes.byref(n)))
        return off.value, n.value

    def set_fault_slip(self, slip):
        """Slip [num_lag, 3] (opening, left-lateral, reverse): a host array or a CUDA tensor."""
        if isinstance(slip, torch.Tensor) and slip.is_cuda:
            if slip.dtype != torch.float64:
                raise PlxError("slip must be float64")
            self._check(self.lib.plx_dm_set_fault_slip_device(self.h, _ptr(slip.contiguous())))
        else:
            s, sp = _f64(slip)
            self._check(self.lib.plx_dm_set_fault_slip(self.h, sp))

    def compute_residual_cohesive(self, locX, locF, mode=L.ADD_VALUES, t=0.0):
        if locX.dtype != torch.float64 or locF.dtype != torch.float64:
            raise PlxError("cohesive cells are computed in float64")
        self._check(self.lib.plx_compute_residual_cohesive(self.h, t, _ptr(locX), _ptr(locF), mode))

    def compute_jacobian_action_cohesive(self, locY, locF, mode=L.ADD_VALUES):
        if locY.dtype != torch.float64 or locF.dtype != torch.float64:
            raise PlxError("cohesive cells are computed in float64")
        self._check(self.lib.plx_compute_jacobian_action_cohesive(self.h, _ptr(locY), _ptr(locF), mode))

    # --- field scatters ------------------------------------------------------------------------
    def local_to_global(self, loc, glob, mode=L.ADD_VALUES):
        self._check(self.lib.plx_local_to_global(self.h, _stype(loc), _ptr(loc), _ptr(glob), mode))

    def global_to_local(self, glob, loc):
        self._check(self.lib.plx_global_to_local(self.h, _stype(loc), _ptr(glob), _ptr(loc)))

    def insert_boundary_values(self, values, loc):
        self._check(self.lib.plx_insert_boundary_values(self.h, _stype(loc), _ptr(values), _ptr(loc)))

    def dot(self, x, y):
        r = ctypes.c_double()
        self._check(self.lib.plx_vec_dot(self.h, _stype(x), x.numel(), _ptr(x), _ptr(y),
                                         ctypes.byref(r)))
        return r.value

    def norm2(self, x):
        r = ctypes.c_double()
        self._check(self.lib.plx_vec_norm2(self.h, _stype(x), x.numel(), _ptr(x), ctypes.byref(r)))
        return r.value

    # --- event timers ---------------------------------------------------------------------------
    def set_event_logging(self, on):
        self._check(self.lib.plx_dm_set_event_logging(self.h, int(on)))

    def event_stats(self, name):
        n, ms = ctypes.c_int(), ctypes.c_double()
        self._check(self.lib.plx_dm_get_event_stats(self.h, name.encode(), ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value

    def reset_event_stats(self):
        self._check(self.lib.plx_dm_reset_event_stats(self.h))


class Solver:
    """Matrix-free Krylov solver of one key's operator A = -J (plx_solver_* in the C ABI)."""

    def __init__(self, dm, label_value=0, part=0):
        self.dm = dm
        self.lib = dm.lib
        h = ctypes.c_void_p()
        dm._check(self.lib.plx_solver_create(ctypes.byref(h), dm.h, label_value, part))
        self.h = h

    def set_inner(self, rtol=1e-4, maxit=2000, restart=30):
        self.dm._check(self.lib.plx_solver_set_inner(self.h, rtol, maxit, restart))

    def set_preconditioner(self, kind):
        """'pbjacobi' (inverse 3x3 point blocks, the default) or 'amg' (smoothed-aggregation
        multigrid V-cycle on the assembled operator; Q1 keys covering every cell)."""
        code = {"pbjacobi": L.PC_PBJACOBI, "amg": L.PC_AMG}.get(kind, kind)
        self.dm._check(self.lib.plx_solver_set_preconditioner(self.h, int(code)))

    def setup(self):
        self.dm._check(self.lib.plx_solver_setup(self.h))

    def solve(self, x, b=None, method=L.KSP_FGMRES_IR, rtol=1e-10, maxit=10000):
        if x.dtype != torch.float64:
            raise PlxError("solutions are FP64 local vectors")
        st = L.SolveStats()
        self.dm._check(self.lib.plx_solver_solve(self.h, method, _ptr(b), _ptr(x), rtol, maxit,
                                                 ctypes.byref(st)))
        return st.as_dict()

    def close(self):
        if getattr(self, "h", None):
            self.lib.plx_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
