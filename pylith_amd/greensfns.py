"""Batched Green's functions over the C ABI: many fault-slip impulses solved with one operator.

Mirrors pylith.problems.GreensFns (libsrc/pylith/problems/GreensFns.cc:236-331): the operator
is assembled once, then every impulse of the fault's impulse basis (FaultCohesiveImpulses,
faults/FaultCohesiveImpulses.cc:263-300) is solved.  Impulses are solved in blocks of k columns
(FP32 block PCG inside FP64 iterative refinement), k autotuned over {8, 16, 32, 64}; the PCG is
preconditioned by point-block Jacobi or by a smoothed-aggregation multigrid V-cycle.
"""
import ctypes

import numpy as np
import torch

from . import _lib as L
from .dm import PlxError, _ptr


class GreensFunctions:
    def __init__(self, dm, label_value=0, part=0):
        self.dm = dm
        self.lib = dm.lib
        h = ctypes.c_void_p()
        err = self.lib.plx_gf_create(ctypes.byref(h), dm.h, label_value, part)
        if err:
            raise PlxError(f"plx_gf_create failed ({err}): {self.lib.plx_dm_last_error(dm.h).decode()}")
        self.h = h

    def _check(self, err):
        if err:
            raise PlxError(f"error {err}: {self.lib.plx_dm_last_error(self.dm.h).decode()}")

    def set_options(self, block=0, inner_rtol=1e-4, inner_maxit=4000, use_graph=True):
        self._check(self.lib.plx_gf_set_options(self.h, block, inner_rtol, inner_maxit, int(use_graph)))

    def set_preconditioner(self, kind):
        """'jacobi' (point-block Jacobi, the default) or 'amg' (smoothed-aggregation multigrid,
        the GAMG role of materials/Elasticity.cc:307-318); builds the hierarchy on first use."""
        code = {"jacobi": L.GF_PC_JACOBI, "amg": L.GF_PC_AMG}.get(kind, kind)
        self._check(self.lib.plx_gf_set_preconditioner(self.h, int(code)))

    def autotune(self):
        best = ctypes.c_int()
        self._check(self.lib.plx_gf_autotune(self.h, ctypes.byref(best)))
        times = (ctypes.c_double * 4)()
        self._check(self.lib.plx_gf_probe_times(self.h, times))
        return best.value, {k: times[i] for i, k in enumerate((8, 16, 32, 64))}

    def apply(self, X, Y, k):
        """Y = A X for a [dof, k] block (float64 or float32 CUDA tensors)."""
        st = L.FP64 if X.dtype == torch.float64 else L.FP32
        self._check(self.lib.plx_gf_apply(self.h, k, st, _ptr(X), _ptr(Y)))

    def solve(self, slip, u, rtol=1e-10, maxit=20):
        """slip: [n, num_lag, 3] float64 CUDA tensor; u: [n, local size] float64 CUDA tensor."""
        if slip.dtype != torch.float64 or u.dtype != torch.float64:
            raise PlxError("slip and u must be float64")
        st = L.GfStats()
        self._check(self.lib.plx_gf_solve(self.h, slip.shape[0], _ptr(slip), _ptr(u), rtol, maxit,
                                          ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in L.GfStats._fields_}

    # --- continuous batching: k independent slots of one block ---------------------------------
    def slots_open(self, k, rtol, maxit=30):
        self._check(self.lib.plx_gf_slots_open(self.h, k, rtol, maxit))
        self.k = k

    def slots_load(self, slot_ids, slip):
        """slip: [n, num_lag, 3] float64 CUDA tensor for the empty slots `slot_ids`."""
        ids = np.ascontiguousarray(slot_ids, np.int32)
        if slip.dtype != torch.float64 or slip.shape[0] != ids.size:
            raise PlxError("slip must be float64 with one row per slot")
        self._check(self.lib.plx_gf_slots_load(self.h, ids.size, ids.ctypes.data_as(L._i32p), _ptr(slip)))

    def slots_step(self):
        """One step of every running slot; returns the slot states (L.SLOT_*)."""
        st = np.zeros(self.k, np.int32)
        self._check(self.lib.plx_gf_slots_step(self.h, st.ctypes.data_as(L._i32p)))
        return st

    def slot_take(self, slot, u=None):
        """Result of a finished slot into u ([local size] float64 CUDA tensor or None)."""
        stats = L.SlotStats()
        self._check(self.lib.plx_gf_slot_take(self.h, slot, _ptr(u), ctypes.byref(stats)))
        return {f: getattr(stats, f) for f, _ in L.SlotStats._fields_}

    def close(self):
        if getattr(self, "h", None):
            self.lib.plx_gf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
