"""Allocation-free residual entry point (plx_compute_residual_noalloc) through the C ABI.

Its results must equal plx_compute_residual's (the same kernels; only the shared-node partial
sums live in the library's static scratch array): bitwise with coloured accumulation, to the
parity bar with LDS atomics (FP64 1e-12, FP32 2e-5 relative), and FP64 matches the oracle at
1e-12.  Two calls in a row share the scratch, the second on a smaller mesh; a key whose partial
sums exceed the scratch is rejected with an error instead of running.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from pylith_amd import _lib as L  # noqa: E402
from pylith_amd import meshgen as mg  # noqa: E402

pytestmark = pytest.mark.gpu
F1, JF3 = 1, 2


def relerr(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def make_dm(m, seed):
    from pylith_amd.dm import DM
    aux, sizes = mg.elastic_aux(m.num_cells, seed=seed, vs_sigma=0.2, vp_vs=1.732)
    dm = DM(m)
    dm.add_residual(0, 0, 0, F1)
    dm.add_jacobian(0, 0, 0, 0, 0, JF3)
    dm.set_auxiliary(0, 0, aux, sizes)
    return dm, aux, sizes


def test_noalloc_matches_residual_on_two_calls(oracle):
    # first call on the larger mesh, second on a smaller one (the scratch keeps the first
    # call's partial sums, which the second must not see)
    meshes = [mg.tet_p2_box(12, (0, 0, 0), (6000.0, 6000.0, 6000.0), brick=4),
              mg.hex_box(6, (0, 0, 0), (6000.0, 6000.0, 6000.0), brick=4)]
    for i, m in enumerate(meshes):
        print(f"[noalloc] call {i}: cell_type={m.cell_type} cells={m.num_cells}", flush=True)
        dm, aux, sizes = make_dm(m, seed=11 + i)
        u = mg.random_field(m.num_nodes, 20 + i)
        ref = oracle.Problem(m, aux, sizes).residual(u, F1)
        for acc in (L.ACCUMULATE_COLORED, L.ACCUMULATE_LDS_ATOMIC):
            dm.set_kernel_options(L.ALGO_AUTO, acc, 0)
            for dt, tol in ((torch.float64, 1e-12), (torch.float32, 2e-5)):
                x = torch.from_numpy(u).to("cuda", dt)
                for mode in (L.INSERT_VALUES, L.ADD_VALUES):
                    base = torch.full_like(x, 0.5)
                    r0, r1 = base.clone(), base.clone()
                    dm.compute_residual(0, 0, x, r0, mode=mode)
                    dm.compute_residual_noalloc(0, 0, x, r1, mode=mode)
                    torch.cuda.synchronize()
                    a, b = r1.double().cpu().numpy(), r0.double().cpu().numpy()
                    if acc == L.ACCUMULATE_COLORED:
                        assert torch.equal(r0, r1), (i, acc, dt, mode)
                    else:
                        assert relerr(a - 0.5 * (mode == L.ADD_VALUES), b - 0.5 * (mode == L.ADD_VALUES)) <= tol
                    got = a - (0.5 if mode == L.ADD_VALUES else 0.0)
                    assert relerr(got, ref) <= tol, (i, acc, dt, mode)
        dm.close()


def test_noalloc_rejects_key_beyond_scratch():
    # 176^3 Q1 in 8^3 tiles: about 4.0M shared-node slots = 95 MB of FP64 partial sums, beyond
    # the 64 MiB scratch; FP32 (48 MB) fits
    from pylith_amd.dm import PlxError
    m = mg.hex_box(176, (0, 0, 0), (6000.0, 6000.0, 6000.0), brick=8)
    print(f"[noalloc] oversized key: cells={m.num_cells}", flush=True)
    dm, _, _ = make_dm(m, seed=3)
    x = torch.zeros(3 * m.num_nodes, dtype=torch.float64, device="cuda")
    r = torch.zeros_like(x)
    with pytest.raises(PlxError, match="exceed"):
        dm.compute_residual_noalloc(0, 0, x, r, mode=L.INSERT_VALUES)
    x32, r32 = x.float(), r.float()
    dm.compute_residual_noalloc(0, 0, x32, r32, mode=L.INSERT_VALUES)
    torch.cuda.synchronize()
    assert float(r32.abs().max()) == 0.0
    dm.close()
