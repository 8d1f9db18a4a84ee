"""The C++ IntegratorDomain mirror (include/pylith_amd_feassemble.hpp), driven by a C++ program
the way PyLith's problem layer drives its integrator, against the oracle (FP64, 1e-12)."""
import os
import subprocess

import numpy as np
import pytest

from pylith_amd import meshgen as mg

pytestmark = pytest.mark.gpu
F1, JF3 = 1, 2
DRIVER = os.path.join(os.path.dirname(__file__), "cpp", "integrator_driver")


def test_cpp_integrator_matches_oracle(oracle, tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail("tests/cpp/integrator_driver is not built (run __graft_entry__.build())")
    m = mg.hex_box((7, 5, 4), (0, 0, 0), (700.0, 500.0, 400.0), brick=4, shear=0.1)
    aux, sizes = mg.elastic_aux(m.num_cells, seed=2, vs_sigma=0.2, vp_vs=1.8)
    u = mg.random_field(m.num_nodes, 31)
    path_in, path_out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(path_in, "wb") as f:
        np.array([m.cell_type, m.num_cells, m.num_vertices, m.num_nodes, sizes.size], np.int32).tofile(f)
        np.ascontiguousarray(m.closure, np.int32).tofile(f)
        np.ascontiguousarray(m.coords, np.float64).tofile(f)
        np.ascontiguousarray(sizes, np.int32).tofile(f)
        np.ascontiguousarray(aux, np.float64).tofile(f)
        np.ascontiguousarray(u, np.float64).tofile(f)
    print("[cpp integrator] running driver", flush=True)
    res = subprocess.run([DRIVER, str(path_in), str(path_out)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    assert "refused" in res.stdout
    out = np.fromfile(path_out, np.float64)
    r, jv = out[:u.size], out[u.size:]
    P = oracle.Problem(m, aux, sizes)
    ref = P.residual(u, F1)
    assert np.abs(r - ref).max() <= 1e-12 * np.abs(ref).max()
    jref = P.jacobian_action(u, JF3)
    assert np.abs(jv - jref).max() <= 1e-12 * np.abs(jref).max()
