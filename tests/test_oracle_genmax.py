"""Pins of the oracle's generalized Maxwell kernels and state update: the axial-strain relaxation
solution of tests/fullscale/viscoelasticity/nofaults-3d/axialstrain_genmaxwell_soln.py, the
elastic limit, the Jacobian as the exact linearization of f1, and the exponential decay of the
viscous strain at constant strain.  No GPU."""
import numpy as np

from pylith_amd import meshgen as mg

F1_ISO, F1_GM, JF3_GM, VISC, TOTAL = 1, 3, 4, 5, 6

# axialstrain_genmaxwell_soln.py material and loading
RHO, VS, VP = 2500.0, 3464.1016, 6000.0
ETA = (9.46728e17, 4.73364e17, 1.893456e18)
RATIO = (0.25, 0.25, 0.25)
U0 = 1.0
E0 = U0 / 8000.0
DT = 0.025 * mg.YEAR


def gm_update(P, u):
    """New state from the old aux: viscous strain and total strain projected at u."""
    new = P.update_state(u, VISC, 5)
    tot = P.update_state(u, TOTAL, 6)
    new[:, 27:33] = tot[:, 27:33]
    return new


def test_axial_strain_relaxation_matches_analytic(oracle):
    m = mg.hex_box(2, (-4000.0, -4000.0, -8000.0), (4000.0, 4000.0, 0.0), brick=2)
    eps0 = np.array([0, 0, E0, 0, 0, 0.0])
    dev0 = eps0 - np.array([1, 1, 1, 0, 0, 0]) * E0 / 3.0
    aux, sizes = mg.genmax_aux(m.num_cells, 8, RHO, VS, VP, ETA, RATIO, viscous_strain=dev0,
                               total_strain=eps0)
    u = np.zeros(3 * m.num_nodes)
    u[2::3] = E0 * (m.node_coords[:, 2] + 8000.0)
    mu, K = mg.moduli(RHO, VS, VP)
    lam = K - 2.0 / 3.0 * mu
    tau = np.array(ETA) / (mu * np.array(RATIO))
    r0 = 1.0 - sum(RATIO)
    top = np.isclose(m.node_coords[:, 2], 0.0)
    xpos = np.isclose(m.node_coords[:, 0], 4000.0)
    interior = np.all(np.abs(m.node_coords[:, :2]) < 3999.0, 1) & (m.node_coords[:, 2] < -1) & \
        (m.node_coords[:, 2] > -7999)
    area = 8000.0 * 8000.0
    for n in range(1, 21):
        t = n * DT
        P = oracle.Problem(m, aux, sizes, per_qp=True, constants=[DT])
        r = P.residual(u, F1_GM).reshape(-1, 3)
        fac = r0 + sum(RATIO[i] * np.exp(-t / tau[i]) for i in range(3))
        szz = 2.0 * mu * (2.0 * E0 / 3.0) * fac + E0 * (3.0 * lam + 2.0 * mu) / 3.0
        sxx = 2.0 * mu * (-E0 / 3.0) * fac + E0 * (3.0 * lam + 2.0 * mu) / 3.0
        assert np.isclose(r[top, 2].sum(), -szz * area, rtol=1e-11)
        assert np.isclose(r[xpos, 0].sum(), -sxx * area, rtol=1e-11)
        assert np.abs(r[interior]).max() <= 1e-11 * szz * area
        aux = gm_update(P, u)
        # constant strain: the viscous strains decay exactly as exp(-t / tau_i)
        for i in range(3):
            assert np.allclose(aux[:, 9 + 6 * i:15 + 6 * i], dev0 * np.exp(-t / tau[i]), rtol=1e-12,
                               atol=1e-20)
        assert np.allclose(aux[:, 27:33], eps0, rtol=1e-12, atol=1e-20)


def test_elastic_limit(oracle):
    # zero shear ratios and zero state: the elastic kernel
    rng = np.random.default_rng(3)
    aux, sizes = mg.genmax_aux(1, 1, RHO, VS, VP, ETA, (1e-300, 1e-300, 1e-300))
    aux[0, 6:9] = 0.0  # zero shear ratios (finite Maxwell times)
    G = rng.standard_normal(9) * 1e-4
    f_gm = oracle.point_f1(F1_GM, G, aux[0], sizes, [DT])[:9]
    f_el = oracle.point_f1(F1_ISO, G, aux[0, :3], np.array([1, 1, 1], np.int32))[:9]
    assert np.allclose(f_gm, f_el, rtol=1e-13, atol=1e-13 * np.abs(f_el).max())


def test_jacobian_is_the_linearization_of_f1(oracle):
    rng = np.random.default_rng(4)
    aux, sizes = mg.genmax_aux(1, 1, RHO, VS, VP, ETA, (0.3, 0.2, 0.1),
                               viscous_strain=rng.standard_normal(18) * 1e-4,
                               total_strain=rng.standard_normal(6) * 1e-4)
    J = oracle.point_jf3(JF3_GM, aux[0], sizes, [DT]).reshape(3, 3, 3, 3)  # [f][g][df][dg]
    G0 = rng.standard_normal(9) * 1e-4
    f0 = oracle.point_f1(F1_GM, G0, aux[0], sizes, [DT])[:9]
    for _ in range(3):
        d = rng.standard_normal((3, 3)) * 1e-4
        f = oracle.point_f1(F1_GM, G0 + d.ravel(), aux[0], sizes, [DT])[:9]
        lin = np.einsum("fgij,gj->fi", J, d).ravel()
        assert np.allclose(f - f0, lin, rtol=1e-9, atol=1e-12 * np.abs(f - f0).max())
    # shear factor: mu (r0 + sum r_i dq_i), dq_i = tau_i (1 - exp(-dt/tau_i)) / dt
    mu, K = aux[0, 1], aux[0, 2]
    tau, r = aux[0, 3:6], aux[0, 6:9]
    sf = mu * (1 - r.sum()) + np.sum(mu * r * tau * (1 - np.exp(-DT / tau)) / DT)
    assert np.isclose(-J[0, 0, 1, 1], sf, rtol=1e-14)                    # C_xyxy
    assert np.isclose(-J[0, 0, 0, 0], K + 4.0 * sf / 3.0, rtol=1e-14)    # C_xxxx
    assert np.isclose(-J[0, 1, 0, 1], K - 2.0 * sf / 3.0, rtol=1e-14)    # C_xxyy


def test_viscous_strain_update_closed_form(oracle):
    # single cell, linear displacement (uniform strain e) and arbitrary previous state:
    # ev_new = x ev + dq (dev e - dev et), et_new = e at every quadrature point
    m = mg.hex_box(1, (0, 0, 0), (2.0, 3.0, 1.5), brick=1)
    rng = np.random.default_rng(8)
    vs0, et0 = rng.standard_normal(18) * 1e-4, rng.standard_normal(6) * 1e-4
    aux, sizes = mg.genmax_aux(1, 8, RHO, VS, VP, (1e17, 5e17, 3e18), (0.3, 0.2, 0.1),
                               viscous_strain=vs0, total_strain=et0)
    Gu = rng.standard_normal((3, 3)) * 1e-4
    u = (m.node_coords @ Gu.T).ravel()
    P = oracle.Problem(m, aux, sizes, per_qp=True, constants=[DT])
    new = gm_update(P, u)
    e = np.array([Gu[0, 0], Gu[1, 1], Gu[2, 2], 0.5 * (Gu[0, 1] + Gu[1, 0]),
                  0.5 * (Gu[1, 2] + Gu[2, 1]), 0.5 * (Gu[0, 2] + Gu[2, 0])])

    def dev(v):
        return v - np.array([1, 1, 1, 0, 0, 0]) * v[:3].sum() / 3.0

    for i in range(3):
        tau = aux[0, 3 + i]
        x = np.exp(-DT / tau)
        dq = tau * (1 - x) / DT
        exp = x * vs0[6 * i:6 * i + 6] + dq * (dev(e) - dev(et0))
        assert np.allclose(new[:, 9 + 6 * i:15 + 6 * i], exp, rtol=1e-10, atol=1e-18)
    assert np.allclose(new[:, 27:33], e, rtol=1e-10, atol=1e-18)


# axialstrainrate_genmaxwell_soln.py: the same column and material, U = U0 + V0 t on the top
# face (V0 = 2 m/yr), 20 steps of dt = 0.025 yr from the state at t = 0
V0 = 2.0 / mg.YEAR


def strain_rate_analytic(t, mu, lam, tau):
    """The reference's closed-form solution at time t (axialstrainrate_genmaxwell_soln.py:57-139):
    total strain, the three viscous strains (eVis[:, :, i], lines 126-128) and sigma_xx, sigma_zz
    (lines 131-143)."""
    e0, edot0 = U0 / 8000.0, V0 / 8000.0
    ezz = e0 + edot0 * t
    eps = np.array([0.0, 0.0, ezz, 0.0, 0.0, 0.0])
    one = np.array([1, 1, 1, 0, 0, 0.0])
    edev0 = np.array([0, 0, e0, 0, 0, 0.0]) - one * e0 / 3.0
    edevdot = np.array([0, 0, edot0, 0, 0, 0.0]) - one * edot0 / 3.0
    evis = [edev0 * np.exp(-t / tau[i]) + edevdot * tau[i] * (1.0 - np.exp(-t / tau[i])) for i in range(3)]
    edev = eps - one * ezz / 3.0
    sdev = 2.0 * mu * ((1.0 - sum(RATIO)) * edev + sum(RATIO[i] * evis[i] for i in range(3)))
    smean = ezz / 3.0 * (3.0 * lam + 2.0 * mu)
    return eps, evis, sdev[0] + smean, sdev[2] + smean


def run_strain_rate(oracle, steps=20):
    """Step the oracle through the strain-rate problem at the exact displacement of each step;
    returns the largest relative deviations from the analytic viscous strains, the top-face and
    x-face reactions, and the interior residual."""
    m = mg.hex_box(2, (-4000.0, -4000.0, -8000.0), (4000.0, 4000.0, 0.0), brick=2)
    mu, K = mg.moduli(RHO, VS, VP)
    lam = K - 2.0 / 3.0 * mu
    tau = np.array(ETA) / (mu * np.array(RATIO))
    eps0, evis0, _, _ = strain_rate_analytic(0.0, mu, lam, tau)
    aux, sizes = mg.genmax_aux(m.num_cells, 8, RHO, VS, VP, ETA, RATIO,
                               viscous_strain=np.concatenate(evis0), total_strain=eps0)
    X = m.node_coords
    top, xpos = np.isclose(X[:, 2], 0.0), np.isclose(X[:, 0], 4000.0)
    interior = np.all(np.abs(X[:, :2]) < 3999.0, 1) & (X[:, 2] < -1) & (X[:, 2] > -7999)
    area = 8000.0 * 8000.0
    worst = {"visc": 0.0, "szz": 0.0, "sxx": 0.0, "interior": 0.0}
    for n in range(1, steps + 1):
        t = n * DT
        eps, evis, sxx, szz = strain_rate_analytic(t, mu, lam, tau)
        u = np.zeros(3 * m.num_nodes)
        u[2::3] = eps[2] * (X[:, 2] + 8000.0)
        P = oracle.Problem(m, aux, sizes, per_qp=True, constants=[DT])
        r = P.residual(u, F1_GM).reshape(-1, 3)
        worst["szz"] = max(worst["szz"], abs(r[top, 2].sum() / (-szz * area) - 1.0))
        worst["sxx"] = max(worst["sxx"], abs(r[xpos, 0].sum() / (-sxx * area) - 1.0))
        worst["interior"] = max(worst["interior"], np.abs(r[interior]).max() / (abs(szz) * area))
        aux = gm_update(P, u)
        scale = max(np.abs(v).max() for v in evis)
        for i in range(3):
            worst["visc"] = max(worst["visc"], np.abs(aux[:, 9 + 6 * i:15 + 6 * i] - evis[i]).max() / scale)
        assert np.abs(aux[:, 27:33] - eps).max() <= 1e-12 * eps[2]
    return worst


def test_axial_strain_rate_matches_analytic(oracle):
    """The viscous-strain increment dq (dev e - dev e_t) is nonzero at every step of the
    strain-rate problem, and with strain linear in time inside a step the discrete update of
    IsotropicLinearMaxwell.hh:220-223,256-278 reproduces the reference's closed form exactly."""
    worst = run_strain_rate(oracle)
    assert worst["visc"] <= 1e-11, worst
    assert worst["szz"] <= 1e-11 and worst["sxx"] <= 1e-11, worst
    assert worst["interior"] <= 1e-11, worst


def test_strain_rate_pin_catches_planted_update_errors(oracle):
    """A mistake planted in the oracle's viscous-strain update fails the strain-rate pin: the
    coefficient without tau/dt, or dev(e) in place of dev(e) - dev(e_t)."""
    try:
        for kind in (1, 2):
            oracle.plant_maxwell_error(kind)
            worst = run_strain_rate(oracle, steps=4)
            assert worst["visc"] > 1e-3 and worst["szz"] > 1e-6, (kind, worst)
    finally:
        oracle.plant_maxwell_error(0)
    assert run_strain_rate(oracle, steps=2)["visc"] <= 1e-11
