"""The oracle's Langevin OVRVO step with SHAKE/RATTLE, pinned by closed forms:

- OU half-steps with explicit noise on a force-free atom:
  v' = c1 v + c2 xi, c1 = exp(-gamma dt/2), c2 = sqrt((1 - c1^2) 418.4 kB T/m)
  (nhs_amber_fused.cu:1713-1744), drift x += dt v (:3780-3793);
- gamma = 0: the step is velocity Verlet, v1 = v0 + dt/2 * 418.4/m (F0 + F1)
  (kick :3667-3703, unit half impulse 0.2092 = 0.5 * 0.001 * 418.4,
  fused_engine.rs:4723-4731);
- SHAKE / RATTLE leave every X-H bond at its length and every relative
  velocity orthogonal to its bond (nhs_amber_fused.cu:1889-2360);
- the constrained NVE energy error is second order in dt.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from toy import make_topology

KB = 0.001987204
CONV = 418.4


def test_ou_and_drift_closed_form(pm, oracle, rng):
    m, T, gamma, dt = 12.0, 300.0, 5.0, 0.004
    t = make_topology(pm, [[0, 0, 0]], masses=[m])
    p = pm.default_params(dt_ps=dt, temperature_K=T, end_temperature_K=T, gamma_ps=gamma, equilibration_steps=0)
    x0 = np.zeros((1, 3))
    v0 = np.array([[1.0, -2.0, 0.5]])
    nz = rng.standard_normal((1, 2, 1, 3)).astype(np.float32)
    x1, v1, _ = oracle.run(t, p, x0, v0, 0, 1, nz)
    c1 = math.exp(-gamma * dt / 2)
    c2 = math.sqrt((1 - c1 * c1) * CONV * KB * T / m)
    va = c1 * v0 + c2 * nz[0, 0].astype(np.float64)
    assert np.allclose(x1, x0 + dt * va, rtol=0, atol=1e-12)
    assert np.allclose(v1, c1 * va + c2 * nz[0, 1].astype(np.float64), rtol=0, atol=1e-12)


def test_ou_stationary_variance(pm, oracle):
    # a force-free atom under the Philox-driven thermostat samples
    # v ~ N(0, 418.4 kB T / m) per axis; samples 50 steps (> 4/gamma) apart
    m, T = 12.0, 300.0
    t = make_topology(pm, [[0, 0, 0]], masses=[m])
    p = pm.default_params(temperature_K=T, end_temperature_K=T, gamma_ps=20.0, equilibration_steps=0,
                          com_remove_every=0)
    x, v = np.zeros((1, 3)), np.zeros((1, 3))
    samples = []
    for k in range(400):
        x, v, _ = oracle.run(t, p, x, v, 50 * k, 50)
        samples.append(v[0].copy())
    var = np.var(np.array(samples))
    assert var == pytest.approx(CONV * KB * T / m, rel=0.1)


def test_free_particle_nve(pm, oracle):
    t = make_topology(pm, [[0, 0, 0]], masses=[3.0])
    p = pm.default_params(gamma_ps=0.0, equilibration_steps=0)
    x, v, _ = oracle.run(t, p, np.zeros((1, 3)), np.array([[1.0, 2.0, 3.0]]), 0, 50)
    assert np.allclose(x, 50 * p.dt_ps * np.array([1.0, 2.0, 3.0]), atol=1e-12)
    assert np.allclose(v, [1.0, 2.0, 3.0], atol=0)


def test_velocity_verlet_units(pm, oracle):
    m = 10.0
    t = make_topology(pm, [[0, 0, 0], [1.8, 0, 0]], masses=[m, m], bonds=[(0, 1, 1.5, 300.0)],
                      exclusions=[[1], [0]])
    p = pm.default_params(dt_ps=0.001, gamma_ps=0.0, equilibration_steps=0)
    x0 = np.array([[0, 0, 0], [1.8, 0, 0]], np.float64)
    F0, _ = oracle.forces(t, x0, 31, False)
    x1, v1, _ = oracle.run(t, p, x0, np.zeros((2, 3)), 0, 1)
    F1, _ = oracle.forces(t, x1, 31, False)
    half = 0.5 * 0.001 * CONV
    assert half == pytest.approx(0.2092)
    assert np.allclose(v1, half / m * (F0 + F1), rtol=1e-12, atol=1e-12)
    assert np.allclose(x1, x0 + 0.001 * half / m * F0, rtol=1e-12, atol=1e-12)


def _clusters(top):
    c = top["cluster_central"]
    h = top["cluster_hydrogens"].reshape(-1, 3)
    d = top["cluster_bond_lengths"].reshape(-1, 3)
    n = top["cluster_n_hydrogens"]
    return c, h, d, n


def test_shake_and_rattle_satisfy_constraints(oracle, small_top, rng):
    x0 = small_top["positions"].reshape(-1, 3).astype(np.float64)
    c, h, d, nh = _clusters(small_top)
    dt = 0.004
    for k in range(len(c)):
        v = rng.normal(0, 8.0, x0.shape)
        xs = x0.copy()
        x = (x0 + dt * v).reshape(-1)
        vv = v.reshape(-1).copy()
        it = oracle.shake_cluster(small_top, k, x, xs.reshape(-1), vv, dt, 1e-10)
        assert 0 < it <= 100
        xr = x.reshape(-1, 3)
        for s in range(nh[k]):
            r = np.linalg.norm(xr[h[k, s]] - xr[c[k]])
            dk = float(d[k, s])
            assert abs(r * r - dk * dk) / (dk * dk) <= 1e-10
        it = oracle.rattle_cluster(small_top, k, x, vv, 1e-12)
        assert 0 < it <= 100
        vr = vv.reshape(-1, 3)
        for s in range(nh[k]):
            b = xr[h[k, s]] - xr[c[k]]
            assert abs(np.dot(b, vr[h[k, s]] - vr[c[k]])) <= 1e-10 * np.linalg.norm(b) * max(
                1.0, np.linalg.norm(vr[h[k, s]] - vr[c[k]]))


def test_shake_conserves_momentum(oracle, small_top, rng):
    x0 = small_top["positions"].reshape(-1, 3).astype(np.float64)
    m = small_top["masses"].astype(np.float64)
    c, h, d, nh = _clusters(small_top)
    k = int(np.argmax(nh))
    atoms = [c[k]] + [h[k, s] for s in range(nh[k])]
    v = rng.normal(0, 8.0, x0.shape)
    x = (x0 + 0.004 * v).reshape(-1)
    vv = v.reshape(-1).copy()
    oracle.shake_cluster(small_top, k, x, x0.reshape(-1).copy(), vv, 0.004, 1e-10)
    # the cluster stores float inverse masses: momentum holds to float rounding
    p0 = sum(m[a] * v[a] for a in atoms)
    p1 = sum(m[a] * vv.reshape(-1, 3)[a] for a in atoms)
    assert np.allclose(p0, p1, rtol=1e-6, atol=1e-6)


def _nve_error(pm, oracle, top, dt, t_ps):
    p = pm.default_params(dt_ps=dt, gamma_ps=0.0, equilibration_steps=0, com_remove_every=0)
    x = top["positions"].reshape(-1, 3).astype(np.float64)  # X-H bonds at their lengths
    v = pm.maxwell_boltzmann(top["masses"], 50.0, 7).astype(np.float64)
    for k in range(len(top["cluster_central"])):
        oracle.rattle_cluster(top, k, x.reshape(-1), v.reshape(-1), 1e-12)
    e0 = sum(oracle.energies(top, x, v))
    err = 0.0
    n = int(round(t_ps / dt))
    for _ in range(n // 40):
        x, v, _ = oracle.run(top, p, x, v, 0, 40)
        err = max(err, abs(sum(oracle.energies(top, x, v)) - e0))
    return err


def test_constrained_nve_second_order(pm, oracle, small_top):
    """Velocity Verlet + SHAKE/RATTLE: the energy error of a fixed trajectory
    segment shrinks ~4x per halving of dt (a dropped or mis-scaled term
    would leave a first-order error)."""
    e1 = _nve_error(pm, oracle, small_top, 0.0005, 0.2)
    e2 = _nve_error(pm, oracle, small_top, 0.00025, 0.2)
    assert e2 < 0.4 * e1
    assert e2 < 0.5


def test_work_ledger_closes_the_energy_balance(pm, oracle):
    """W books every non-Hamiltonian kinetic-energy change (OU half-steps,
    the RATTLE after OU_N and the final RATTLE, force and velocity clamps;
    nhs_amber_fused.cu:1713-1788, :2245-2270, :11148-11186 with the routing of
    fused_engine.rs:19273-19277).  The remaining moves -- kicks, drift, the
    pre-drift and pre-OU_NP1 RATTLE and SHAKE -- are the constrained
    Hamiltonian flow, so E(t) - E(0) - W is the integrator's shadow error:
    second order in dt, while W itself is hundreds of kcal/mol.  A dropped,
    extra or sign-flipped entry leaves a residual of the size of W."""
    top = pm.Topology.generate(192, seed=11)
    x0 = top["positions"].reshape(-1, 3).astype(np.float64)
    v0 = pm.maxwell_boltzmann(top["masses"], 300.0, 5).astype(np.float64)
    pe0, ke0 = oracle.energies(top, x0, v0)
    res = []
    for dt in (0.001, 0.0005):
        p = pm.default_params(dt_ps=dt, gamma_ps=20.0, equilibration_steps=0, com_remove_every=0)
        x, v, w = oracle.run(top, p, x0, v0, 0, int(round(0.2 / dt)))
        pe1, ke1 = oracle.energies(top, x, v)
        res.append((pe1 + ke1) - (pe0 + ke0) - w)
        assert abs(w) > 100.0
        print(f"dt {dt}: W {w:.3f}, dE - W {res[-1]:.4f} kcal/mol")
    assert abs(res[1]) < 1e-3 * abs(w)
    assert abs(res[0]) / abs(res[1]) > 3.0  # second order (4 in the limit)


def test_work_ledger_is_zero_without_bath_or_clamps(pm, oracle):
    """gamma = 0: exact kicks, no OU, no RATTLE-after-OU change (the state is
    already constrained), and no velocity clamp at thermal speeds: W == 0."""
    top = pm.Topology.generate(192, seed=11)
    p = pm.default_params(dt_ps=0.001, gamma_ps=0.0, equilibration_steps=0, com_remove_every=0)
    x = top["positions"].reshape(-1, 3).astype(np.float64)
    v = pm.maxwell_boltzmann(top["masses"], 300.0, 5).astype(np.float64)
    x, v, _ = oracle.run(top, p, x, v, 0, 1)  # project the velocities onto the constraints
    x, v, w = oracle.run(top, p, x, v, 1, 50)
    assert abs(w) < 1e-9
