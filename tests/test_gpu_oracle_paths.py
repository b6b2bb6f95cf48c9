"""GPU paths compared with the FP64 oracle that the per-step parity tests
(explicit noise, fixed schedule) leave open:

- the work ledger W (prism_md_energies out[2]) against oracle_run's work,
  with the force and velocity clamps firing (nhs_amber_fused.cu:1713-1788,
  :2245-2270, :11148-11186);
- the production noise path: device Philox4x32-10 keyed by (seed, stream),
  counter (system-local atom, step, purpose), Box-Muller on the hardware
  transcendentals, against oracle_run(noise=None), which draws the same keyed
  stream in FP64 (pinned by the Random123 known answers) -- including a step
  counter above 2^32 and a nonzero stream;
- a step inside the director's temperature ramp and the friction ramp
  (protocol_director.cu:58-160);
- the initial Maxwell-Boltzmann velocities (fused_engine.rs:16066-16116):
  the purpose-3 Philox normals scaled by sqrt(418.4 kB T / m), COM-free;
- the reference's own constraint tolerances (SHAKE 1e-4, RATTLE 1e-5,
  nhs_amber_fused.cu:1889-2360);
- a restored checkpoint stepped against the oracle from the saved state.

Bars are written in each test; the per-step ones are those of the
explicit-noise parity test (|dx| <= 2e-5 A, v relative RMS <= 1e-4) unless
the tolerance of the compared solver sets a wider one.
"""
from __future__ import annotations

import struct

import numpy as np
import pytest
from stepcheck import compare_step

pytestmark = pytest.mark.gpu

KB, CONV = 0.001987204, 418.4


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / max(1e-300, (b ** 2).sum())))


@pytest.fixture(scope="module")
def chain(pm):
    top = pm.Topology.generate(2000, seed=7)
    with pm.FusedEngine(top, pm.default_params()) as eng:
        eng.minimize(1000, 1.0)
        x = eng.get_positions()
    return top.with_positions(x), x


def _compare_step(eng, oracle, top, params, xs, vs, step0, n, noise=None, dx_bar=2e-5, dv_bar=1e-4, tag=""):
    """n steps on the GPU and in the oracle from (xs, vs); stepcheck holds
    the atoms of pairs at the cutoff's rounding edge to their own bound."""
    eng.step(n, noise=noise)
    xg, vg = eng.get_positions(), eng.get_velocities()
    xo, vo, w = oracle.run(top.with_positions(xs), params, xs, vs, step0, n, noise)
    compare_step(top, params, xg, vg, xo, vo, dx_bar, dv_bar, f"{tag} step0 {step0} n {n}")
    return xg, vg, vo, w


def test_work_ledger_matches_oracle(pm, oracle, chain):
    """Three explicit-noise Langevin steps (2 fs, gamma = 10/ps) from a state
    where one constraint group is nudged toward its closest non-bonded heavy
    neighbour until the LJ wall beats the 1000 kcal/mol/A kick clamp, and ten
    heavy atoms move at 25x their thermal speed (the 20x velocity clamp
    fires).  Each step's GPU W increment equals the oracle's to within the
    kinetic-energy change that the measured GPU-oracle velocity differences
    can carry, 3 sum_i m_i |v_i| |dv_i| / 418.4 (a first-order bound), plus
    1e-6 of the increment: a missing, extra or sign-flipped ledger entry moves
    W by the size of that entry, far outside the bound."""
    top, x = chain
    n = top.n_atoms
    rng = np.random.default_rng(3)
    p = pm.default_params(dt_ps=0.002, equilibration_steps=0, com_remove_every=0)
    m = top["masses"].astype(np.float64)
    c = top["cluster_central"]
    hyd = top["cluster_hydrogens"].reshape(-1, 3)
    x64 = x.astype(np.float64)
    # the closest non-bonded pair of heavy atoms that are not 1-4 partners
    pairs = oracle.pairs(top, x, x, 1)
    heavy = np.zeros(n, bool)
    heavy[c] = True
    exc = {tuple(sorted(e)) for e in top["exception_atoms"].reshape(-1, 2).tolist()}
    cand = np.array([(i, j) for i, j in pairs if heavy[i] and heavy[j] and (i, j) not in exc])
    dist = np.linalg.norm(x64[cand[:, 0]] - x64[cand[:, 1]], axis=1)
    a, b = cand[np.argmin(dist)]
    ci = int(np.nonzero(c == a)[0][0])
    group = [a] + [int(h) for h in hyd[ci] if h >= 0]
    u = (x64[b] - x64[a]) / np.linalg.norm(x64[b] - x64[a])
    for shift in np.arange(0.05, 2.0, 0.05):
        xt = x64.copy()
        xt[group] += shift * u
        F, _ = oracle.forces(top.with_positions(xt), xt)
        if np.linalg.norm(F, axis=1).max() > 1500.0:
            break
    x = xt.astype(np.float32)
    print(f"[work] group of atom {a} nudged {shift:.2f} A toward atom {b}: max|F| "
          f"{np.linalg.norm(F, axis=1).max():.0f}", flush=True)
    v = pm.maxwell_boltzmann(top["masses"], 300.0, 9)
    fast = rng.choice(np.setdiff1d(c, group), 10, replace=False)
    v[fast] = 25.0 * np.sqrt(CONV * KB * 300.0 / m[fast])[:, None] * np.array([0.6, 0.0, 0.8])
    nz = rng.standard_normal((3, 2, n, 3)).astype(np.float32)
    with pm.FusedEngine(top.with_positions(x), p) as eng:
        eng.set_state(x=x, v=v)
        F, _ = eng.compute_forces()
        assert np.linalg.norm(F, axis=1).max() > 1000.0  # the kick clamp is exercised
        xs, vs = eng.get_positions(), eng.get_velocities()
        for k in range(3):
            w_prev = eng.energies()["work"]
            xg, vg, vo, wo = _compare_step(eng, oracle, top, p, xs, vs, k, 1, nz[k:k + 1], 1e-4, 1e-4, "work")
            dw = eng.energies()["work"] - w_prev
            bound = 3.0 * (m * np.linalg.norm(vg, axis=1) * np.linalg.norm(vg - vo, axis=1)).sum() / CONV
            print(f"[work] step {k}: GPU dW {dw:.6f}  oracle dW {wo:.6f}  bound {bound:.2e}", flush=True)
            assert abs(dw - wo) <= bound + 1e-6 * abs(wo), (k, dw, wo, bound)
            assert abs(wo) > 100.0 * bound  # the comparison resolves the ledger
            xs, vs = xg, vg


def test_production_philox_step_matches_oracle(pm, oracle, chain):
    """noise = NULL: the device draws its own normals.  Three steps from the
    GPU's state at step 40 against oracle_run(noise=None) with the same
    (seed, stream) key and step counters, at the explicit-noise bars: a
    slip in the counter packing, the purpose word or the atom offset would
    change every normal by O(1)."""
    top, x = chain
    p = pm.default_params(seed=0x1234ABCD9876, stream_id=7, equilibration_steps=0, com_remove_every=0)
    with pm.FusedEngine(top, p) as eng:
        eng.step(40)
        xs, vs = eng.get_positions(), eng.get_velocities()
        _compare_step(eng, oracle, top, p, xs, vs, 40, 3, None, tag="philox")


def test_philox_counter_above_2_32(pm, oracle, chain):
    """Step counters above 2^32 put bits 32..47 into the third counter word.
    The step counter is set through a checkpoint (its header carries the
    system state, whose first field is the step)."""
    top, x = chain
    p = pm.default_params(seed=99, stream_id=3, equilibration_steps=0, com_remove_every=0)
    big = (1 << 32) + (5 << 32) + 17
    with pm.FusedEngine(top, p) as eng:
        eng.step(2)
        blob = bytearray(eng.checkpoint())
        off = 8 + 4 + 4 + 4 + 4 + 8 + 8 + 8 + 8  # magic .. work: SysState follows
        (step,) = struct.unpack_from("<Q", blob, off)
        assert step == 2
        struct.pack_into("<Q", blob, off, big)
        eng.restore(bytes(blob))
        assert eng.stats()["steps_done"] == big
        xs, vs = eng.get_positions(), eng.get_velocities()
        _compare_step(eng, oracle, top, p, xs, vs, big, 2, None, tag="philox 2^32")


def test_temperature_and_friction_ramp_step(pm, oracle, chain):
    """Steps inside the director's heating ramp (250 -> 350 K between steps 2
    and 60) and the equilibration friction ramp (1000/ps decaying over 400
    steps) against the oracle's schedule, with the device's own noise."""
    top, x = chain
    p = pm.default_params(temperature_K=250.0, end_temperature_K=350.0, cold_hold_end=2, ramp_end=60,
                          warm_hold_end=80, ramp_down_end=120, equilibration_steps=400,
                          equilibration_gamma_ps=1000.0, com_remove_every=0)
    T, g, _ = oracle.schedule(p, 31)
    assert 250.0 < T < 350.0 and 10.0 < g < 1000.0
    with pm.FusedEngine(top, p) as eng:
        eng.step(30)
        xs, vs = eng.get_positions(), eng.get_velocities()
        _compare_step(eng, oracle, top, p, xs, vs, 30, 2, None, tag="ramp")
        # and in the cooling ramp
        eng.step(100 - 32)
        xs, vs = eng.get_positions(), eng.get_velocities()
        _compare_step(eng, oracle, top, p, xs, vs, 100, 2, None, tag="cooling")


def test_initial_velocities(pm, oracle):
    """The engine's initial velocities are the Maxwell-Boltzmann draw of the
    Philox stream (purpose 3, step 0): v = sqrt(418.4 kB T / m) z minus the
    centre-of-mass velocity, with z the oracle's keyed normals; the sample
    temperature is T within its statistical spread and the momentum is zero."""
    top = pm.Topology.generate(3000, seed=21)
    T = 310.0
    p = pm.default_params(seed=4242, stream_id=9, temperature_K=T, end_temperature_K=T)
    with pm.FusedEngine(top, p) as eng:
        v = eng.get_velocities().astype(np.float64)
    m = top["masses"].astype(np.float64)
    n = top.n_atoms
    z = np.array([oracle.langevin_normals(p.seed, p.stream_id, i, 0, 3) for i in range(n)])
    want = np.sqrt(CONV * KB * T / m)[:, None] * z
    want -= (m[:, None] * want).sum(0) / m.sum()
    assert np.abs(v - want).max() <= 1e-5 * np.abs(want).max()
    p_tot = np.abs((m[:, None] * v).sum(0)).max()
    assert p_tot <= 1e-3 * np.sqrt((m[:, None] * v * v).sum())
    t_sample = (m[:, None] * v * v).sum() / (3 * n * KB * CONV)
    assert abs(t_sample / T - 1.0) < 4.0 * np.sqrt(2.0 / (3 * n))


def test_reference_constraint_tolerances(pm, oracle, chain):
    """At the reference's tolerances (SHAKE |r^2 - d^2|/d^2 <= 1e-4, RATTLE
    1e-5) the GPU's exact/Newton solves and the oracle's Gauss-Seidel sweeps
    stop at different points inside the tolerance, so positions agree to
    ~tol * d (bar 2e-4 A) and velocities to the SHAKE correction / dt (bar
    3e-3 relative RMS); every X-H bond meets the reference's tolerance."""
    top, x = chain
    n = top.n_atoms
    p = pm.default_params(shake_tol=1e-4, rattle_tol=1e-5, equilibration_steps=0, com_remove_every=0)
    nz = np.random.default_rng(5).standard_normal((2, 2, n, 3)).astype(np.float32)
    with pm.FusedEngine(top, p) as eng:
        xs, vs = eng.get_positions(), eng.get_velocities()
        for k in range(2):
            xs, vs, _, _ = _compare_step(eng, oracle, top, p, xs, vs, k, 1, nz[k:k + 1], 2e-4, 3e-3, "ref tol")
        c = top["cluster_central"]
        h = top["cluster_hydrogens"].reshape(-1, 3)
        d = top["cluster_bond_lengths"].reshape(-1, 3)
        for s in range(3):
            ok = h[:, s] >= 0
            if not ok.any():
                continue
            r2 = ((xs[h[ok, s]].astype(np.float64) - xs[c[ok]]) ** 2).sum(1)
            assert (np.abs(r2 - d[ok, s].astype(np.float64) ** 2) / d[ok, s] ** 2).max() <= 1.5e-4


def test_restored_checkpoint_steps_like_oracle(pm, oracle, chain):
    """A checkpoint saved after 150 steps and restored into a fresh engine:
    one explicit-noise step from the restored state matches the oracle's FP64
    step from the saved positions and velocities, and W continues from its
    saved value."""
    top, x = chain
    n = top.n_atoms
    p = pm.default_params(seed=77, stream_id=2)
    with pm.FusedEngine(top, p) as a:
        a.step(150)
        blob = a.checkpoint()
        xs, vs, w_saved = a.get_positions(), a.get_velocities(), a.energies()["work"]
    nz = np.random.default_rng(8).standard_normal((1, 2, n, 3)).astype(np.float32)
    with pm.FusedEngine(top, p) as b:
        b.restore(blob)
        assert np.array_equal(b.get_positions(), xs) and np.array_equal(b.get_velocities(), vs)
        assert b.energies()["work"] == pytest.approx(w_saved, rel=1e-12, abs=1e-9)
        _compare_step(b, oracle, top, p, xs, vs, 150, 1, nz, tag="restored")
