"""One MD step of the GPU against the FP64 oracle from the same state.

The nonbonded force is discontinuous at the cutoff: the potential is shifted,
the force is not (nhs_amber_fused.cu:631-717).  After the drift the GPU holds
FP32 positions and the oracle FP64 ones, so a pair whose r^2 lies within
rounding of rc^2 can be inside the cutoff on one side and outside on the
other.  Its two atoms then receive different second half-kicks, by that
pair's force at the cutoff, and RATTLE passes the difference on to their
constraint partners.  (Seen at 2,000 atoms: FP32 r^2 = 100.0 exactly at the
GPU's positions, 99.999992 at the oracle's; those atoms differed by 0.08 A/ps,
every other atom by 1e-4 of the RMS velocity.)

`compare_step` therefore finds the pairs within 1e-4 A^2 of rc^2 at either
side's new positions (an FP32 ulp of r^2 near 100 is 7.6e-6; one step's
GPU - oracle position difference of ~1e-6 A moves r^2 by ~2e-5), holds the
atoms of their constraint groups to the impulse that pair's cutoff force
carries, 1.5 |F(rc)| dt/2 418.4 / m, and every other atom to the parity bar
(relative RMS over all atoms as before).
Positions are unaffected: the step's drift and SHAKE precede its force
evaluation.
"""
from __future__ import annotations

import numpy as np
from scipy.spatial import cKDTree

COULOMB, CONV = 332.0636, 418.4


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / max(1e-300, (b ** 2).sum())))


def boundary_pairs(xa, xb, rc, slack=1e-4):
    """Pairs (i, j) with |r^2 - rc^2| <= slack at positions xa or xb."""
    xa = np.asarray(xa, np.float64)
    xb = np.asarray(xb, np.float64)
    pr = cKDTree(xa).query_pairs(rc + 0.05, output_type="ndarray")
    if len(pr) == 0:
        return pr.reshape(0, 2)

    def r2(x):
        d = x[pr[:, 0]] - x[pr[:, 1]]
        return (d * d).sum(1)

    near = (np.abs(r2(xa) - rc * rc) <= slack) | (np.abs(r2(xb) - rc * rc) <= slack)
    return pr[near]


def cutoff_force(top, i, j, rc):
    """|F| of the pair (i, j) at r = rc: Lennard-Jones (Lorentz-Berthelot) plus Coulomb."""
    q = top["charges"].astype(np.float64)
    s = top["sigma"].astype(np.float64)
    e = top["epsilon"].astype(np.float64)
    sig, eps = 0.5 * (s[i] + s[j]), np.sqrt(e[i] * e[j])
    sr6 = (sig / rc) ** 6
    f_over_r = 24.0 * eps * (2.0 * sr6 * sr6 - sr6) / rc ** 2 + COULOMB * q[i] * q[j] / rc ** 3
    return abs(f_over_r) * rc


def groups_of(top):
    """Atom -> list of the atoms of its constraint group."""
    n = top.n_atoms
    out = {a: [a] for a in range(n)}
    c = np.asarray(top["cluster_central"])
    h = np.asarray(top["cluster_hydrogens"]).reshape(-1, 3)
    for k, a in enumerate(c):
        g = [int(a)] + [int(x) for x in h[k] if x >= 0]
        for b in g:
            out[b] = g
    return out


def compare_step(top, params, xg, vg, xo, vo, dx_bar=2e-5, dv_bar=1e-4, tag=""):
    """Assert one step's GPU state (xg, vg) against the oracle's (xo, vo)."""
    rc = float(params.cutoff_A)
    dx = float(np.abs(xg - xo).max())
    m = top["masses"].astype(np.float64)
    vg = vg.astype(np.float64)
    bp = boundary_pairs(xg, xo, rc)
    affected = np.zeros(len(m), bool)
    allow = np.zeros(len(m))
    if len(bp):
        grp = groups_of(top)
        for i, j in bp:
            impulse = 1.5 * cutoff_force(top, i, j, rc) * 0.5 * params.dt_ps * CONV
            for a in set(grp[int(i)]) | set(grp[int(j)]):
                affected[a] = True
                allow[a] += impulse / m[a]
    dv_atom = np.linalg.norm(vg - vo, axis=1)
    keep = ~affected
    dv = float(np.sqrt((dv_atom[keep] ** 2).sum() / max(1e-300, (vo ** 2).sum())))
    worst = float((dv_atom[affected] / allow[affected]).max()) if affected.any() else 0.0
    print(f"[{tag}] max|dx| {dx:.2e} A, v rel RMS {dv:.2e} ({len(bp)} pairs at the cutoff, "
          f"their groups' |dv| at {worst:.2f} of the cutoff-force bound)", flush=True)
    assert len(bp) <= 8 + len(m) // 500, (tag, len(bp))  # ~1.3 expected per 2,000 atoms
    assert dx <= dx_bar and dv <= dv_bar, (tag, dx, dv)
    assert worst <= 1.0, (tag, worst)
    return dx, dv
