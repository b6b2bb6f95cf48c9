"""Hand-built tiny topologies for closed-form checks (test helper)."""
from __future__ import annotations

import numpy as np


def make_topology(pm, positions, masses=None, charges=None, sigma=None, epsilon=None, bonds=(), angles=(),
                  dihedrals=(), exclusions=None, exceptions=(), clusters=()):
    """positions [n,3]; bonds (i, j, r0, k); angles (i, j, k, theta0, k);
    dihedrals (i, j, k, l, n, phase, k); exceptions (i, j, qq, sigma, eps);
    clusters (central, [h...], [d...]) with HMR-free inverse masses taken from
    `masses`; exclusions: list of partner lists (symmetric) or None (none)."""
    x = np.asarray(positions, np.float32).reshape(-1, 3)
    n = len(x)
    f = lambda v, d: np.full(n, d, np.float32) if v is None else np.asarray(v, np.float32)
    a = {"positions": x.reshape(-1), "masses": f(masses, 12.0), "charges": f(charges, 0.0),
         "sigma": f(sigma, 1.0), "epsilon": f(epsilon, 0.0)}
    a["bond_atoms"] = np.array([[b[0], b[1]] for b in bonds], np.int32).reshape(-1)
    a["bond_r0"] = np.array([b[2] for b in bonds], np.float32)
    a["bond_k"] = np.array([b[3] for b in bonds], np.float32)
    a["angle_atoms"] = np.array([[g[0], g[1], g[2]] for g in angles], np.int32).reshape(-1)
    a["angle_theta0"] = np.array([g[3] for g in angles], np.float32)
    a["angle_k"] = np.array([g[4] for g in angles], np.float32)
    a["dihedral_atoms"] = np.array([[d[0], d[1], d[2], d[3]] for d in dihedrals], np.int32).reshape(-1)
    a["dihedral_periodicity"] = np.array([d[4] for d in dihedrals], np.int32)
    a["dihedral_phase"] = np.array([d[5] for d in dihedrals], np.float32)
    a["dihedral_k"] = np.array([d[6] for d in dihedrals], np.float32)
    a["dihedral_improper"] = np.zeros(len(dihedrals), np.uint8)
    a["exception_atoms"] = np.array([[e[0], e[1]] for e in exceptions], np.int32).reshape(-1)
    a["exception_charge_product"] = np.array([e[2] for e in exceptions], np.float32)
    a["exception_sigma"] = np.array([e[3] for e in exceptions], np.float32)
    a["exception_epsilon"] = np.array([e[4] for e in exceptions], np.float32)
    part = exclusions if exclusions is not None else [[] for _ in range(n)]
    off = np.zeros(n + 1, np.int32)
    for i in range(n):
        off[i + 1] = off[i] + len(part[i])
    a["exclusion_offsets"] = off
    a["exclusion_atoms"] = np.array([j for p in part for j in sorted(p)], np.int32)
    m = a["masses"]
    a["cluster_central"] = np.array([c[0] for c in clusters], np.int32)
    a["cluster_hydrogens"] = np.array([list(c[1]) + [-1] * (3 - len(c[1])) for c in clusters], np.int32).reshape(-1)
    a["cluster_bond_lengths"] = np.array([list(c[2]) + [0.0] * (3 - len(c[2])) for c in clusters],
                                         np.float32).reshape(-1)
    a["cluster_n_hydrogens"] = np.array([len(c[1]) for c in clusters], np.int32)
    a["cluster_inv_mass_central"] = np.array([1.0 / m[c[0]] for c in clusters], np.float32)
    a["cluster_inv_mass_h"] = np.array([1.0 / m[c[1][0]] for c in clusters], np.float32)
    return pm.Topology(a)
