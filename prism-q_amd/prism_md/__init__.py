"""Python harness over the prism_md C ABI (include/prism_md.h).

The class names mirror the reference host engine so tests read like the
reference's own: ``FusedEngine`` plays ``NhsAmberFusedEngine``
(crates/prism-nhs/src/fused_engine.rs:6554) with ``step`` / ``run`` /
``get_positions`` / ``get_velocities`` / ``rebuild_neighbor_lists``.  All
compute happens in the HIP library ``libprism_md.so``; this module only moves
arrays.  A missing library raises ``LibraryMissing`` -- there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.dirname(_HERE)

OK = 0
E_INVALID, E_CAPACITY, E_HIP, E_NONFINITE, E_CONSTRAINT, E_NOMEM, E_UNSUPPORTED, E_NODEVICE, E_DIVERGED = range(-1, -10, -1)
TERM_BOND, TERM_ANGLE, TERM_DIHEDRAL, TERM_NONBONDED, TERM_EXCEPTION, TERM_ALL = 1, 2, 4, 8, 16, 31
TERM_NAMES = ("bond", "angle", "dihedral", "nonbonded", "exception")
PAIRS_REFERENCE, PAIRS_BRUTE = 0, 1
FORCE_FIXED_BITS = 24


class PrismMdError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"prism_md rc={rc}: {msg}")
        self.rc = rc


class LibraryMissing(RuntimeError):
    pass


_f32p = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)


class TopologyV1(C.Structure):
    _fields_ = [
        ("n_atoms", C.c_int32), ("positions", _f32p), ("masses", _f32p), ("charges", _f32p),
        ("sigma", _f32p), ("epsilon", _f32p),
        ("n_bonds", C.c_int32), ("bond_atoms", _i32p), ("bond_r0", _f32p), ("bond_k", _f32p),
        ("n_angles", C.c_int32), ("angle_atoms", _i32p), ("angle_theta0", _f32p), ("angle_k", _f32p),
        ("n_dihedrals", C.c_int32), ("dihedral_atoms", _i32p), ("dihedral_periodicity", _i32p),
        ("dihedral_phase", _f32p), ("dihedral_k", _f32p), ("dihedral_improper", _u8p),
        ("exclusion_offsets", _i32p), ("exclusion_atoms", _i32p),
        ("n_exceptions", C.c_int32), ("exception_atoms", _i32p), ("exception_charge_product", _f32p),
        ("exception_sigma", _f32p), ("exception_epsilon", _f32p),
        ("n_clusters", C.c_int32), ("cluster_central", _i32p), ("cluster_hydrogens", _i32p),
        ("cluster_bond_lengths", _f32p), ("cluster_n_hydrogens", _i32p),
        ("cluster_inv_mass_central", _f32p), ("cluster_inv_mass_h", _f32p),
    ]


class ParamsV1(C.Structure):
    _fields_ = [
        ("dt_ps", C.c_double), ("temperature_K", C.c_double), ("gamma_ps", C.c_double),
        ("cutoff_A", C.c_double), ("skin_A", C.c_double), ("rebuild_every", C.c_int32),
        ("com_remove_every", C.c_int32), ("seed", C.c_uint64), ("stream_id", C.c_uint32),
        ("constraint_max_iter", C.c_int32), ("shake_tol", C.c_double), ("rattle_tol", C.c_double),
        ("end_temperature_K", C.c_double), ("cold_hold_end", C.c_int32), ("ramp_end", C.c_int32),
        ("warm_hold_end", C.c_int32), ("ramp_down_end", C.c_int32), ("warm_tail", C.c_int32),
        ("equilibration_steps", C.c_int32), ("equilibration_gamma_ps", C.c_double),
        ("resort_every", C.c_int32), ("sort_cell_A", C.c_float), ("chunk_tiles", C.c_int32),
        ("reserved", C.c_int32 * 5),
    ]


class SlotCapsV1(C.Structure):
    _fields_ = [("atoms", C.c_int32), ("bonds", C.c_int32), ("angles", C.c_int32), ("dihedrals", C.c_int32),
                ("exceptions", C.c_int32), ("exclusions", C.c_int32), ("reserved", C.c_int32 * 2)]


class StatsV1(C.Structure):
    _fields_ = [("steps_done", C.c_uint64), ("pair_rebuilds", C.c_uint64), ("force_evaluations", C.c_uint64),
                ("n_tiles", C.c_int32), ("n_blocks", C.c_int32), ("status", C.c_int32),
                ("max_shake_sweeps", C.c_int32), ("max_rattle_sweeps", C.c_int32), ("reserved", C.c_int32)]


class GenParamsV1(C.Structure):
    _fields_ = [("n_atoms", C.c_int32), ("seed", C.c_uint64)] + [
        (n, C.c_double) for n in ("density", "heavy_bond_A", "h_bond_A", "heavy_angle_deg", "hmr_factor",
                                  "charge_max", "sigma_min", "sigma_max", "epsilon_min", "epsilon_max",
                                  "bond_k_min", "bond_k_max", "angle_k_min", "angle_k_max", "torsion_vn_max",
                                  "scale14_lj", "scale14_qq")
    ] + [("n_candidates", C.c_int32), ("reserved", C.c_int32 * 7)]


class GenStatsV1(C.Structure):
    _fields_ = [("n_heavy", C.c_int32), ("n_hydrogen", C.c_int32), ("sphere_radius_A", C.c_double),
                ("min_nonbonded_distance_A", C.c_double), ("mean_nearest_nonbonded_A", C.c_double),
                ("total_charge", C.c_double), ("total_mass", C.c_double)]


class TopoLoadOptionsV1(C.Structure):
    _fields_ = [("harmonic_convention", C.c_int32), ("apply_hmr", C.c_int32), ("hmr_factor", C.c_double),
                ("reserved", C.c_int32 * 8)]


class TopoInfoV1(C.Structure):
    _fields_ = [("n_atoms", C.c_int32), ("n_residues", C.c_int32), ("n_chains", C.c_int32),
                ("hmr_in_file", C.c_int32), ("hmr_applied_here", C.c_int32), ("hmr_skipped", C.c_int32),
                ("harmonic_convention", C.c_int32), ("has_exception_contract", C.c_int32),
                ("recommended_timestep_fs", C.c_double), ("reserved", C.c_int32 * 8)]


_LIBS: dict = {}


def library_path(name: str) -> str:
    return os.path.join(_PKG, name)


def _load(name: str):
    if name in _LIBS:
        return _LIBS[name]
    path = library_path(name)
    if not os.path.exists(path):
        raise LibraryMissing(f"{path} is not built; run __graft_entry__.build()")
    lib = C.CDLL(path)
    _LIBS[name] = lib
    return lib


def md_lib():
    lib = _load("libprism_md.so")
    if not getattr(lib, "_typed", False):
        H = C.c_void_p
        sig = {
            "prism_md_abi_version": ([], C.c_int),
            "prism_md_default_params": ([C.POINTER(ParamsV1)], C.c_int),
            "prism_md_create": ([C.POINTER(TopologyV1), C.POINTER(ParamsV1), C.c_void_p, C.POINTER(H)], C.c_int),
            "prism_md_destroy": ([H], C.c_int),
            "prism_md_set_state": ([H, _f32p, _f32p, C.c_int], C.c_int),
            "prism_md_get_state": ([H, _f32p, _f32p, C.c_int], C.c_int),
            "prism_md_build_pairs": ([H, C.c_int], C.c_int),
            "prism_md_build_pairs_static": ([H, C.c_int], C.c_int),
            "prism_md_set_step_limit": ([H, C.c_uint64], C.c_int),
            "prism_md_query_latency": ([H, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_uint64)],
                                       C.c_int),
            "prism_md_batch_set_step_limit": ([H, C.c_int32, C.c_uint64], C.c_int),
            "prism_md_batch_query_latency": ([H, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                              C.POINTER(C.c_uint64)], C.c_int),
            "prism_md_batch_latency_samples": ([H, C.c_int32, _f32p, C.c_size_t, C.POINTER(C.c_size_t)], C.c_int),
            "prism_md_get_pairs": ([H, _i32p, _i32p, C.c_size_t, C.POINTER(C.c_size_t)], C.c_int),
            "prism_md_compute_forces": ([H, C.c_uint, _f32p, C.POINTER(C.c_double)], C.c_int),
            "prism_md_compute_forces_fixed": ([H, C.c_uint, C.POINTER(C.c_int64)], C.c_int),
            "prism_md_step": ([H, C.c_uint64, _f32p], C.c_int),
            "prism_md_energies": ([H, C.POINTER(C.c_double)], C.c_int),
            "prism_md_kinetic_temperature": ([H, C.POINTER(C.c_double)], C.c_int),
            "prism_md_batch_kinetic_temperature": ([H, C.c_int32, C.POINTER(C.c_double)], C.c_int),
            "prism_md_get_stats": ([H, C.POINTER(StatsV1)], C.c_int),
            "prism_md_last_error": ([H], C.c_char_p),
            "prism_md_count_pairs": ([H, C.POINTER(C.c_uint64)], C.c_int),
            "prism_md_time_nonbonded": ([H, C.c_int, C.POINTER(C.c_double)], C.c_int),
            "prism_md_checkpoint_size": ([H, C.POINTER(C.c_size_t)], C.c_int),
            "prism_md_checkpoint_save": ([H, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)], C.c_int),
            "prism_md_checkpoint_restore": ([H, C.c_void_p, C.c_size_t], C.c_int),
            "prism_md_batch_checkpoint_size": ([H, C.c_int32, C.POINTER(C.c_size_t)], C.c_int),
            "prism_md_batch_checkpoint_save": ([H, C.c_int32, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)], C.c_int),
            "prism_md_batch_checkpoint_restore": ([H, C.c_int32, C.c_void_p, C.c_size_t], C.c_int),
            "prism_md_kernel_timing": ([H, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_uint64)], C.c_int),
            "prism_md_batch_create": ([C.POINTER(C.POINTER(TopologyV1)), C.POINTER(ParamsV1), C.c_int32, C.c_void_p,
                                       C.POINTER(H)], C.c_int),
            "prism_md_batch_destroy": ([H], C.c_int),
            "prism_md_batch_size": ([H, C.POINTER(C.c_int32), C.POINTER(C.c_int64)], C.c_int),
            "prism_md_batch_step": ([H, C.c_uint64, C.c_int32], C.c_int),
            "prism_md_batch_minimize": ([H, C.c_int, C.c_double, C.POINTER(C.c_double)], C.c_int),
            "prism_md_batch_set_active": ([H, C.c_int32, C.c_int], C.c_int),
            "prism_md_batch_set_state": ([H, C.c_int32, _f32p, _f32p, C.c_int], C.c_int),
            "prism_md_batch_get_state": ([H, C.c_int32, _f32p, _f32p, C.c_int], C.c_int),
            "prism_md_batch_compute_forces": ([H, C.c_int32, C.c_uint, _f32p, C.POINTER(C.c_double)], C.c_int),
            "prism_md_batch_energies": ([H, C.c_int32, C.POINTER(C.c_double)], C.c_int),
            "prism_md_batch_get_stats": ([H, C.c_int32, C.POINTER(StatsV1)], C.c_int),
            "prism_md_batch_last_error": ([H], C.c_char_p),
            "prism_md_slot_caps_for": ([C.POINTER(TopologyV1), C.POINTER(SlotCapsV1)], C.c_int),
            "prism_md_pool_create": ([C.POINTER(SlotCapsV1), C.c_int32, C.c_void_p, C.POINTER(H)], C.c_int),
            "prism_md_batch_admit": ([H, C.c_int32, C.POINTER(TopologyV1), C.POINTER(ParamsV1)], C.c_int),
            "prism_md_batch_retire": ([H, C.c_int32], C.c_int),
            "prism_md_batch_slot_info": ([H, C.c_int32, C.POINTER(SlotCapsV1), C.POINTER(C.c_int32)], C.c_int),
            "prism_md_batch_device_bytes": ([H, C.POINTER(C.c_uint64)], C.c_int),
            "prism_md_minimize": ([H, C.c_int, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        lib._typed = True
    return lib


def gen_lib():
    lib = _load("libprism_gen.so")
    if not getattr(lib, "_typed", False):
        H = C.c_void_p
        lib.prism_gen_default_params.argtypes = [C.POINTER(GenParamsV1)]
        lib.prism_gen_create.argtypes = [C.POINTER(GenParamsV1), C.POINTER(H)]
        lib.prism_gen_topology.argtypes = [H]
        lib.prism_gen_topology.restype = C.POINTER(TopologyV1)
        lib.prism_gen_stats.argtypes = [H, C.POINTER(GenStatsV1)]
        lib.prism_gen_destroy.argtypes = [H]
        lib.prism_gen_maxwell_boltzmann.argtypes = [_f32p, C.c_int32, C.c_double, C.c_uint64, _f32p]
        lib.prism_topo_default_options.argtypes = [C.POINTER(TopoLoadOptionsV1)]
        lib.prism_topo_load_json.argtypes = [C.c_char_p, C.POINTER(TopoLoadOptionsV1), C.POINTER(H)]
        lib.prism_topo_parse_json.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(TopoLoadOptionsV1), C.POINTER(H)]
        lib.prism_topo_view.argtypes = [H]
        lib.prism_topo_view.restype = C.POINTER(TopologyV1)
        lib.prism_topo_info.argtypes = [H, C.POINTER(TopoInfoV1)]
        lib.prism_topo_destroy.argtypes = [H]
        lib.prism_topo_last_error.argtypes = []
        lib.prism_topo_last_error.restype = C.c_char_p
        lib.prism_topo_write_json.argtypes = [C.POINTER(TopologyV1), C.c_int32, C.c_char_p]
        lib.prism_topo_sha256_hex.argtypes = [C.c_void_p, C.c_size_t, C.c_char_p]
        lib._typed = True
    return lib


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None and a.size else None


_ARRAYS = {  # field -> (dtype, arity, count field)
    "positions": (np.float32, 3, "n_atoms"), "masses": (np.float32, 1, "n_atoms"),
    "charges": (np.float32, 1, "n_atoms"), "sigma": (np.float32, 1, "n_atoms"),
    "epsilon": (np.float32, 1, "n_atoms"),
    "bond_atoms": (np.int32, 2, "n_bonds"), "bond_r0": (np.float32, 1, "n_bonds"), "bond_k": (np.float32, 1, "n_bonds"),
    "angle_atoms": (np.int32, 3, "n_angles"), "angle_theta0": (np.float32, 1, "n_angles"),
    "angle_k": (np.float32, 1, "n_angles"),
    "dihedral_atoms": (np.int32, 4, "n_dihedrals"), "dihedral_periodicity": (np.int32, 1, "n_dihedrals"),
    "dihedral_phase": (np.float32, 1, "n_dihedrals"), "dihedral_k": (np.float32, 1, "n_dihedrals"),
    "dihedral_improper": (np.uint8, 1, "n_dihedrals"),
    "exception_atoms": (np.int32, 2, "n_exceptions"), "exception_charge_product": (np.float32, 1, "n_exceptions"),
    "exception_sigma": (np.float32, 1, "n_exceptions"), "exception_epsilon": (np.float32, 1, "n_exceptions"),
    "cluster_central": (np.int32, 1, "n_clusters"), "cluster_hydrogens": (np.int32, 3, "n_clusters"),
    "cluster_bond_lengths": (np.float32, 3, "n_clusters"), "cluster_n_hydrogens": (np.int32, 1, "n_clusters"),
    "cluster_inv_mass_central": (np.float32, 1, "n_clusters"), "cluster_inv_mass_h": (np.float32, 1, "n_clusters"),
}
_CTYPE = {np.float32: C.c_float, np.int32: C.c_int32, np.uint8: C.c_uint8}


def _copy_struct(t: TopologyV1) -> dict:
    """numpy copies of the arrays a PrismMdTopologyV1 points to."""
    arrays = {}
    for name, (dt, ar, cnt) in _ARRAYS.items():
        n = getattr(t, cnt)
        p = getattr(t, name)
        arrays[name] = np.ctypeslib.as_array(p, shape=(n * ar,)).astype(dt).copy() if n and p else np.zeros(0, dt)
    n = t.n_atoms
    off = np.ctypeslib.as_array(t.exclusion_offsets, shape=(n + 1,)).astype(np.int32).copy()
    arrays["exclusion_offsets"] = off
    arrays["exclusion_atoms"] = (np.ctypeslib.as_array(t.exclusion_atoms, shape=(int(off[-1]),))
                                 .astype(np.int32).copy() if off[-1] else np.zeros(0, np.int32))
    return arrays


class Topology:
    """Owning numpy copy of a PrismMdTopologyV1 (the PrismPrepTopology fields the step reads)."""

    def __init__(self, arrays: dict):
        self.a = {k: np.ascontiguousarray(v) for k, v in arrays.items()}
        self._struct = None

    @property
    def n_atoms(self) -> int:
        return int(self.a["masses"].shape[0])

    def __getitem__(self, k):
        return self.a[k]

    @classmethod
    def generate(cls, n_atoms: int, seed: int = 0x50524953, **overrides) -> "Topology":
        """Synthetic AMBER-typed chain from libprism_gen (see include/prism_gen.h)."""
        lib = gen_lib()
        gp = GenParamsV1()
        lib.prism_gen_default_params(C.byref(gp))
        gp.n_atoms = int(n_atoms)
        gp.seed = int(seed)
        for k, v in overrides.items():
            setattr(gp, k, v)
        h = C.c_void_p()
        rc = lib.prism_gen_create(C.byref(gp), C.byref(h))
        if rc != OK:
            raise PrismMdError(rc, "generator failed")
        try:
            st = GenStatsV1()
            lib.prism_gen_stats(h, C.byref(st))
            top = cls(_copy_struct(lib.prism_gen_topology(h).contents))
            top.gen_stats = {f: getattr(st, f) for f, _ in GenStatsV1._fields_}
            return top
        finally:
            lib.prism_gen_destroy(h)

    @classmethod
    def from_json(cls, path: str, convention: str | None = None, apply_hmr: bool = True,
                  hmr_factor: float = 3.0) -> "Topology":
        """Read a PrismPrepTopology JSON file (the reference's on-disk topology,
        crates/prism-nhs/src/input.rs:1027-1182) through prism_topo_load_json:
        the reference's validation, harmonic lowering and HMR.  `convention`
        None requires the file's harmonic_parameter_contract; "amber" or
        "openmm" states the coefficient convention of a file without one."""
        lib = gen_lib()
        o = TopoLoadOptionsV1()
        lib.prism_topo_default_options(C.byref(o))
        o.harmonic_convention = {None: 0, "amber": 1, "openmm": 2}[convention]
        o.apply_hmr = int(bool(apply_hmr))
        o.hmr_factor = float(hmr_factor)
        h = C.c_void_p()
        rc = lib.prism_topo_load_json(os.fsencode(path), C.byref(o), C.byref(h))
        if rc != OK:
            raise PrismMdError(rc, lib.prism_topo_last_error().decode())
        try:
            info = TopoInfoV1()
            lib.prism_topo_info(h, C.byref(info))
            top = cls(_copy_struct(lib.prism_topo_view(h).contents))
            top.json_info = {f: getattr(info, f) for f, _ in TopoInfoV1._fields_ if f != "reserved"}
            return top
        finally:
            lib.prism_topo_destroy(h)

    def to_json(self, path: str, hmr_applied: bool | None = True):
        """Write this topology as a PrismPrepTopology JSON document
        (prism_topo_write_json); hmr_applied None leaves the key out."""
        flag = -1 if hmr_applied is None else int(bool(hmr_applied))
        lib = gen_lib()
        rc = lib.prism_topo_write_json(C.byref(self.struct()), flag, os.fsencode(path))
        if rc != OK:
            raise PrismMdError(rc, lib.prism_topo_last_error().decode())

    def struct(self) -> TopologyV1:
        if self._struct is None:
            s = TopologyV1()
            s.n_atoms = self.n_atoms
            counts = {"n_atoms": self.n_atoms, "n_bonds": len(self.a["bond_r0"]), "n_angles": len(self.a["angle_k"]),
                      "n_dihedrals": len(self.a["dihedral_k"]), "n_exceptions": len(self.a["exception_sigma"]),
                      "n_clusters": len(self.a["cluster_central"])}
            for k, v in counts.items():
                setattr(s, k, v)
            for name, (dt, _, _) in _ARRAYS.items():
                setattr(s, name, _ptr(self.a[name], _CTYPE[dt]))
            s.exclusion_offsets = _ptr(self.a["exclusion_offsets"], C.c_int32)
            s.exclusion_atoms = _ptr(self.a["exclusion_atoms"], C.c_int32)
            self._struct = s
        return self._struct

    def with_positions(self, x) -> "Topology":
        arrays = dict(self.a)
        arrays["positions"] = np.asarray(x, np.float32).reshape(-1)
        t = Topology(arrays)
        t.gen_stats = getattr(self, "gen_stats", None)
        return t


def default_params(**kw) -> ParamsV1:
    p = ParamsV1()
    md_lib().prism_md_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def maxwell_boltzmann(masses, temperature_K: float, seed: int) -> np.ndarray:
    m = np.ascontiguousarray(masses, np.float32)
    out = np.zeros(3 * m.size, np.float32)
    rc = gen_lib().prism_gen_maxwell_boltzmann(_ptr(m, C.c_float), m.size, temperature_K, seed, _ptr(out, C.c_float))
    if rc != OK:
        raise PrismMdError(rc, "maxwell_boltzmann")
    return out.reshape(-1, 3)


class FusedEngine:
    """One system on the GPU through the C ABI (role of NhsAmberFusedEngine)."""

    def __init__(self, topology: Topology, params: ParamsV1 | None = None, hip_stream: int | None = None):
        self.lib = md_lib()
        self.topology = topology
        self.params = params if params is not None else default_params()
        self.n_atoms = topology.n_atoms
        h = C.c_void_p()
        rc = self.lib.prism_md_create(C.byref(topology.struct()), C.byref(self.params),
                                      C.c_void_p(hip_stream) if hip_stream else None, C.byref(h))
        if rc != OK:
            self.h = None
            raise PrismMdError(rc, "prism_md_create: " + self.lib.prism_md_last_error(None).decode())
        self.h = h

    def _check(self, rc, what):
        if rc != OK:
            raise PrismMdError(rc, f"{what}: {self.lib.prism_md_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.prism_md_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state
    def set_state(self, x=None, v=None):
        xa = None if x is None else np.ascontiguousarray(x, np.float32).reshape(-1)
        va = None if v is None else np.ascontiguousarray(v, np.float32).reshape(-1)
        self._check(self.lib.prism_md_set_state(self.h, _ptr(xa, C.c_float), _ptr(va, C.c_float), 0), "set_state")

    def get_positions(self) -> np.ndarray:
        x = np.zeros(3 * self.n_atoms, np.float32)
        self._check(self.lib.prism_md_get_state(self.h, _ptr(x, C.c_float), None, 0), "get_state")
        return x.reshape(-1, 3)

    def get_velocities(self) -> np.ndarray:
        v = np.zeros(3 * self.n_atoms, np.float32)
        self._check(self.lib.prism_md_get_state(self.h, None, _ptr(v, C.c_float), 0), "get_state")
        return v.reshape(-1, 3)

    # -- pair list
    def rebuild_neighbor_lists(self, mode: int = PAIRS_BRUTE):
        self._check(self.lib.prism_md_build_pairs(self.h, mode), "build_pairs")

    def pairs(self, mode: int = PAIRS_BRUTE, static: bool = False) -> np.ndarray:
        """Unique pairs (i < j) of the list built at the current positions;
        static=True stages them in the library's static device array
        (prism_md_build_pairs_static, no allocation)."""
        if static:
            self._check(self.lib.prism_md_build_pairs_static(self.h, mode), "build_pairs_static")
        else:
            self.rebuild_neighbor_lists(mode)
        n = C.c_size_t()
        self.lib.prism_md_get_pairs(self.h, None, None, 0, C.byref(n))
        i = np.zeros(max(1, n.value), np.int32)
        j = np.zeros(max(1, n.value), np.int32)
        self._check(self.lib.prism_md_get_pairs(self.h, _ptr(i, C.c_int32), _ptr(j, C.c_int32), n.value, C.byref(n)),
                    "get_pairs")
        return np.stack([i[: n.value], j[: n.value]], axis=1)

    # -- forces / energies
    def compute_forces(self, mask: int = TERM_ALL, energies: bool = True):
        F = np.zeros(3 * self.n_atoms, np.float32)
        E = (C.c_double * 5)()
        self._check(self.lib.prism_md_compute_forces(self.h, mask, _ptr(F, C.c_float), E if energies else None),
                    "compute_forces")
        return F.reshape(-1, 3), np.array(list(E))

    def compute_forces_fixed(self, mask: int = TERM_ALL) -> np.ndarray:
        F = np.zeros(3 * self.n_atoms, np.int64)
        self._check(self.lib.prism_md_compute_forces_fixed(self.h, mask, F.ctypes.data_as(C.POINTER(C.c_int64))),
                    "compute_forces_fixed")
        return F.reshape(-1, 3)

    def energies(self):
        """Potential and kinetic energy and the cumulative work ledger W
        (kcal/mol), plus the kinetic temperature (K)."""
        out = (C.c_double * 3)()
        self._check(self.lib.prism_md_energies(self.h, out), "energies")
        T = C.c_double()
        self._check(self.lib.prism_md_kinetic_temperature(self.h, C.byref(T)), "kinetic_temperature")
        return {"potential": out[0], "kinetic": out[1], "work": out[2], "temperature": T.value}

    # -- stepping
    def step(self, n_steps: int = 1, noise=None):
        nz = None if noise is None else np.ascontiguousarray(noise, np.float32).reshape(-1)
        if nz is not None and nz.size != n_steps * 2 * self.n_atoms * 3:
            raise ValueError("noise must be [n_steps][2][n_atoms][3]")
        self._check(self.lib.prism_md_step(self.h, n_steps, _ptr(nz, C.c_float)), "step")

    run = step

    def minimize(self, max_iter: int = 3000, f_tol: float = 1.0):
        """FIRE minimisation on the GPU; returns (potential energy, max |F|)."""
        e, f = C.c_double(), C.c_double()
        self._check(self.lib.prism_md_minimize(self.h, max_iter, f_tol, C.byref(e), C.byref(f)), "minimize")
        return e.value, f.value

    def stats(self) -> dict:
        s = StatsV1()
        self._check(self.lib.prism_md_get_stats(self.h, C.byref(s)), "stats")
        return {f: getattr(s, f) for f, _ in StatsV1._fields_ if f != "reserved"}

    def count_pairs(self) -> int:
        n = C.c_uint64()
        self._check(self.lib.prism_md_count_pairs(self.h, C.byref(n)), "count_pairs")
        return int(n.value)

    def kernel_timing(self, every: int):
        """(Re)start sampling the nonbonded launch of every `every`-th step
        (0 stops); returns (mean ms, samples) of the samples so far."""
        ms, n = C.c_double(), C.c_uint64()
        self._check(self.lib.prism_md_kernel_timing(self.h, every, C.byref(ms), C.byref(n)), "kernel_timing")
        return float(ms.value), int(n.value)

    def set_step_limit(self, end_step: int):
        """Stop advancing at step count end_step (0: no limit)."""
        self._check(self.lib.prism_md_set_step_limit(self.h, end_step), "set_step_limit")

    def latency(self) -> dict:
        """p50 / p99 per-step latency (ms) from the device's step-end stamps."""
        p50, p99, n = C.c_double(), C.c_double(), C.c_uint64()
        self._check(self.lib.prism_md_query_latency(self.h, C.byref(p50), C.byref(p99), C.byref(n)), "query_latency")
        return {"p50_ms": p50.value, "p99_ms": p99.value, "samples": int(n.value)}

    def checkpoint(self) -> bytes:
        """The system's dynamic state (prism_md_checkpoint_save): restoring it
        into an engine of the same topology and parameters continues the
        trajectory bitwise."""
        n = C.c_size_t()
        self._check(self.lib.prism_md_checkpoint_size(self.h, C.byref(n)), "checkpoint_size")
        buf = C.create_string_buffer(n.value)
        w = C.c_size_t()
        self._check(self.lib.prism_md_checkpoint_save(self.h, buf, n, C.byref(w)), "checkpoint_save")
        return buf.raw[:w.value]

    def restore(self, blob: bytes):
        self._check(self.lib.prism_md_checkpoint_restore(self.h, blob, len(blob)), "checkpoint_restore")

    def time_nonbonded(self, iters: int = 20) -> float:
        ms = C.c_double()
        self._check(self.lib.prism_md_time_nonbonded(self.h, iters, C.byref(ms)), "time_nonbonded")
        return float(ms.value)


class Ensemble:
    """Many independent systems in one engine (prism_md_batch_*): one launch
    per phase over all systems, stepped through a captured hipGraph.  Each
    system keeps its own parameters, noise stream, rebuilds and faults; its
    trajectory is bitwise the one it follows alone."""

    def __init__(self, topologies, params=None, hip_stream: int | None = None, _handle=None):
        self.lib = md_lib()
        if _handle is not None:  # a pool: slots filled by admit()
            self.h, self.topologies, self.params, self._structs = _handle, list(topologies), list(params), []
            return
        self.topologies = list(topologies)
        n = len(self.topologies)
        if params is None:
            params = [default_params() for _ in range(n)]
        elif isinstance(params, ParamsV1):
            params = [params] * n
        self.params = list(params)
        self._structs = [t.struct() for t in self.topologies]
        tops = (C.POINTER(TopologyV1) * n)(*[C.pointer(s) for s in self._structs])
        prm = (ParamsV1 * n)(*self.params)
        h = C.c_void_p()
        rc = self.lib.prism_md_batch_create(tops, prm, n, C.c_void_p(hip_stream) if hip_stream else None, C.byref(h))
        if rc != OK:
            self.h = None
            raise PrismMdError(rc, "prism_md_batch_create: " + self.lib.prism_md_batch_last_error(None).decode())
        self.h = h

    def _check(self, rc, what):
        if rc != OK:
            raise PrismMdError(rc, f"{what}: {self.lib.prism_md_batch_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.prism_md_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def n_systems(self) -> int:
        return len(self.topologies)

    def n_atoms(self) -> int:
        n, a = C.c_int32(), C.c_int64()
        self._check(self.lib.prism_md_batch_size(self.h, C.byref(n), C.byref(a)), "batch_size")
        return int(a.value)

    def step(self, n_steps: int, graph_steps: int = 10):
        self._check(self.lib.prism_md_batch_step(self.h, n_steps, graph_steps), "batch_step")

    def minimize(self, max_iter: int = 2000, f_tol: float = 1.0) -> float:
        f = C.c_double()
        self._check(self.lib.prism_md_batch_minimize(self.h, max_iter, f_tol, C.byref(f)), "batch_minimize")
        return f.value

    def set_active(self, s: int, active: bool):
        self._check(self.lib.prism_md_batch_set_active(self.h, s, int(bool(active))), "batch_set_active")

    def set_state(self, s: int, x=None, v=None):
        xa = None if x is None else np.ascontiguousarray(x, np.float32).reshape(-1)
        va = None if v is None else np.ascontiguousarray(v, np.float32).reshape(-1)
        self._check(self.lib.prism_md_batch_set_state(self.h, s, _ptr(xa, C.c_float), _ptr(va, C.c_float), 0),
                    "batch_set_state")

    def get_positions(self, s: int) -> np.ndarray:
        x = np.zeros(3 * self.topologies[s].n_atoms, np.float32)
        self._check(self.lib.prism_md_batch_get_state(self.h, s, _ptr(x, C.c_float), None, 0), "batch_get_state")
        return x.reshape(-1, 3)

    def get_velocities(self, s: int) -> np.ndarray:
        v = np.zeros(3 * self.topologies[s].n_atoms, np.float32)
        self._check(self.lib.prism_md_batch_get_state(self.h, s, None, _ptr(v, C.c_float), 0), "batch_get_state")
        return v.reshape(-1, 3)

    def compute_forces(self, s: int, mask: int = TERM_ALL):
        F = np.zeros(3 * self.topologies[s].n_atoms, np.float32)
        E = (C.c_double * 5)()
        self._check(self.lib.prism_md_batch_compute_forces(self.h, s, mask, _ptr(F, C.c_float), E),
                    "batch_compute_forces")
        return F.reshape(-1, 3), np.array(list(E))

    def energies(self, s: int):
        out = (C.c_double * 3)()
        self._check(self.lib.prism_md_batch_energies(self.h, s, out), "batch_energies")
        T = C.c_double()
        self._check(self.lib.prism_md_batch_kinetic_temperature(self.h, s, C.byref(T)), "batch_kinetic_temperature")
        return {"potential": out[0], "kinetic": out[1], "work": out[2], "temperature": T.value}

    def set_step_limit(self, s: int, end_step: int):
        self._check(self.lib.prism_md_batch_set_step_limit(self.h, s, end_step), "batch_set_step_limit")

    def latency(self, s: int) -> dict:
        p50, p99, n = C.c_double(), C.c_double(), C.c_uint64()
        self._check(self.lib.prism_md_batch_query_latency(self.h, s, C.byref(p50), C.byref(p99), C.byref(n)),
                    "batch_query_latency")
        return {"p50_ms": p50.value, "p99_ms": p99.value, "samples": int(n.value)}

    def latency_samples(self, s: int) -> np.ndarray:
        n = C.c_size_t()
        self._check(self.lib.prism_md_batch_latency_samples(self.h, s, None, 0, C.byref(n)), "latency_samples")
        out = np.zeros(max(1, n.value), np.float32)
        self._check(self.lib.prism_md_batch_latency_samples(self.h, s, _ptr(out, C.c_float), n.value, C.byref(n)),
                    "latency_samples")
        return out[: n.value]

    def checkpoint(self, s: int) -> bytes:
        n = C.c_size_t()
        self._check(self.lib.prism_md_batch_checkpoint_size(self.h, s, C.byref(n)), "batch_checkpoint_size")
        buf = C.create_string_buffer(n.value)
        w = C.c_size_t()
        self._check(self.lib.prism_md_batch_checkpoint_save(self.h, s, buf, n, C.byref(w)), "batch_checkpoint_save")
        return buf.raw[:w.value]

    def restore(self, s: int, blob: bytes):
        self._check(self.lib.prism_md_batch_checkpoint_restore(self.h, s, blob, len(blob)), "batch_checkpoint_restore")

    def stats(self, s: int) -> dict:
        st = StatsV1()
        self._check(self.lib.prism_md_batch_get_stats(self.h, s, C.byref(st)), "batch_stats")
        return {f: getattr(st, f) for f, _ in StatsV1._fields_ if f != "reserved"}


def slot_caps_for(topology: Topology, atoms: int | None = None, scale: float = 1.0) -> SlotCapsV1:
    """Exact capacities of `topology`; with `atoms`, the same per-atom
    proportions scaled to hold `atoms - 1` atoms (times `scale` headroom)."""
    c = SlotCapsV1()
    rc = md_lib().prism_md_slot_caps_for(C.byref(topology.struct()), C.byref(c))
    if rc != OK:
        raise PrismMdError(rc, "slot_caps_for")
    if atoms is not None:
        f = scale * (atoms - 1) / max(1, topology.n_atoms)
        for k in ("bonds", "angles", "dihedrals", "exceptions", "exclusions"):
            setattr(c, k, int(np.ceil(getattr(c, k) * f)))
        c.atoms = int(atoms)
    return c


class ServingPool(Ensemble):
    """Slots of fixed capacity, filled and emptied while the others keep
    running (prism_md_pool_create / batch_admit / batch_retire)."""

    def __init__(self, caps, hip_stream: int | None = None):
        lib = md_lib()
        arr = (SlotCapsV1 * len(caps))(*caps)
        h = C.c_void_p()
        rc = lib.prism_md_pool_create(arr, len(caps), C.c_void_p(hip_stream) if hip_stream else None, C.byref(h))
        if rc != OK:
            raise PrismMdError(rc, "prism_md_pool_create: " + lib.prism_md_batch_last_error(None).decode())
        super().__init__([None] * len(caps), [None] * len(caps), _handle=h)
        self.caps = list(caps)

    def admit(self, slot: int, topology: Topology, params: ParamsV1 | None = None):
        p = params if params is not None else default_params()
        self._check(self.lib.prism_md_batch_admit(self.h, slot, C.byref(topology.struct()), C.byref(p)), "admit")
        self.topologies[slot] = topology
        self.params[slot] = p

    def retire(self, slot: int):
        self._check(self.lib.prism_md_batch_retire(self.h, slot), "retire")

    def occupied(self, slot: int) -> bool:
        occ = C.c_int32()
        self._check(self.lib.prism_md_batch_slot_info(self.h, slot, None, C.byref(occ)), "slot_info")
        return bool(occ.value)

    def fits(self, slot: int, topology: Topology) -> bool:
        need = slot_caps_for(topology)
        c = self.caps[slot]
        return (topology.n_atoms < c.atoms and need.bonds <= c.bonds and need.angles <= c.angles
                and need.dihedrals <= c.dihedrals and need.exceptions <= c.exceptions
                and need.exclusions <= c.exclusions)

    def device_bytes(self) -> int:
        b = C.c_uint64()
        self._check(self.lib.prism_md_batch_device_bytes(self.h, C.byref(b)), "device_bytes")
        return int(b.value)
