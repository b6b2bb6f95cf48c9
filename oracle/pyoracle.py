"""ctypes wrapper of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_HERE)
sys.path.insert(0, os.path.join(_ROOT, "prism-q_amd"))
from prism_md import ParamsV1, TopologyV1  # noqa: E402  (struct layouts only)

_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        fp = C.POINTER(C.c_float)
        ip = C.POINTER(C.c_int32)
        up = C.POINTER(C.c_uint32)
        L.oracle_forces.argtypes = [C.POINTER(TopologyV1), dp, C.c_uint, C.c_int, dp, dp]
        L.oracle_pairs.argtypes = [C.POINTER(TopologyV1), fp, fp, C.c_int, ip, ip, C.c_int64]
        L.oracle_pairs.restype = C.c_int64
        L.oracle_run.argtypes = [C.POINTER(TopologyV1), C.POINTER(ParamsV1), dp, dp, C.c_int64, C.c_int64, fp, dp]
        L.oracle_energies.argtypes = [C.POINTER(TopologyV1), dp, dp, dp, dp]
        L.oracle_philox4x32_10.argtypes = [up, up, up]
        L.oracle_langevin_normals.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32, dp]
        L.oracle_schedule.argtypes = [C.POINTER(ParamsV1), C.c_uint32, dp, dp, C.POINTER(C.c_int)]
        L.oracle_shake_cluster.argtypes = [C.POINTER(TopologyV1), C.c_int, dp, dp, dp, C.c_double, C.c_double, C.c_int]
        L.oracle_rattle_cluster.argtypes = [C.POINTER(TopologyV1), C.c_int, dp, dp, C.c_double, C.c_int]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


def forces(top, x, mask=31, decide_f32=True):
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    F = np.zeros_like(x)
    E = np.zeros(5)
    rc = lib().oracle_forces(C.byref(top.struct()), _dp(x), mask, int(decide_f32), _dp(F), _dp(E))
    if rc != 0:
        raise RuntimeError(f"oracle_forces rc={rc}")
    return F.reshape(-1, 3), E


def pairs(top, x, x_grid, mode):
    x = np.ascontiguousarray(x, np.float32).reshape(-1)
    xg = np.ascontiguousarray(x_grid, np.float32).reshape(-1)
    n = lib().oracle_pairs(C.byref(top.struct()), _fp(x), _fp(xg), mode, None, None, 0)
    if n < 0:
        return int(n)
    i = np.zeros(max(1, n), np.int32)
    j = np.zeros(max(1, n), np.int32)
    lib().oracle_pairs(C.byref(top.struct()), _fp(x), _fp(xg), mode, i.ctypes.data_as(C.POINTER(C.c_int32)),
                       j.ctypes.data_as(C.POINTER(C.c_int32)), n)
    return np.stack([i[:n], j[:n]], axis=1)


def run(top, params, x, v, step0, n_steps, noise=None):
    x = np.array(x, np.float64).reshape(-1)
    v = np.array(v, np.float64).reshape(-1)
    nz = None if noise is None else np.ascontiguousarray(noise, np.float32).reshape(-1)
    work = np.zeros(1)
    rc = lib().oracle_run(C.byref(top.struct()), C.byref(params), _dp(x), _dp(v), step0, n_steps, _fp(nz), _dp(work))
    if rc != 0:
        raise RuntimeError(f"oracle_run rc={rc}")
    return x.reshape(-1, 3), v.reshape(-1, 3), float(work[0])


def energies(top, x, v):
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    v = np.ascontiguousarray(v, np.float64).reshape(-1)
    pe, ke = C.c_double(), C.c_double()
    rc = lib().oracle_energies(C.byref(top.struct()), _dp(x), _dp(v), C.byref(pe), C.byref(ke))
    if rc != 0:
        raise RuntimeError(f"oracle_energies rc={rc}")
    return pe.value, ke.value


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().oracle_philox4x32_10(c, k, o)
    return tuple(o)


def langevin_normals(seed, stream, atom, step, purpose):
    out = np.zeros(3)
    lib().oracle_langevin_normals(seed, stream, atom, step, purpose, _dp(out))
    return out


def schedule(params, step):
    T, g, h = C.c_double(), C.c_double(), C.c_int()
    lib().oracle_schedule(C.byref(params), step, C.byref(T), C.byref(g), C.byref(h))
    return T.value, g.value, h.value


def shake_cluster(top, c, x, xs, v, dt, tol, max_iter=100):
    return lib().oracle_shake_cluster(C.byref(top.struct()), c, _dp(x), _dp(xs), _dp(v), dt, tol, max_iter)


def rattle_cluster(top, c, x, v, tol, max_iter=100):
    return lib().oracle_rattle_cluster(C.byref(top.struct()), c, _dp(x), _dp(v), tol, max_iter)
