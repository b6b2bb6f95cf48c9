"""The C-ABI boundary on the host: every function declared in include/*.h is
exported by the built libraries, the ctypes mirrors of the structs have the C
layout (checked against gcc on the headers), and the host-side argument
checks return the documented error classes.  No device work is started."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from conftest import ROOT

HEADERS = {"prism_md.h": "libprism_md.so", "prism_gen.h": "libprism_gen.so",
           "prism_topology.h": "libprism_gen.so"}
INCLUDE = os.path.join(ROOT, "include")
PKG = os.path.join(ROOT, "prism-q_amd")


def declared_functions(header):
    text = open(os.path.join(INCLUDE, header)).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"#define[^\n]*", " ", text)
    return sorted(set(re.findall(r"\b(prism_(?:md|gen|topo)_\w+)\s*\(", text)))


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_every_declared_symbol_is_exported(header):
    names = declared_functions(header)
    assert len(names) >= 5
    lib = C.CDLL(os.path.join(PKG, HEADERS[header]))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    nm = subprocess.run(["nm", "-D", "--defined-only", os.path.join(PKG, HEADERS[header])], capture_output=True,
                        text=True, check=True).stdout
    exported = set(re.findall(r"\s[TW]\s+(prism_\w+)", nm))
    assert set(names) <= exported


def _c_layout(struct_fields):
    """sizeof / offsetof of the C structs, from gcc on the public headers."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "prism_md.h"', '#include "prism_gen.h"',
             '#include "prism_topology.h"', "int main(void) {"]
    for s, fields in struct_fields.items():
        lines.append(f'printf("{s} %zu\\n", sizeof({s}));')
        for f in fields:
            lines.append(f'printf("{s}.{f} %zu\\n", offsetof({s}, {f}));')
    lines += ["return 0;", "}"]
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "l.c"), os.path.join(d, "l")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-std=c99", "-I", INCLUDE, src, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    return {k: int(v) for k, v in (ln.split() for ln in out.splitlines())}


def test_struct_layouts_match_ctypes(pm):
    structs = {"PrismMdTopologyV1": pm.TopologyV1, "PrismMdParamsV1": pm.ParamsV1, "PrismMdStatsV1": pm.StatsV1,
               "PrismMdSlotCapsV1": pm.SlotCapsV1,
               "PrismGenParamsV1": pm.GenParamsV1, "PrismGenStatsV1": pm.GenStatsV1,
               "PrismTopoLoadOptionsV1": pm.TopoLoadOptionsV1, "PrismTopoInfoV1": pm.TopoInfoV1}
    fields = {s: [f for f, _ in t._fields_] for s, t in structs.items()}
    c = _c_layout(fields)
    for s, t in structs.items():
        assert C.sizeof(t) == c[s], s
        for f, _ in t._fields_:
            assert getattr(t, f).offset == c[f"{s}.{f}"], (s, f)


def test_defaults_and_version(pm):
    lib = pm.md_lib()
    assert lib.prism_md_abi_version() == 2
    p = pm.default_params()
    assert (p.dt_ps, p.temperature_K, p.cutoff_A, p.skin_A, p.rebuild_every) == (0.004, 300.0, 10.0, 2.0, 20)
    assert (p.gamma_ps, p.equilibration_steps, p.equilibration_gamma_ps) == (10.0, 10000, 1000.0)
    assert p.com_remove_every == 100 and p.constraint_max_iter == 100 and p.resort_every == 8


def _create(pm, top, params):
    lib = pm.md_lib()
    h = C.c_void_p()
    rc = lib.prism_md_create(C.byref(top.struct()), C.byref(params), None, C.byref(h))
    return rc, h, lib.prism_md_last_error(None).decode()


def test_invalid_inputs_rejected_before_device(pm, small_top):
    bad = pm.default_params(dt_ps=-1.0)
    rc, h, msg = _create(pm, small_top, bad)
    assert rc == pm.E_INVALID and not h.value and "param" in msg
    for chunk in (4, 33, -1):  # row tiles per nonbonded unit: 0 (default) or 8..32
        rc, h, msg = _create(pm, small_top, pm.default_params(chunk_tiles=chunk))
        assert rc == pm.E_INVALID and not h.value and "param" in msg, chunk
    t = small_top.with_positions(small_top["positions"].copy())
    t.a["positions"][5] = float("nan")
    rc, h, msg = _create(pm, t, pm.default_params())
    assert rc == pm.E_INVALID and "position" in msg
    t = small_top.with_positions(small_top["positions"])
    t.a = dict(t.a)
    t.a["bond_atoms"] = t.a["bond_atoms"].copy()
    t.a["bond_atoms"][0] = small_top.n_atoms + 3
    t._struct = None
    rc, h, msg = _create(pm, t, pm.default_params())
    assert rc == pm.E_INVALID and "bond" in msg


def test_null_arguments(pm, small_top):
    lib = pm.md_lib()
    h = C.c_void_p()
    assert lib.prism_md_create(None, None, None, C.byref(h)) == pm.E_INVALID
    assert lib.prism_md_destroy(None) == pm.OK  # like free(NULL)
    assert lib.prism_md_step(None, 1, None) == pm.E_INVALID
    assert lib.prism_md_compute_forces(None, 31, None, None) == pm.E_INVALID


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a HIP device is visible")
def test_no_device_is_an_error_not_a_fallback(pm, small_top):
    rc, h, msg = _create(pm, small_top, pm.default_params())
    assert rc == pm.E_NODEVICE and not h.value
