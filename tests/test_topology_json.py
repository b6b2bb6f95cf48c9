"""PrismPrepTopology JSON ingest (prism_topo_*, include/prism_topology.h), on
the CPU.

The reader restates PrismPrepTopology::load / validate / the harmonic lowering
and HMR (crates/prism-nhs/src/input.rs:300-323, :467-538, :1934-2300,
:2494-2565).  Pinned by: the FIPS 180-2 SHA-256 examples and Python's hashlib
(the digest), Python's json.dumps(sort_keys, compact) (the canonical contract
text the reference hashes, python_canonical_json input.rs:44-110), struct
packing of the exception and strict-exclusion rows (input.rs:1934-1968), a
round trip of generated systems whose oracle forces must not change, and the
reference's refusal messages for each malformed input."""
from __future__ import annotations

import copy
import ctypes as C
import hashlib
import json
import struct

import numpy as np
import pytest


@pytest.fixture(scope="module")
def gen_top(pm):
    return pm.Topology.generate(700, seed=21)


@pytest.fixture(scope="module")
def doc(gen_top, tmp_path_factory):
    path = tmp_path_factory.mktemp("topo") / "gen.json"
    gen_top.to_json(str(path), hmr_applied=True)
    with open(path) as f:
        return json.load(f)


def load_doc(pm, d, tmp_path, name="t.json", **kw):
    p = tmp_path / name
    with open(p, "w") as f:
        json.dump(d, f)
    return pm.Topology.from_json(str(p), **kw)


def refused(pm, d, tmp_path, match, **kw):
    with pytest.raises(pm.PrismMdError) as ex:
        load_doc(pm, d, tmp_path, **kw)
    assert ex.value.rc == pm.E_INVALID
    assert match in str(ex.value), str(ex.value)


def sha_c(pm, data: bytes) -> str:
    out = C.create_string_buffer(65)
    assert pm.gen_lib().prism_topo_sha256_hex(data, len(data), out) == 0
    return out.value.decode()


def test_sha256_known_answers(pm):
    # FIPS 180-2 appendix B examples
    assert sha_c(pm, b"abc") == "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"
    assert sha_c(pm, b"") == "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855"
    assert sha_c(pm, b"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq") == \
        "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1"
    rng = np.random.default_rng(3)
    for n in (1, 55, 56, 63, 64, 65, 119, 1000, 4097):
        b = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert sha_c(pm, b) == hashlib.sha256(b).hexdigest()


def test_written_contracts_match_python_digests(doc):
    hc = dict(doc["harmonic_parameter_contract"])
    digest = hc.pop("contract_sha256")
    text = json.dumps(hc, sort_keys=True, separators=(",", ":"))
    assert digest == hashlib.sha256(text.encode()).hexdigest()
    xc = doc["nonbonded_exception_contract"]
    rows = sorted(doc["nonbonded_exceptions"], key=lambda r: (r["i"], r["j"]))
    h = hashlib.sha256(b"PRISMOS_NBEX_ACTIVE_V1\0" + struct.pack("<Q", len(rows)))
    for r in rows:
        h.update(struct.pack("<QQddd", r["i"], r["j"], r["charge_product_e2"], r["sigma_A"], r["epsilon_kcal_mol"]))
    assert xc["pair_digest_sha256"] == h.hexdigest()
    strict = sorted({(min(i, j), max(i, j)) for i, row in enumerate(doc["exclusions"]) for j in row})
    h = hashlib.sha256(b"PRISMOS_NBEX_STRICT_V1\0" + struct.pack("<Q", len(strict)))
    for i, j in strict:
        h.update(struct.pack("<QQ", i, j))
    assert xc["strict_exclusion_digest_sha256"] == h.hexdigest()
    assert xc["strict_exclusion_pair_count"] == len(strict)
    assert not {(r["i"], r["j"]) for r in rows} & set(strict)


def test_round_trip_keeps_the_system(pm, oracle, gen_top, doc, tmp_path):
    top = load_doc(pm, doc, tmp_path)
    assert top.json_info["harmonic_convention"] == 1 and top.json_info["hmr_in_file"] == 1
    assert top.json_info["hmr_applied_here"] == 0
    for k, a in gen_top.a.items():
        if k.startswith("exclusion"):
            continue
        assert np.array_equal(a, top.a[k]), k
    # strict exclusions = generator exclusions minus exception partners
    exc = {tuple(sorted(p)) for p in gen_top["exception_atoms"].reshape(-1, 2).tolist()}
    off, at = gen_top["exclusion_offsets"], gen_top["exclusion_atoms"]
    want = [[int(j) for j in at[off[i]:off[i + 1]] if tuple(sorted((i, int(j)))) not in exc]
            for i in range(gen_top.n_atoms)]
    off2, at2 = top["exclusion_offsets"], top["exclusion_atoms"]
    got = [at2[off2[i]:off2[i + 1]].tolist() for i in range(top.n_atoms)]
    assert got == want
    # the step sees the same system: identical FP64 oracle forces and energies
    x = gen_top["positions"].reshape(-1, 3)
    F0, E0 = oracle.forces(gen_top, x)
    F1, E1 = oracle.forces(top, x)
    assert np.array_equal(F0, F1) and np.array_equal(E0, E1)


def test_missing_harmonic_contract(pm, gen_top, doc, tmp_path):
    d = copy.deepcopy(doc)
    del d["harmonic_parameter_contract"]
    refused(pm, d, tmp_path, "harmonic source convention is absent")
    amber = load_doc(pm, d, tmp_path, convention="amber")
    assert np.array_equal(amber["bond_k"], gen_top["bond_k"])
    openmm = load_doc(pm, d, tmp_path, convention="openmm")  # U = k d^2 / 2 -> halved
    assert np.array_equal(openmm["bond_k"], (gen_top["bond_k"].astype(np.float64) * 0.5).astype(np.float32))
    assert np.array_equal(openmm["angle_k"], (gen_top["angle_k"].astype(np.float64) * 0.5).astype(np.float32))


def test_harmonic_contract_refusals(pm, doc, tmp_path):
    d = copy.deepcopy(doc)
    d["harmonic_parameter_contract"]["bond_count"] += 1
    refused(pm, d, tmp_path, "unsupported schema, units, provider or term counts")
    d = copy.deepcopy(doc)
    d["harmonic_parameter_contract"]["source_provider"] = "someone else"
    refused(pm, d, tmp_path, "canonical digest mismatch")
    d = copy.deepcopy(doc)
    d["harmonic_parameter_contract"]["extra"] = 1
    refused(pm, d, tmp_path, "unknown field")
    # an OpenMM contract with a correct digest halves the coefficients
    d = copy.deepcopy(doc)
    hc = d["harmonic_parameter_contract"]
    hc["bond_convention"] = hc["angle_convention"] = "OPENMM_FULL_K_HALF_ENERGY"
    body = {k: v for k, v in hc.items() if k != "contract_sha256"}
    hc["contract_sha256"] = hashlib.sha256(json.dumps(body, sort_keys=True, separators=(",", ":")).encode()).hexdigest()
    t = load_doc(pm, d, tmp_path)
    assert t.json_info["harmonic_convention"] == 2
    assert np.allclose(t["bond_k"], np.asarray([b["k"] for b in doc["bonds"]]) * 0.5, rtol=1e-7)
    refused(pm, d, tmp_path, "disagrees with the file's contract", convention="amber")


def test_term_and_exclusion_refusals(pm, doc, tmp_path):
    cases = []
    d = copy.deepcopy(doc); d["bonds"][3]["j"] = d["bonds"][3]["i"]
    cases.append((d, "invalid bond term at index 3"))
    d = copy.deepcopy(doc); d["angles"][5]["theta0"] = 3.5
    cases.append((d, "invalid angle term at index 5"))
    d = copy.deepcopy(doc); d["dihedrals"][2]["periodicity"] = 0
    cases.append((d, "invalid dihedral term at index 2"))
    d = copy.deepcopy(doc); d["masses"][7] = 0.0
    cases.append((d, "mass value 7 must be finite and positive"))
    d = copy.deepcopy(doc); d["lj_params"][4]["sigma"] = -1.0
    cases.append((d, "invalid LJ parameters for atom 4"))
    d = copy.deepcopy(doc); d["positions"] = d["positions"][:-1]
    cases.append((d, "positions length"))
    i = next(k for k, row in enumerate(doc["exclusions"]) if row)
    j = doc["exclusions"][i][0]
    d = copy.deepcopy(doc); d["exclusions"][j].remove(i)
    cases.append((d, "asymmetric strict exclusion"))
    d = copy.deepcopy(doc); d["exclusions"][i].append(i)
    cases.append((d, "invalid, self, or duplicate strict exclusion"))
    for d, msg in cases:
        refused(pm, d, tmp_path, msg)


def test_exception_refusals(pm, doc, tmp_path):
    e0 = doc["nonbonded_exceptions"][0]
    d = copy.deepcopy(doc); d["nonbonded_exceptions"][0]["i"], d["nonbonded_exceptions"][0]["j"] = e0["j"], e0["i"]
    refused(pm, d, tmp_path, "must be canonical i < j")
    d = copy.deepcopy(doc); d["nonbonded_exceptions"].append(dict(e0))
    refused(pm, d, tmp_path, "duplicate nonbonded exception pair")
    d = copy.deepcopy(doc); d["nonbonded_exceptions"][0]["charge_product_e2"] += 0.25
    refused(pm, d, tmp_path, "nonbonded exception digest mismatch")
    d = copy.deepcopy(doc); d["nonbonded_exceptions"][0]["sigma_A"] = 1e-50
    refused(pm, d, tmp_path, "not safely representable as f32")
    d = copy.deepcopy(doc); d["exclusions"][e0["i"]].append(e0["j"]); d["exclusions"][e0["j"]].append(e0["i"])
    refused(pm, d, tmp_path, "overlaps strict exclusion")
    d = copy.deepcopy(doc); del d["nonbonded_exception_contract"]
    refused(pm, d, tmp_path, "require a versioned completeness/digest contract")
    d = copy.deepcopy(doc); d["nonbonded_exception_contract"]["fallback_used"] = True
    refused(pm, d, tmp_path, "incomplete or records fallback")
    d = copy.deepcopy(doc); d["generated_prmtop_sha256"] = "xyz"
    refused(pm, d, tmp_path, "source prmtop hash")


def test_malformed_json(pm, tmp_path):
    p = tmp_path / "bad.json"
    p.write_text('{"n_atoms": 3, "positions": [1, 2,, 3]}')
    with pytest.raises(pm.PrismMdError) as ex:
        pm.Topology.from_json(str(p))
    assert "Failed to parse topology JSON" in str(ex.value)
    with pytest.raises(pm.PrismMdError) as ex:
        pm.Topology.from_json(str(tmp_path / "absent.json"))
    assert "Failed to read topology file" in str(ex.value)


def test_hmr_on_load(pm, tmp_path):
    """apply_hmr semantics (input.rs:2494-2565): a file without a verified
    hmr_applied = true is repartitioned on load (H x3, the heavy atom pays);
    a verified one is left alone."""
    plain = pm.Topology.generate(500, seed=9, hmr_factor=1.0)
    hmr = pm.Topology.generate(500, seed=9)
    for flag in (None, False):
        p = tmp_path / f"plain_{flag}.json"
        plain.to_json(str(p), hmr_applied=flag)
        t = pm.Topology.from_json(str(p))
        assert t.json_info["hmr_in_file"] == (-1 if flag is None else 0)
        assert t.json_info["hmr_applied_here"] == len(plain["cluster_central"])
        assert t.json_info["recommended_timestep_fs"] == 4.0
        assert np.allclose(t["masses"], hmr["masses"], rtol=2e-7, atol=0)
        assert np.allclose(t["cluster_inv_mass_h"], hmr["cluster_inv_mass_h"], rtol=2e-7)
        assert np.allclose(t["cluster_inv_mass_central"], hmr["cluster_inv_mass_central"], rtol=2e-7)
        kept = pm.Topology.from_json(str(p), apply_hmr=False)
        assert np.array_equal(kept["masses"], plain["masses"])
    p = tmp_path / "hmr_true.json"
    plain.to_json(str(p), hmr_applied=True)
    t = pm.Topology.from_json(str(p))
    assert t.json_info["hmr_applied_here"] == 0 and np.array_equal(t["masses"], plain["masses"])
