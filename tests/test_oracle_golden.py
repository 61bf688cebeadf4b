"""Pins the oracle's composed step() against vectors written by the
reference's own CPUFDTDEngine (tests/golden/gen_reference_vectors.mjs).

The f32 oracle evaluates every expression in double and rounds on store,
exactly like the JavaScript Float32Array code, so the comparison is
bit-exact: any dropped term, wrong sign, swapped index or reordered hook
changes bits.  The f64 oracle must agree to float32 rounding level.
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
VEC = os.path.join(HERE, "golden", "reference_step_vectors.json")


def load_cases():
    with open(VEC) as f:
        return {c["name"]: c for c in json.load(f)["cases"]}


CASES = load_cases()


def build(orc, case, fp64):
    e = orc.Oracle(case["dims"], case["coeffs"], fp64=fp64)
    e.uploadFields(case["volt0"], case["curr0"])
    if "pml" in case:
        e.configurePML(case["pml"])
        e.configureExcitation(case["excitation"])
        e.configureMur(case["mur"])
        e.configureLorentz(case["ade"])
        e.configureTFSF(case["tfsf"])
        e.configureRLC(case["rlc"])
    return e


@pytest.mark.parametrize("name", sorted(CASES))
def test_composed_step_bit_exact(oracle_mod, name):
    case = CASES[name]
    e = build(oracle_mod, case, fp64=False)
    e.iterate(case["steps"])
    f = e.getFields()
    out = case["out"]
    assert e.numTS == out["numTS"]
    np.testing.assert_array_equal(f["volt"], np.array(out["volt"]))
    np.testing.assert_array_equal(f["curr"], np.array(out["curr"]))
    for r, fl in enumerate(out.get("pml_flux", [])):
        v, c = e.pmlFlux(r)
        np.testing.assert_array_equal(v, np.array(fl["volt"]))
        np.testing.assert_array_equal(c, np.array(fl["curr"]))
    k = 0
    for order in out.get("ade", []):
        for d in order:
            s = e.adeState(k)
            for key in ("volt_ADE", "volt_Lor_ADE", "curr_ADE", "curr_Lor_ADE"):
                np.testing.assert_array_equal(s[key], np.array(d[key]), err_msg=f"ade {k} {key}")
            k += 1
    if "rlc" in out:
        s = e.rlcState()
        for key in ("Vdn", "Jn", "VIl"):
            np.testing.assert_array_equal(s[key], np.array(out["rlc"][key]), err_msg=key)


@pytest.mark.parametrize("name", sorted(CASES))
def test_fp64_oracle_tracks_reference(oracle_mod, name):
    case = CASES[name]
    e = build(oracle_mod, case, fp64=True)
    e.iterate(case["steps"])
    f = e.getFields()
    for key in ("volt", "curr"):
        ref = np.array(case["out"][key])
        err = np.linalg.norm(f[key] - ref) / max(np.linalg.norm(ref), 1e-30)
        assert err < 1e-5, (key, err)


def test_planted_mistake_is_caught(oracle_mod):
    """A perturbed coefficient (one ulp-scale change at one cell) must break
    bit-exactness: the comparison above is sensitive enough to see it."""
    case = dict(CASES["all"])
    coeffs = {k: list(v) for k, v in case["coeffs"].items()}
    coeffs["vi"][100] = float(np.float32(coeffs["vi"][100]) * np.float32(1.0000001))
    case["coeffs"] = coeffs
    e = build(oracle_mod, case, fp64=False)
    e.iterate(case["steps"])
    assert not np.array_equal(e.getFields()["volt"], np.array(case["out"]["volt"]))
