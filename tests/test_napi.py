"""The Node (N-API) binding: loads and exposes the WebGPUEngine surface
(CPU), and runs the reference step vectors on the GPU through JavaScript."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from cases import golden_cases, rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
ADDON = os.path.join(ROOT, "efs_amd", "js", "efs_engine.node")

needs_node = pytest.mark.skipif(NODE is None or not os.path.exists(ADDON), reason="node or the addon is missing")


@needs_node
def test_js_binding_surface():
    out = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "check_api.mjs")], capture_output=True, text=True,
                         timeout=120, cwd=ROOT)
    assert out.returncode == 0, out.stderr
    info = json.loads(out.stdout.strip().splitlines()[-1])
    assert info["missing"] == []
    assert info["version"].startswith("efs-mi355x")
    if info["devices"] == 0:
        assert info["initGPU"] is False  # the reference resolves false without a GPU


def _jsonable(case):
    def conv(v):
        if isinstance(v, np.ndarray):
            return v.tolist()
        if isinstance(v, dict):
            return {k: conv(x) for k, x in v.items()}
        if isinstance(v, list):
            return [conv(x) for x in v]
        if isinstance(v, (np.integer,)):
            return int(v)
        if isinstance(v, (np.floating,)):
            return float(v)
        return v
    return conv(case)


@needs_node
@pytest.mark.gpu
@pytest.mark.parametrize("name", ["core", "all"])
def test_js_binding_reference_vectors(tmp_path, name):
    case = dict(golden_cases()[name])
    out = case.pop("out")
    src, dst = tmp_path / "case.json", tmp_path / "out.json"
    src.write_text(json.dumps(_jsonable(case)))
    print("node case", name, flush=True)
    res = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "run_case.mjs"), str(src), str(dst)],
                         capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert res.returncode == 0, res.stderr
    got = json.loads(dst.read_text())
    assert got["numTS"] == out["numTS"]
    for k in ("volt", "curr"):
        assert rel_l2(got[k], out[k]) <= 1e-5, k
