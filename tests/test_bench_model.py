"""The benchmark's algorithmic byte count per sweep launch (bench.py):
FP32 fields 60 B per cell; FP16 fields 9 x 2 B plus 6 coefficients, at 2 B
when they are indexed (byte or 16-bit indices) and 4 B as FP32 arrays;
60 B per UPML cell on top in every form."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_sweep_bytes_per_form():
    n, p = 1000, 100
    assert bench.sweep_bytes(n, p, "fp32", ("coef_idx",)) == 60 * n + 60 * p
    assert bench.sweep_bytes(n, p, "fp32") == 60 * n + 60 * p
    assert bench.sweep_bytes(n, p, "fp16", ("coef_idx",)) == 30 * n + 60 * p
    assert bench.sweep_bytes(n, p, "fp16", ("coef_idx16", "f16")) == 30 * n + 60 * p
    assert bench.sweep_bytes(n, p, "fp16", ("f16", "fused")) == 42 * n + 60 * p


def test_byte_model_names_the_form():
    assert bench.byte_model("fp32").startswith("60 B/cell")
    assert bench.byte_model("fp16", ("coef_idx",)).startswith("30 B/cell")
    assert bench.byte_model("fp16", ()).startswith("42 B/cell")
