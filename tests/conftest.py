"""Shared test setup.

Markers: ``gpu`` tests need an MI355X (they call the HIP library through its
C ABI); everything else runs on the CPU.  The oracle (oracle/oracle.py) is the
checker in both tiers.
"""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ccpi-regularisation-toolkit_amd")
for p in (PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X); calls the HIP library")


def _gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


class Golden:
    """The reference fixtures written by tests/golden/make_golden.py."""

    def __init__(self):
        with open(os.path.join(GOLDEN_DIR, "ref_cases.json")) as fh:
            self.cases = json.load(fh)
        self.arrays = dict(np.load(os.path.join(GOLDEN_DIR, "ref_cases.npz")))

    def input(self, case):
        return self.arrays["in_" + case["input"]]

    def params(self, case):
        """The case's parameters; FGP-dTV cases also get their reference image as ``ref``."""
        p = dict(case["params"])
        k = case.get("ref_input")
        if k is not None:
            p["ref"] = self.arrays["in_" + k]
        return p

    def out_init(self, case):
        k = case.get("out_init")
        return None if k is None else self.arrays["in_" + k].copy()

    def expected(self, case):
        return self.arrays[f"c{case['id']}_out"], self.arrays[f"c{case['id']}_info"]


_golden = None


def golden():
    global _golden
    if _golden is None:
        _golden = Golden()
    return _golden


def case_id(case):
    p = case["params"]
    extra = ",".join(f"{k}={v}" for k, v in p.items() if k not in ("iters",))
    return f"c{case['id']}-{case['method']}-{case['input']}-{p['iters']}it-{extra}"


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
