"""Shared test set-up: import paths, the ``gpu`` marker, golden fixtures."""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs through librt6d.so")


def golden_index():
    with open(os.path.join(GOLDEN_DIR, "index.json")) as f:
        return json.load(f)


def golden_case(entry):
    """Rebuild the inputs of a golden entry and attach its dt."""
    from raytracer_3d_amd import cases
    kw = dict(entry["kwargs"])
    kw["n"] = tuple(kw["n"])
    if "fixed" in kw:
        kw["fixed"] = tuple(kw["fixed"])
    c = getattr(cases, entry["ctor"])(**kw)
    c.dt = entry["dt"]
    return c


def golden_data(entry):
    with np.load(os.path.join(GOLDEN_DIR, entry["stem"] + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    if nb == 0.0:
        return float(np.linalg.norm(a))
    return float(np.linalg.norm(a - b) / nb)


@pytest.fixture(scope="session")
def gpu_available():
    from raytracer_3d_amd import rt6d
    try:
        n = rt6d.device_count()
    except RuntimeError as e:
        pytest.fail(f"librt6d.so is not loadable on the GPU host: {e}")
    if n < 1:
        pytest.fail("no HIP device visible to librt6d.so")
    return True
