import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X)")


def golden(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def gpu_device():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")
    return torch.device("cuda:0")


@pytest.fixture
def knobs():
    """Setter for the library's launch-tuning knobs (lbm_set_tuning): call
    knobs(two_step=0, ...); every knob set is restored at teardown."""
    from lettuce_amd import _native as N
    touched = set()

    def set_knobs(**kw):
        for name, value in kw.items():
            N.check(N.lib().lbm_set_tuning(name.encode(), int(value)), "lbm_set_tuning")
            touched.add(name)
    yield set_knobs
    for name in touched:
        N.lib().lbm_set_tuning(name.encode(), -1)


def pytest_sessionfinish(session, exitstatus):
    """with LBM_ERR_LOG set, the max-normalised and entrywise errors of every
    rel_err comparison of the parity tests, one JSON line each"""
    path = os.environ.get("LBM_ERR_LOG")
    if not path:
        return
    import json
    # pytest imports the test file under its own name, other tests import
    # it as tests.test_parity_gpu: two module objects, two logs
    with open(path, "w") as fh:
        for name in ("test_parity_gpu", "tests.test_parity_gpu"):
            mod = sys.modules.get(name)
            for test, rel, elem in getattr(mod, "ERR_LOG", []):
                fh.write(json.dumps({"test": test, "rel": rel, "elem": elem}) + "\n")
