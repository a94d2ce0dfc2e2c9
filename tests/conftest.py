import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "phasefield-precipitate-aging_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (MI355X) and the built libpfpa.so")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def thermo_golden():
    import numpy as np
    return dict(np.load(os.path.join(ROOT, "tests", "golden", "thermo_golden.npz")))


@pytest.fixture(scope="session")
def step_golden():
    import numpy as np
    return dict(np.load(os.path.join(ROOT, "tests", "golden", "step_golden.npz")))


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    import pfpa
    pfpa.lib()
    return torch
