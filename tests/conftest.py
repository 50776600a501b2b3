import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

POT = os.path.join(ROOT, "tests", "golden", "potentials")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "reference: needs the reference library built from its sources")


def pot_files(name, types):
    return {t: os.path.join(POT, "%s.%s.nn" % (name, t)) for t in types}


TI_O = {"Ti": os.path.join(POT, "Ti.15t-15t.nn"), "O": os.path.join(POT, "O.15t-15t.nn")}

# Ti2O4 of python3/aenet/tests/test_ase.py:31-43
TI2O4_POS = np.array([(0.42651082, 2.82851745, 1.22027956), (1.76480257, -0.94296447, 3.66083922),
                      (0.14225662, 0.94278546, 0.79509432), (2.04905661, 0.94276737, 4.08602446),
                      (-0.14223328, -0.94276538, 3.23565398), (2.33354646, 2.82831821, 1.64546480)])
TI2O4_CELL = np.array([(3.81382872, 0.0, 0.0), (0.56886601, 3.77116440, 0.0),
                       (-2.19138173, -1.88561177, 4.88111931)])
TI2O4_TYPES = np.array([1, 1, 2, 2, 2, 2], np.int32)


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def device():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("the gpu tests need a GPU (run with -m 'not gpu' on machines without one)")
    return torch.device("cuda:0")
