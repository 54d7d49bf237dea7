import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (REPO, os.path.join(REPO, "qdk-chemistry_amd"), os.path.join(REPO, "qdk-chemistry_amd", "csrc")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def water():
    from qdk_chemistry_amd.molecule import Structure
    # examples/data/water.structure.xyz of the reference
    return Structure.from_angstrom(["O", "H", "H"], [[0.0, 0.0, 0.0], [0.758602, 0.0, 0.504284],
                                                     [-0.758602, 0.0, 0.504284]])


def random_density(nbf, seed, scale=1.0):
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(seed))
    A = rng.uniform(-1.0, 1.0, (nbf, nbf)) * scale
    return 0.5 * (A + A.T)
