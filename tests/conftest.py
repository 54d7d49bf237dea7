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


def decaying_density(basis, seed, scale=1.0, width=0.5):
    """A symmetric density-like matrix whose elements fall off with the
    distance r of their two AO centres, P_ij = scale u_ij exp(-width r^2),
    u uniform in (-1, 1) from PCG64(seed): blocks of distant shells are small,
    as in a molecular density, so the density test screens realistically.
    Reproducible from numpy alone."""
    import numpy as np
    off = basis.shell_offsets
    cen = np.zeros((basis.nbf, 3))
    for s, sh in enumerate(basis.shells):
        cen[off[s]:off[s + 1]] = sh.center
    r2 = np.sum((cen[:, None, :] - cen[None, :, :]) ** 2, axis=2)
    rng = np.random.Generator(np.random.PCG64(seed))
    A = rng.uniform(-1.0, 1.0, (basis.nbf, basis.nbf))
    return scale * 0.5 * (A + A.T) * np.exp(-width * r2)
