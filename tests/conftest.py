import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "vumat-pd_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")


# BASELINE material: Johnson-Cook A=350 B=275 n=0.36 C=0.022 m=1, rate0 = 1/s = 1e-3/ms,
# Troom 293 K, Tmelt 1793 K, E=200 GPa, nu=0.3, rho 7850 kg/m3, beta 0.9, cp 500, workToHeat 1.
JC_PROPS = [200000.0, 0.3, 350.0, 275.0, 0.022, 0.36, 1.0, 1.0e-3, 293.0, 1793.0, 7850.0, 0.9,
            500.0, 1.0]
# The reference's configs/johnson_cook.json props (B=500, C=0.025, E=210 GPa, rate0=1/ms).
JC_PROPS_REFERENCE = [210000.0, 0.30, 350.0, 500.0, 0.025, 0.36, 1.0, 1.0, 293.0, 1793.0, 7800.0,
                      0.9, 500.0, 1.0]


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def ref(orc):
    if not orc.ref_available():
        pytest.skip("oracle/_ref is not built here (needs the reference sources)")
    return orc
