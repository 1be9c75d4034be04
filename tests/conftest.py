import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
SYSTEMS = os.path.join(GOLDEN, "systems")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def fixture_system(name):
    """Paths (sys, atm) of a reference check fixture copied under tests/golden/systems."""
    d = os.path.join(SYSTEMS, name)
    return os.path.join(d, "sys-params.txt"), os.path.join(d, "prof-1-atmos-night0.txt")


FIXTURES = ["nx456_nLayers2_wfs6_Nssp10", "nx456_nLayers10_wfs6_Nssp10"]
