import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built librbh.so")


@pytest.fixture(scope="session")
def oracle_bank32():
    """Shapes with oracle-built SDFs (CPU), resolution 32."""
    from oracle.sdf import create_signed_distance
    from rainbow_amd import scenes
    return scenes.ShapeBank(resolution=32, sdf_builder=create_signed_distance)
