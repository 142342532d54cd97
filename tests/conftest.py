"""Shared pytest setup: the `gpu` marker and paths to the built libraries."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "build", "liboracle.so")
ORC_CLI = os.path.join(ORACLE_DIR, "build", "orc_cli")
GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


@pytest.fixture(scope="session")
def oracle_built():
    """Build the oracle (test infrastructure) if it is missing."""
    if not (os.path.exists(ORACLE_LIB) and os.path.exists(ORC_CLI)):
        subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
    return ORACLE_LIB
