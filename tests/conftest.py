"""Shared test configuration: the `gpu` marker and repo paths."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libfmmh on the device)")
    config.addinivalue_line("markers", "slow: larger CPU oracle cases")
