"""Shared pytest setup: the `gpu` marker and import paths.

`-m "not gpu"` runs the oracle, host-logic and ABI tests on any machine;
`-m gpu` runs the parity tests that call the HIP library on an MI355X.
"""
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", "pacmap-rs_amd"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def load_golden(name):
    with open(os.path.join(GOLDEN, name), encoding="utf-8") as fh:
        return json.load(fh)


@pytest.fixture(scope="session")
def golden():
    return load_golden
