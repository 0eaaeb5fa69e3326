import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "user-eph_amd"),
          os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


@pytest.fixture(scope="session")
def oracle():
    import eph_oracle
    eph_oracle.lib()
    return eph_oracle


# Figures the parity tests measure (errors against the oracle per workload),
# printed after the session's summary so that they stay in a run's tail.
_REPORT = []


@pytest.fixture(scope="session")
def parity_report():
    return _REPORT


def pytest_terminal_summary(terminalreporter):
    if _REPORT:
        terminalreporter.write_sep("-", "parity figures")
        for line in _REPORT:
            terminalreporter.write_line(line)
