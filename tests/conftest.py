import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ceres-raytracer_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libcrt_hip.so")


# The parity tests print what they measured (agreement, classified
# differences, errors).  Under output capture (the default, also with -q)
# those lines would be lost with a passing test; they are collected here and
# printed in the terminal summary.
_FIGURES = []


def pytest_runtest_logreport(report):
    if report.when == "call" and report.passed and report.capstdout:
        for line in report.capstdout.splitlines():
            if line.strip():
                _FIGURES.append((report.nodeid.split("::", 1)[-1], line.strip()))


def pytest_terminal_summary(terminalreporter):
    if _FIGURES:
        terminalreporter.section("measured figures (stdout of passed tests)")
        for test, line in _FIGURES:
            terminalreporter.write_line("%s: %s" % (test, line))
