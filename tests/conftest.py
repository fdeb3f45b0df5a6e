import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests through the C ABI")
    # Build what is missing (the GPU box receives the libraries already built).
    for folder, lib in (("oracle", "liboracle.so"), ("rosy_amd", "librosy_amd.so")):
        if not os.path.exists(os.path.join(ROOT, folder, lib)):
            subprocess.run(["make", "-j8", "-C", os.path.join(ROOT, folder)], check=True)
