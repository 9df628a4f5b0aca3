import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# the batch engine runs one stream per (L, field) class; with fewer hardware
# queues than classes two classes share one and run in turn (set before the
# HIP runtime starts)
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"
for sub in ("spin-glass_amd", "oracle"):
    p = os.path.join(REPO, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libsg's C-ABI)")
    config.addinivalue_line("markers", "slow: long CPU run (statistical parity)")
