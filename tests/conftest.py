"""Test configuration: the `gpu` marker and import paths for the oracle and the product wrapper."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

# Proofs in flight run on streams of their own, which HIP maps onto GPU_MAX_HW_QUEUES hardware
# queues (read once, when the runtime initialises): the tests run with the benchmark's 16.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")
