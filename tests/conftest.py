import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# The oracle runs OpenMP over every core.  Under pytest-xdist each worker
# gets its share of the cores instead: n workers with all cores each would
# leave n x cores spinning threads on the machine (the libgomp default waits
# by spinning), which slows the suite down many times over.
_workers = int(os.environ.get("PYTEST_XDIST_WORKER_COUNT", "1") or 1)
if _workers > 1 and "OMP_NUM_THREADS" not in os.environ:
    os.environ["OMP_NUM_THREADS"] = str(max(1, (os.cpu_count() or 1) // _workers))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libpizza_amd.so)")
