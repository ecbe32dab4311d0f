"""Test configuration: puts the package and the oracle on sys.path and
registers the `gpu` marker (tests that need an MI355X).  Under pytest-xdist
each worker runs the oracle on one thread: several workers each spinning a
full OpenMP team on the same cores slow the oracle's barriers by orders of
magnitude."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "volumefluxes.jl_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X, gfx950)")
    if os.environ.get("PYTEST_XDIST_WORKER"):
        import oracle
        oracle.set_threads(1)
