import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "build-up-phase_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built librtx.so")
    config.addinivalue_line(
        "markers",
        "process_flags: builds non-opaque or AABB geometry, after which every later traversal in the process takes "
        "the flags variant (and renders the per-vertex wavefront); such tests run last")


def pytest_collection_modifyitems(session, config, items):
    # tests that switch the process to the flags traversal variant run after all others (stable order otherwise)
    items.sort(key=lambda it: it.get_closest_marker("process_flags") is not None)
