import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "ref: needs the reference checker library built from the reference sources")


def pytest_collection_modifyitems(config, items):
    from oracle import oracle as O

    skip_ref = pytest.mark.skip(reason="reference checker library not built here (oracle/_ref)")
    for item in items:
        if "ref" in item.keywords and not O.ref_available():
            item.add_marker(skip_ref)
