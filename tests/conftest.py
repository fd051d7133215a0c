import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "ref: needs the compiled reference bridge oracle/_ref (built where the reference tree exists)")


def pytest_collection_modifyitems(config, items):
    from oracle_lib import ref_available
    skip_ref = pytest.mark.skip(reason="oracle/_ref/libzimgref.so not built (reference tree absent)")
    for item in items:
        if "ref" in item.keywords and not ref_available():
            item.add_marker(skip_ref)
