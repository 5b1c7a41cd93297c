import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


@pytest.fixture(scope="session")
def ctx():
    import rssieve_amd as R
    c = R.Context(0)
    yield c
    c.close()
