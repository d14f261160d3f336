import os
import sys

import pytest

# One device for the whole suite, set before any HIP call: the sharded drop-in runs shard g
# on device g % device_count, so on a box exposing several GPUs the multi-shard tests
# would otherwise touch a second one. An existing device list keeps its first entry (HIP
# numbers the devices it was given from 0).
for _var in ("HIP_VISIBLE_DEVICES", "CUDA_VISIBLE_DEVICES"):
    if os.environ.get(_var):
        os.environ[_var] = os.environ[_var].split(",")[0]
if not (os.environ.get("HIP_VISIBLE_DEVICES") or os.environ.get("CUDA_VISIBLE_DEVICES")):
    os.environ["HIP_VISIBLE_DEVICES"] = "0"

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    return oracle_lib.load()
