import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"
FIXTURES = ["test_1", "test_2", "test_3", "test_4"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def fixture_paths(name: str) -> dict:
    d = GOLDEN / "data" / name
    return {
        "graph": str(d / "graph.totem"),
        "vertex": str(d / "vertex_data" / "graph_vertex_data.totem"),
        "pattern_dir": str(d / "pattern_1"),
        "pattern": str(d / "pattern_1" / "pattern_edge"),
        "pattern_vertex": str(d / "pattern_1" / "pattern_vertex_data"),
    }


@pytest.fixture(scope="session")
def pj():
    import pattenmatching_sharedmemory_amd as m

    return m


@pytest.fixture(scope="session")
def orc():
    import oracle

    return oracle
