import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def ctx():
    import falcondb_amd as fa
    c = fa.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def ctx_part():
    """A context that takes the radix-partitioned join for every probe size."""
    import falcondb_amd as fa
    c = fa.Context(0)
    c.set_tuning("join_part_min_rows", 1)
    yield c
    c.close()


@pytest.fixture(scope="session")
def ctx_part_compact():
    """The radix-partitioned join for every probe size, with compact record
    keys wherever the build key range allows them (payload records too)."""
    import falcondb_amd as fa
    c = fa.Context(0)
    c.set_tuning("join_part_min_rows", 1)
    c.set_tuning("join_compact", 2)
    yield c
    c.close()


@pytest.fixture(scope="session")
def ctx_part_hash64():
    """The radix-partitioned join for every probe size, always with 8-byte
    hash64 records (the compact 4-byte record keys switched off)."""
    import falcondb_amd as fa
    c = fa.Context(0)
    c.set_tuning("join_part_min_rows", 1)
    c.set_tuning("join_compact", 0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def ctx_direct():
    """A context that always probes the open-addressing index directly."""
    import falcondb_amd as fa
    c = fa.Context(0)
    c.set_tuning("join_partition", 0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def ctx_gbp():
    """A context that takes the radix-partitioned GROUP BY for every size."""
    import falcondb_amd as fa
    c = fa.Context(0)
    c.set_tuning("gb_part_min_rows", 1)
    c.mode = "partitioned"
    yield c
    c.close()


@pytest.fixture(scope="session")
def ctx_gbp_small():
    """Partitioned GROUP BY with an 8-bucket hot-key cache: most rows take the
    cold path, and inserts race for the few ways."""
    import falcondb_amd as fa
    c = fa.Context(0)
    c.set_tuning("gb_part_min_rows", 1)
    c.set_tuning("gbp_cache_slots", 64)
    c.mode = "partitioned"
    yield c
    c.close()


@pytest.fixture(scope="session")
def ctx_gbh():
    """A context that keeps high-cardinality GROUP BY on the single hash table."""
    import falcondb_amd as fa
    c = fa.Context(0)
    c.set_tuning("gb_partition", 0)
    c.mode = "hash"
    yield c
    c.close()
