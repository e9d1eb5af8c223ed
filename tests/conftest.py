import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def hb():
    import hasteboots_amd
    return hasteboots_amd


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


def stream(seed: int, tag: str) -> np.random.Generator:
    """Independent numpy generator per (seed, purpose)."""
    h = int.from_bytes(tag.encode()[:8].ljust(8, b"\0"), "little")
    return np.random.default_rng([seed, h])


def uniform_field(rng: np.random.Generator, q: int, shape) -> np.ndarray:
    """Uniform canonical words in [0, q)."""
    return rng.integers(0, q, size=shape, dtype=np.uint64, endpoint=False)


_KEYS = {}


@pytest.fixture(scope="session")
def keyset(hb):
    """(params, sk, bsk, ksk) per parameter set from seed 0, generated on the host once."""
    def get(set_id: int, seed: int = 0):
        k = (set_id, seed)
        if k not in _KEYS:
            P = hb.Parameters.of(set_id)
            sk = hb.KeyGen.generate_secret_key(P, seed=seed)
            bsk, ksk = sk.generate_keys_host()
            _KEYS[k] = (P, sk, bsk, ksk)
        return _KEYS[k]
    return get


def gpu_available() -> bool:
    try:
        import hasteboots_amd
        return hasteboots_amd.device_count() > 0
    except Exception:
        return False
