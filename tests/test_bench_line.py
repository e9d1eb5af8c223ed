"""bench.py's JSON line keeps each leg's result under its own key (no GPU needed): the
single-NAND leg's per-set table reaches the line as a dict beside the CPU baseline."""
import importlib.util
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_legs_keep_their_own_keys(bench):
    single = {"goldilocks": {"gpu_p50_ms": 18.0, "cpu_port_single_thread_ms": 136.0,
                             "published_ms": 107.96}}
    cpu = {"value": 117.0, "unit": "bootstraps/s", "cores": 16, "kind": "port",
           "sample": "..."}
    line = bench.compose_line({"metric": "m", "value": 1.0},
                              {"single_nand": single, "cpu_baseline": cpu})
    assert line["single_nand"] is single and isinstance(line["single_nand"], dict)
    assert line["cpu_baseline"] is cpu
    assert line["serving"] is None and line["metric"] == "m"
    assert list(line)[-len(bench.LEGS):] == list(bench.LEGS)


def test_unknown_leg_is_refused(bench):
    with pytest.raises(KeyError):
        bench.compose_line({}, {"singel_nand": {}})
