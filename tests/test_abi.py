"""CPU checks of the C ABI: both libraries load, every function declared in
include/*.h is exported, and device entry points fail cleanly (a status,
never an exit) when no GPU is visible."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

from hetegeneous_parallel_programming_amd import _native as nat
from hetegeneous_parallel_programming_amd.solver import Context, TCSolver, device_lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions(header: str) -> set:
    text = open(os.path.join(INCLUDE, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"^\s*[\w\s\*]*?\b(tch?_\w+)\s*\(", text, flags=re.M))
    return {n for n in names if not n.endswith("_t")}


def test_headers_declare_the_expected_surface():
    host = declared_functions("tc_host.h")
    dev = declared_functions("tc.h")
    assert {"tch_gen_rmat", "tch_orient", "tch_partition_range", "tch_read_graphminer"} <= host
    assert {"tc_ctx_create", "tc_graph_upload", "tc_count", "tc_count_plain", "tc_count_binned",
            "tc_partition_upload", "tc_batch_create", "tc_batch_count"} <= dev


@pytest.mark.parametrize("header,libpath", [("tc_host.h", nat.HOST_LIB), ("tc.h", nat.HIP_LIB)])
def test_every_declared_symbol_is_exported(header, libpath):
    nat.host()
    lib = ctypes.CDLL(libpath)
    missing = [n for n in sorted(declared_functions(header)) if not hasattr(lib, n)]
    assert not missing, missing


def test_hip_library_loads_without_a_gpu():
    device_lib()


def no_gpu() -> bool:
    return not os.path.exists("/dev/kfd")


@pytest.mark.skipif(not no_gpu(), reason="only meaningful where no GPU is visible")
def test_context_reports_missing_device():
    with pytest.raises(nat.TcError) as e:
        Context(0)
    assert e.value.status == nat.TC_EDEVICE


def test_solver_rejects_multi_gpu():
    from hetegeneous_parallel_programming_amd import Graph

    with pytest.raises(nat.TcError) as e:
        TCSolver(Graph.er(100, 4, 1).oriented(), n_gpus=4)
    assert e.value.status == nat.TC_EINVAL


def test_bench_refuses_more_than_one_gpu():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
