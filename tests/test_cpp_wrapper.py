"""The C++23 host layer (include/aleator_gpu.hpp) through its own test program
(tests/cpp/test_wrapper.cc): exception mapping and cache keys on the CPU; on the device the
grid build and interpolate, the cache round trip, the EwaldIncrementalState sequence and Widom
K_H, each against the oracle."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "_build", "test_wrapper")


def run(mode):
    if not os.path.exists(BIN):  # normally built by __graft_entry__.build()
        subprocess.run(["make", "-C", os.path.join(HERE, "cpp")], check=True, capture_output=True)
    r = subprocess.run([BIN, mode], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout


def test_cpp_host_layer_cpu():
    run("cpu")


@pytest.mark.gpu
def test_cpp_host_layer_gpu():
    print("C++ host layer on the device", flush=True)
    run("gpu")
