"""CPU model of the device task ring's hand-off protocol (tests/model/
ring_model.cpp): sequence words per slot, slot reuse, positions wrapping
modulo 2^32, busy raised before a claim. Many threads donate and claim on a
ring of a few slots; every record must be read by its own claimant and the
searched node count must equal the closed form."""
import json
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def binary(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("ring") / "ring_model")
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-pthread", "-o", out,
                           os.path.join(HERE, "model", "ring_model.cpp")])
    return out


@pytest.mark.parametrize("slots,workers,depth", [(2, 4, 10), (8, 8, 14), (64, 8, 14), (1024, 6, 12)])
def test_ring_hand_off(binary, slots, workers, depth):
    r = subprocess.run([binary, str(slots), str(workers), str(depth)], capture_output=True, text=True,
                       timeout=120)
    res = json.loads(r.stdout)
    assert r.returncode == 0, res
    assert res["nodes"] == res["expect"] and res["bad"] == 0
    assert res["wrapped"] == 1 and res["donated"] > slots, res
