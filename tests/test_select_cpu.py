"""Compiles tests/cpu/select_test.cpp (the sweep's per-lane list compaction,
holomap_amd/csrc/hm_select.hpp, built as host code) and runs it: adversarial
lists (ties, -inf keys, huge and subnormal values) and streamed insertion."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_lane_compaction_cpu(tmp_path):
    exe = tmp_path / "select_test"
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe),
                    os.path.join(HERE, "cpu", "select_test.cpp")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK" in r.stdout
