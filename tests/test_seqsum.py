"""The transducer arithmetic of the exact sequential float sums
(csrc/seqsum.hpp), which reproduces std::accumulate<float> in the OC update
(cpp/src/driver.cpp:181, 188-189, 201), checked on the CPU against real binary32
additions: tests/native/seqsum_check.cpp, built here with g++."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_seqsum_transducers_match_float_adds(tmp_path):
    exe = str(tmp_path / "seqsum_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-o", exe,
                    os.path.join(HERE, "native", "seqsum_check.cpp")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
