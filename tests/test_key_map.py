"""The two-word key mapping that digit_count applies to signed and float keys
(mgpu_key_map in kernel_args.h) equals the mapping of key_transform and the
passes (mgpu_key_to_radix / mgpu_key_from_radix), both ways, on ~12.6M bit
patterns and the special values (CPU; tests/cpp/key_map_check.cpp)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_key_map_matches_key_transform(tmp_path):
    exe = str(tmp_path / "key_map_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "mgpu_old_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "key_map_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "keys agree" in r.stdout
