"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer: the oracle
(every entry point, out-of-range bricks, a truncated buffer, non-resident
bricks in both render forms, over golden fixtures) and the C-ABI producer."""
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT

CASES = [("cfg1_b32", 32), ("microns8_b64", 64), ("rand_b2", 2), ("const_mix_b8", 8), ("u64_b8", 8),
         ("lookback16_b8", 8), ("checker_b16", 16), ("all_zero_b8", 8)]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc missing")
    out = tmp_path_factory.mktemp("san")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize", f"SAN_OUT={out}"], check=True)
    return str(out / "oracle_sanitize_check")


@pytest.mark.parametrize("name,b", CASES)
def test_oracle_is_memory_safe(checker, name, b):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([checker, os.path.join(GOLDEN, name + ".csdb"), str(b)], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stderr[-2000:]


def test_producer_is_memory_safe(tmp_path):
    """The C-ABI producer (recipes, both encoders, pyramids, invalid input)
    under AddressSanitizer + UndefinedBehaviorSanitizer, host code only."""
    if shutil.which("g++") is None:
        pytest.skip("g++ missing")
    exe = str(tmp_path / "producer_sanitize")
    csrc = os.path.join(ROOT, "fast-compressed-segmentation_amd", "csrc")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-fno-omit-frame-pointer", "-I", os.path.join(ROOT, "include"), "-I", csrc,
                    os.path.join(ROOT, "tests", "native", "producer_sanitize.cpp"),
                    os.path.join(csrc, "fcs_producer.cpp"), "-o", exe, "-lpthread"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stderr[-2000:]
