"""The host library under AddressSanitizer and UndefinedBehaviorSanitizer:
tests/sanitize/host_sanitize.cpp drives generation, file I/O, trees, detour
DAGs in every mode, the route writer and the evaluator; any sanitizer report
fails the test."""
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "instantgr_amd", "csrc", "host")
FLAGS = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer",
         "-std=c++17", "-ffp-contract=off", "-pthread", "-I" + os.path.join(ROOT, "include"), "-I" + HOST]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_library_is_sanitizer_clean(tmp_path):
    srcs = [os.path.join(ROOT, "tests", "sanitize", "host_sanitize.cpp")]
    srcs += [s for s in sorted(glob.glob(os.path.join(HOST, "*.cpp"))) if not s.endswith("igr_route.cpp")]

    def compile_one(src):
        obj = str(tmp_path / (os.path.basename(src) + ".o"))
        subprocess.run(["g++", *FLAGS, "-c", src, "-o", obj], check=True, capture_output=True, timeout=600)
        return obj
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_one, srcs))
    exe = str(tmp_path / "host_sanitize")
    subprocess.run(["g++", *FLAGS, *objs, "-o", exe], check=True, capture_output=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    p = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=600, env=env)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-4000:]
    assert "runtime error" not in p.stderr and "AddressSanitizer" not in p.stderr, p.stderr[-4000:]
    assert p.stdout.strip().endswith("clean")
