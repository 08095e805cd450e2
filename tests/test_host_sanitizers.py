"""The host-side C/C++ (generators, transpose, Matrix Market reader, autotuner table) and
the CPU oracle built with AddressSanitizer and UndefinedBehaviorSanitizer by g++/gcc and
run on small inputs, malformed files included (tests/sanitize/host_checks.cpp).  No GPU."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "hsmu-spgemm_amd", "csrc")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1",
       "-fopenmp"]


def test_host_code_under_asan_ubsan(tmp_path):
    inc = ["-I", os.path.join(ROOT, "include")]
    ora = tmp_path / "oracle.o"
    subprocess.check_call(["gcc", "-std=c11", *SAN, "-c", os.path.join(ROOT, "oracle", "hsmu_oracle.c"), "-o",
                           str(ora)])
    exe = tmp_path / "host_checks"
    subprocess.check_call(["g++", "-std=c++17", *SAN, *inc, os.path.join(ROOT, "tests", "sanitize", "host_checks.cpp"),
                           os.path.join(CSRC, "host_inputs.cpp"), os.path.join(CSRC, "tune.cpp"), str(ora), "-o",
                           str(exe), "-lm"])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", OMP_NUM_THREADS="2")
    out = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, env=env, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "HOST_SANITIZE_OK" in out.stdout
