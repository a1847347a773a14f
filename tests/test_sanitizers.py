"""AddressSanitizer + UndefinedBehaviorSanitizer builds (CPU only): the C
oracle over scenes reaching every profile, PSF, mask and noise branch, and
the library's host code (argument validation, workspace layouts, error
messages of every C ABI entry point) with the sanitizers on the host pass
only -- its device code is compiled as usual and never launched here."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
NATIVE = os.path.join(HERE, "native")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ROCM = "/opt/rocm"


def _run(cmd, env=None, timeout=600):
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env, timeout=timeout)
    return res.returncode, res.stdout


def _env(leaks):
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=%d:abort_on_error=0:verify_asan_link_order=0" % leaks
    env["UBSAN_OPTIONS"] = "print_stacktrace=1"
    return env


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
def test_c_oracle_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "oracle_checks")
    rc, out = _run(["gcc", "-std=gnu11", "-fopenmp"] + SAN +
                   [os.path.join(NATIVE, "oracle_checks.c"), os.path.join(ROOT, "oracle", "lensim_oracle.c"),
                    "-lm", "-o", exe])
    assert rc == 0, out
    rc, out = _run([exe], env=_env(1), timeout=300)
    assert rc == 0 and "oracle checks passed" in out, out
    assert "runtime error" not in out and "AddressSanitizer" not in out, out


@pytest.mark.skipif(not os.path.exists(os.path.join(ROCM, "bin", "hipcc")), reason="hipcc not found")
def test_host_layer_under_asan_ubsan(tmp_path):
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    clang = os.path.join(ROCM, "llvm", "bin", "clang")
    inc = "-I" + os.path.join(ROOT, "include")
    lib_o, drv_o, exe = str(tmp_path / "lensim.o"), str(tmp_path / "drv.o"), str(tmp_path / "host_checks")
    # sanitizers on the host pass only (each -fsanitize= behind -Xarch_host)
    host_san = []
    for f in SAN:
        host_san += ["-Xarch_host", f]
    rc, out = _run([hipcc, "--offload-arch=gfx950", "-std=c++17", "-O1", inc,
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"] + host_san +
                   ["-c", os.path.join(ROOT, "jaxtronomy_amd", "csrc", "lensim.hip"), "-o", lib_o])
    assert rc == 0, out
    rc, out = _run([clang, "-std=gnu11", inc] + SAN + ["-c", os.path.join(NATIVE, "host_abi_checks.c"), "-o", drv_o])
    assert rc == 0, out
    rc, out = _run([clang + "++", "-fsanitize=address,undefined", drv_o, lib_o, "-L" + os.path.join(ROCM, "lib"),
                    "-lamdhip64", "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-o", exe])
    assert rc == 0, out
    # the HIP runtime's own allocations outlive main: leak checking off for this one
    rc, out = _run([exe], env=_env(0), timeout=300)
    assert rc == 0 and "host ABI checks passed" in out, out
    assert "runtime error" not in out and "AddressSanitizer" not in out, out
