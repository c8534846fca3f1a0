"""Host code under the address and undefined-behaviour sanitizers (CPU).

1. The library's plan builder (grdzhadzha_amd/csrc/grhd_plan.hpp: box
   views, block tables, internal-ghost exchange list) compiled with g++
   -fsanitize=address,undefined into tests/cpp/plan_check.cpp, which builds
   4096-box levels (tiled, ragged, with a hole) and checks every table
   against brute-force restatements.
2. The oracle (oracle/grhd_oracle.c) built the same way (`make sanitize`)
   and the oracle suite rerun against it in a child process with the ASan
   runtime preloaded; any report aborts the child.  The two tests that
   only rerun the RK4 order study with planted mistakes, and the order
   study itself, are left out for time: every oracle routine they call is
   called by the others.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
SANFLAGS = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all",
            "-fno-omit-frame-pointer", "-g"]
ORACLE_FILES = ["test_oracle_units.py", "test_oracle_anchors.py",
                "test_oracle_diagnostics.py", "test_oracle_excision.py",
                "test_oracle_inverse_dx.py", "test_oracle_mdot.py",
                "test_oracle_rk4_order.py", "test_checkpoint_cpu.py"]


def asan_runtime():
    path = subprocess.run(["gcc", "-print-file-name=libasan.so"],
                          capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(path) or not os.path.exists(path):
        pytest.skip("gcc has no ASan runtime")
    return path


def sanitizer_env():
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    return env


def test_plan_tables_of_4096_box_levels_under_sanitizers():
    asan_runtime()
    exe = os.path.join(CPP, "_build", "plan_check_san")
    os.makedirs(os.path.dirname(exe), exist_ok=True)
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra",
                    "-Werror", *SANFLAGS, "-D__HIP_PLATFORM_AMD__",
                    "-I/opt/rocm/include", "-I" + os.path.join(ROOT,
                                                               "include"),
                    "-o", exe, os.path.join(CPP, "plan_check.cpp")],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True,
                       env=sanitizer_env(), timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "plan_check ok" in r.stdout
    assert "planted exchange mistakes caught" in r.stdout
    assert "tiled: 4096 boxes" in r.stdout
    assert "split: 4096 boxes" in r.stdout
    # the holed level must be caught as unexchangeable
    holed = [l for l in r.stdout.splitlines() if l.startswith("holed")][0]
    assert " 0 orphan" not in holed


def test_oracle_suite_against_sanitized_oracle():
    runtime = asan_runtime()
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"),
                    "sanitize"], check=True)
    lib = os.path.join(ROOT, "oracle", "_san", "libgrhd_oracle.so")
    env = sanitizer_env()
    env["LD_PRELOAD"] = runtime
    env["GRHD_ORACLE_LIB"] = lib
    env["PYTHONPATH"] = ROOT
    # the child loads the sanitized build (which needs the ASan runtime
    # first in the process) and not the ordinary one
    probe = subprocess.run(
        [sys.executable, "-c",
         "import oracle; oracle.lib(); print(oracle.LIB)"],
        capture_output=True, text=True, env=env, cwd=ROOT, timeout=120)
    assert probe.returncode == 0, probe.stderr[-4000:]
    assert probe.stdout.strip() == lib
    r = subprocess.run(
        [sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
         "-m", "not gpu", "-k", "not planted and not fourth_order",
         *[os.path.join(ROOT, "tests", f) for f in ORACLE_FILES]],
        capture_output=True, text=True, env=env, cwd=ROOT, timeout=1200)
    print(r.stdout[-2000:])
    assert r.returncode == 0, (r.stdout[-4000:], r.stderr[-4000:])
    assert " passed" in r.stdout and "failed" not in r.stdout
