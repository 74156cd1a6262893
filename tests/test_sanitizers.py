"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
scene engine and packer (raymotion_amd/librm_hip_san.so, `make -C
raymotion_amd/csrc san`) and the oracle (oracle/_san/liborc.so, `make -C
oracle san`) run a subset of the CPU tests in a child process that preloads
the sanitizer runtimes; any report fails the run. The whole CPU suite passes
the same way: `python -m pytest tests -m "not gpu"` in the environment _env()
builds (about 3 minutes on 8 cores)."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SAN_LIB = ROOT / "raymotion_amd" / "librm_hip_san.so"
SAN_ORC = ROOT / "oracle" / "_san" / "liborc.so"


def _runtime(name):
    r = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True)
    p = r.stdout.strip()
    return p if r.returncode == 0 and os.path.isabs(p) and os.path.exists(p) else None


def _env():
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not (SAN_LIB.exists() and SAN_ORC.exists() and asan and ubsan):
        pytest.skip("sanitizer builds or runtimes absent")
    env = dict(os.environ)
    env.update(LD_PRELOAD=f"{asan} {ubsan}", ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", RAYMOTION_HIP_LIB=str(SAN_LIB),
               RAYMOTION_ORACLE_LIB=str(SAN_ORC))
    return env


def test_sanitized_libraries_are_the_ones_loaded():
    code = ("import sys; sys.path[:0] = [%r, %r]\n"
            "import raymotion_amd as rm, oracle_bindings as ob\n"
            "rm.lib(); ob.orc()\n"
            "maps = open('/proc/self/maps').read()\n"
            "assert 'librm_hip_san.so' in maps and '_san/liborc.so' in maps and 'libasan' in maps, 'not loaded'\n"
            "print('ok')\n") % (str(ROOT), str(ROOT / "tests"))
    r = subprocess.run([sys.executable, "-c", code], env=_env(), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "ok" in r.stdout, r.stderr[-3000:]


def test_cpu_subset_under_sanitizers():
    files = ["test_engine_api.py", "test_engine_vs_reference.py", "test_oracle_kat.py", "test_oracle_pins.py",
             "test_oracle_adjoint.py", "test_pins_oracle.py"]
    args = [sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
            "-k", "not furnace and not f64 and not slab"] + [str(ROOT / "tests" / f) for f in files]
    r = subprocess.run(args, env=_env(), capture_output=True, text=True, timeout=900, cwd=str(ROOT))
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
