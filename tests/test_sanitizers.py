"""AddressSanitizer and UndefinedBehaviorSanitizer over the host code (CPU only).

* The product's equation compiler (gopf_amd/csrc/model.cpp, which restates
  pf/rhsBuilder.go:26-242 and pf/util.go:17-213) is built twice with g++ around
  tests/sanitize/model_driver.cpp, plainly and with -fsanitize=address,undefined (UB aborts),
  and both builds parse the parser-parity corpus, the reference's panicking equations and a
  registered function.  The sanitized run must be clean and print exactly what the plain
  run prints.
* The C oracle (oracle/pf_oracle.c) is built with the same sanitizers and the oracle's own CPU
  tests run against it in a child process (libasan preloaded, as Python itself is not
  instrumented).
"""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)

from test_parser_parity import CORPUS  # noqa: E402

SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g"]


def _script():
    lines = []
    for fields, scalars, eqs in CORPUS:
        lines += [f"F {f}" for f in fields]
        lines += [f"S {k} {float(v)!r} 0" for k, v in scalars.items()]
        lines += [f"E {e}" for e in eqs]
        lines += ["I", "R"]
    # the reference's panics (TestPanicOnUnknownName, TestNameFromLeibniz, model.go:480-485)
    for eq in ("dconc/dt = m1*conc", "dconc/dt = LAP otherField", "dc/dq = LAP conc", "dconc/dt = a = b"):
        lines += ["F conc", f"E {eq}", "I", "R"]
    lines += ["F LAPfield", "R", "S -x 1 0", "R"]
    # a registered function (pf/model.go:400-412) as a derived brick of the right-hand side
    lines += ["F conc", "S gamma 0.3 0", "P CHEMPOT conc -1 3", "E dconc/dt = CHEMPOT + gamma*LAP conc", "I", "R"]
    return "\n".join(lines) + "\n"


def _gxx(out, extra):
    src = [os.path.join(HERE, "sanitize", "model_driver.cpp"), os.path.join(ROOT, "gopf_amd", "csrc", "model.cpp")]
    cmd = ["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "gopf_amd", "csrc")] + extra + src + ["-o", out]
    subprocess.check_call(cmd)


def test_equation_compiler_under_asan_ubsan(tmp_path):
    plain, san = str(tmp_path / "driver"), str(tmp_path / "driver_san")
    _gxx(plain, [])
    _gxx(san, SAN)
    script = _script()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    a = subprocess.run([plain], input=script, capture_output=True, text=True, timeout=300)
    b = subprocess.run([san], input=script, capture_output=True, text=True, timeout=300, env=env)
    assert a.returncode == 0, a.stderr
    assert b.returncode == 0, b.stderr
    assert "runtime error" not in b.stderr and "AddressSanitizer" not in b.stderr, b.stderr
    assert a.stdout == b.stdout
    # the corpus ran: every case initialised, and every panic was reported
    assert a.stdout.count("PANIC") == 6
    assert "term 0 kind 0" in a.stdout and "brick CHEMPOT" in a.stdout


def test_oracle_under_asan_ubsan(tmp_path):
    from oracle import engine as E
    so = E.build(force=True, path=str(tmp_path / "liboracle_san.so"), sanitize=True)
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(os.environ, GPF_ORACLE_LIB=so, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0",
               UBSAN_OPTIONS="print_stacktrace=1", PYTHONPATH=ROOT)
    tests = ["tests/test_oracle_fft.py", "tests/test_oracle_model.py", "tests/test_oracle_terms.py",
             "tests/test_oracle_adjoint.py", "tests/test_oracle_noise.py"]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider"] + tests, cwd=ROOT,
                       capture_output=True, text=True, timeout=900, env=env)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, tail
    assert " passed" in r.stdout
