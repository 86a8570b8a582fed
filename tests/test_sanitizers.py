"""The C host layer (wfl_host.c, wfl_effact.c) and the FP64 oracle
(oracle/wfl_oracle.c) under AddressSanitizer, UndefinedBehaviorSanitizer and
LeakSanitizer: a gcc build of tests/c (sanitize_driver.c, with host_stubs.c
standing in for the device layer) runs every host entry point that needs no
device (loop generator, text and binary loop files, getwl, spline set-up,
EVMean with the floating-point checks, the profiles' closed forms, Tfunc's
analytic branches at verbosity 2, L_eff and the lattice outputs) and the
oracle's batch, spline and closed forms; any sanitizer report fails."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_code_and_oracle_are_sanitizer_clean(tmp_path):
    exe = tmp_path / "sanitize_driver"
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "c"), "sanitize", f"OUT={exe}"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, env=env, timeout=300)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out and "LeakSanitizer" not in out
    assert "sanitize driver OK" in r.stdout
