"""CPU: the oracle and the product's host-only C++ (BRK-A-01 set-up, closed-form Ke0) built with
AddressSanitizer and UndefinedBehaviorSanitizer and run on small cases; any memory error or
undefined behaviour aborts the program (-fno-sanitize-recover=all)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_code_under_asan_ubsan():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"])
    # verify_asan_link_order=0: an environment that preloads another library keeps it; the
    # sanitizer runtime is linked into the program itself
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(ROOT, "oracle", "_build", "sanitize_check")],
                       capture_output=True, text=True, env=env, timeout=600)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "sanitize_check: ok" in r.stdout
