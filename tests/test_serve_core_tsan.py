"""The serving loop's scheduling core (pastix_amd/csrc/serve_core.cpp: the
request queue, slot versions, merged solves, memory admission and worker
threads of pastix_amd_server_*) under ThreadSanitizer, on the CPU: the
slots are a host model of the device work (tests/tsan/serve_core_stress.cpp).
Six clients submit solves and refactorisations (one fails) under a budget
that makes slots wait for memory.  Passes when the driver reports no race
(ThreadSanitizer ends the run with status 66 on a report) and every request
ends with the solution of its version or that version's failure status."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "tsan", "serve_core_tsan")


def test_serve_core_under_thread_sanitizer():
    # (re)built from its sources by make when they are newer
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "pastix_amd", "csrc"), "tsan"], check=True)
    env = dict(os.environ)
    env["TSAN_OPTIONS"] = "halt_on_error=1 exitcode=66 second_deadlock_stack=1"
    for _ in range(3):
        p = subprocess.run([EXE], capture_output=True, text=True, timeout=120, env=env)
        print(p.stdout.strip())
        assert "ThreadSanitizer" not in p.stderr, p.stderr[-4000:]
        assert p.returncode == 0, (p.returncode, p.stderr[-4000:])
        assert "0 failures" in p.stdout
