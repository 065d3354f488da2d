"""The serving loop's host logic under sanitizers, on the CPU.

serve.cpp (queueing, admission by footprint, least-recently-used eviction,
refit on demand, coalescing, release while predictions are queued, spare-model
reuse) is built with g++ against a host-only stand-in of the device calls
(tests/host_stub/device_stub.cpp) and driven by many client threads
(tests/host_stub/serve_stress.cpp), once under ThreadSanitizer and once under
AddressSanitizer + UndefinedBehaviorSanitizer.  The driver checks every
prediction against the stand-in's closed form, the status of failed fits and of
predictions on them, refusal after release, and that the small budget forced
evictions; the sanitizers fail the run on any data race, memory error or
undefined behaviour.
"""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(ROOT, "tests", "host_stub", "serve_stress.cpp"),
       os.path.join(ROOT, "tests", "host_stub", "device_stub.cpp"),
       os.path.join(ROOT, "gplite_amd", "csrc", "serve.cpp")]

SANITIZERS = {
    "thread": ["-fsanitize=thread"],
    "address_undefined": ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"],
}


def _build(tmp, name):
    exe = os.path.join(tmp, "serve_" + name)
    cmd = ["g++", "-std=c++17", "-O1", "-g", *SANITIZERS[name], "-I" + os.path.join(ROOT, "gplite_amd", "csrc"),
           *SRC, "-o", exe, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-4000:]
    return exe


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("name", sorted(SANITIZERS))
@pytest.mark.parametrize("threads,rounds,workers,budget", [(6, 40, 4, 24 << 20), (8, 30, 16, 4_500_000),
                                                           (4, 30, 1, 24 << 20)])
def test_serving_loop_under_sanitizers(tmp_path_factory, name, threads, rounds, workers, budget):
    tmp = tmp_path_factory.getbasetemp()
    exe = os.path.join(str(tmp), "serve_" + name)
    if not os.path.exists(exe):
        exe = _build(str(tmp), name)
    env = dict(os.environ)
    env["TSAN_OPTIONS"] = "halt_on_error=1 second_deadlock_stack=1"
    env["ASAN_OPTIONS"] = "detect_leaks=1 halt_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1 halt_on_error=1"
    r = subprocess.run([exe, str(threads), str(rounds), str(workers), str(budget)], capture_output=True,
                       text=True, timeout=240, env=env)
    if name == "thread" and "unexpected memory mapping" in r.stderr:
        pytest.skip("ThreadSanitizer cannot map its shadow memory on this kernel")
    print(r.stdout)
    assert r.returncode == 0, (r.stdout + r.stderr)[-6000:]
    assert "OK" in r.stdout
