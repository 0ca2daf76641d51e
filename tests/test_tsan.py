"""The threaded host code under ThreadSanitizer (CPU): the kd builder's
subtree threads and the host traversal's worker threads (mpt_kd_build,
mpt_kd_trace), and the oracle's thread pool, built with -fsanitize=thread by
tests/tsan/Makefile.  Each run reports no race and gives results identical to
one thread; a planted race on a plain counter is reported, which shows the
instrumentation is live.  The server's worker thread needs a GPU
(tests/test_gpu_server_threads.py), where it runs without the sanitizer:
the HIP runtime finds no device in a process laid out by it."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
TSAN = os.path.join(HERE, "tsan")
BIN = os.path.join(TSAN, "_build", "race")


def tsan_env():
    env = dict(os.environ)
    env["TSAN_OPTIONS"] = (f"halt_on_error=1 exitcode=66 second_deadlock_stack=1 "
                           f"suppressions={os.path.join(TSAN, 'suppressions.txt')}")
    return env


def run(mode, timeout=600):
    return subprocess.run([BIN, mode], capture_output=True, text=True, timeout=timeout, env=tsan_env())


@pytest.fixture(scope="module")
def built():
    subprocess.run(["make", "-C", TSAN, "-j4"], check=True, capture_output=True, timeout=900)
    assert os.path.exists(BIN)


def test_planted_race_is_reported(built):
    r = run("selftest")
    assert r.returncode == 66, r.stdout + r.stderr
    assert "ThreadSanitizer: data race" in r.stderr


@pytest.mark.parametrize("mode", ["kd", "oracle"])
def test_no_race(built, mode):
    r = run(mode)
    print(r.stdout, flush=True)
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert "identical" in r.stdout
