"""The server driven from several host threads on the GPU: its worker thread
runs batches while four threads submit, poll, wait and cancel; every
finished film equals the job rendered alone, and cancelled jobs are exactly
the ones cancelled.  The driver (tests/tsan/_build/stress, the product
library uninstrumented) is built with the library.  ThreadSanitizer covers
the kd builder, the host traversal and the oracle on the CPU
(tests/test_tsan.py); the HIP runtime finds no device in a process laid out
by ThreadSanitizer, so the server runs here without it."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tsan", "_build", "stress")


def test_server_from_four_threads():
    assert os.path.exists(BIN), "tests/tsan/_build/stress is built with the library (make -C tests/tsan)"
    print("[server threads] start", flush=True)
    r = subprocess.run([BIN, "server"], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert "identical to alone" in r.stdout
