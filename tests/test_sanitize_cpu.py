"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer:
tests/asan_driver.c and oracle/tq_oracle.c compiled with
-fsanitize=address,undefined (host code only; no GPU), run over tiny, odd,
non-power-of-two and padded-8192 spaces with edge rows. Any out-of-bounds
access, leak or undefined behaviour fails the run."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, ".."))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "asan_driver")
    cmd = ["gcc", "-std=c11", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           os.path.join(HERE, "asan_driver.c"), os.path.join(REPO, "oracle", "tq_oracle.c"),
           "-o", exe, "-lm"]
    subprocess.check_call(cmd)
    # verify_asan_link_order=0: the environment may preload a library ahead
    # of the ASan runtime; that is left as it is.
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    p = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, env=env,
                       timeout=600)
    assert p.returncode == 0, p.stderr[-4000:]
    assert "asan-ok" in p.stdout
    assert "runtime error" not in p.stderr and "AddressSanitizer" not in p.stderr
