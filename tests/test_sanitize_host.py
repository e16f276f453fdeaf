"""Host C code of libbnamg under AddressSanitizer and
UndefinedBehaviorSanitizer: tests/sanitize_host.c drives the network
generator, RS/PMIS/HMIS set-up, value-only rebuild, IC(0), dense inverse,
schedules, COO and network-case readers and CSR helpers on small inputs,
built with gcc from the host sources that need no GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "blood_network_amd", "csrc", "host")
SOURCES = ["agg.c", "csr.c", "dense.c", "netgen.c", "nwk.c", "schedule.c", "setup.c"]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    exe = tmp_path / "sanitize_host"
    cmd = ["gcc", "-std=gnu11", "-O1", "-g", "-fno-omit-frame-pointer", "-fopenmp", "-ffp-contract=off",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-I", os.path.join(ROOT, "include"), "-I", HOST,
           os.path.join(ROOT, "tests", "sanitize_host.c")] + [os.path.join(HOST, s) for s in SOURCES] + \
          ["-lm", "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="4")
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert r.stdout.strip().splitlines()[-1] == "ok"
