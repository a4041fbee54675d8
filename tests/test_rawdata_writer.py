"""The driver's ASCII field writer (save_ascii's format, output_styles.cpp:
182-327) formats row blocks on several threads: cuda-rp_amd/bin/rawdata_bench
checks that its bytes equal the serial ostream output (plain and
cylindrical-mirrored layouts, several thread counts) and times a 4096^2
field (CPU only)."""
import os
import subprocess
import tempfile

import cases


def test_parallel_ascii_writer_matches_serial():
    exe = os.path.join(cases.ROOT, "cuda-rp_amd", "bin", "rawdata_bench")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", os.path.join(cases.ROOT, "cuda-rp_amd"), "bin/rawdata_bench"],
                       check=True)
    threads = min(16, os.cpu_count() or 1)
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([exe, d, "4096", str(threads)], capture_output=True, text=True,
                           timeout=600)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr
