"""The BOINC shim of the command line (rssieve_amd/csrc/rs_boinc.h), on the
host: the init_data.xml device tag (main.cpp:207-229), the fraction done
(cl_sieve.cpp:1406-1414) and the daily trickle-up message (cl_sieve.cpp:40-62),
checked by tests/native/boinc_shim_check.cpp.  The command line's use of them
is tested on the GPU in test_gpu_cli.py."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_boinc_shim_host(tmp_path):
    exe = str(tmp_path / "boinc_shim_check")
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-o", exe, os.path.join(HERE, "native", "boinc_shim_check.cpp")],
                   check=True)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
