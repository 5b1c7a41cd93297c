"""CPU test of the binary64 modular product (rs_mont.h fp_mulmod*), the
arithmetic of the sieve's baby and giant steps for p < 2^51: a host build of
the same header is compared with 128-bit integer products on random and
extreme operands (tests/native/fp_mulmod_check.cpp), and the same check is
shown to fail for moduli just above the stated limit."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("fp") / "fp_mulmod_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "native", "fp_mulmod_check.cpp")],
                   check=True)
    return exe


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_fp_mulmod_exact_below_2_51(checker, seed):
    r = subprocess.run([checker, "1000000", str(seed)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert "mismatches 0" in r.stdout


def test_fp_mulmod_check_detects_moduli_above_limit(checker):
    # between 2^51 and 2^52 the quotient estimate can be off by one more: the
    # sieve keeps those primes on the integer Montgomery path
    r = subprocess.run([checker, "200000", "7", "52"], capture_output=True, text=True)
    assert r.returncode != 0 and "mismatches 0" not in r.stdout


def test_fp_sprp2_agrees_with_the_integer_test(tmp_path):
    """The prime generator's binary64 strong-PRP test below 2^51 gives the
    integer Montgomery test's answer (the reference's steps) on random odd N,
    every N = 2^t + 1, known base-2 strong pseudoprimes, every odd N below
    200001 and the odd N just below 2^51."""
    exe = str(tmp_path / "fp_sprp_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "native", "fp_sprp_check.cpp")],
                   check=True)
    r = subprocess.run([exe, "300000", "5"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert "mismatches 0" in r.stdout
    passing = int(r.stdout.split("passing")[1].split()[0])
    assert passing > 18000  # the primes below 200001 among them
