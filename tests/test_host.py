"""CPU tests of librssieve.so's host side (no device calls): the library
loads and exports every entry point the headers declare, exact verification,
wheel offset, ABCD / factor-file / checkpoint formats, the pre-sieve and the
reference BSGS split."""
import math
import os
import random
import re
import subprocess

import numpy as np
import pytest

import oracle as O
import rssieve_amd as R
from util import bitmap_terms, full_bitmap

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    names = set()
    for h in ("rssieve.h", "rssieve_io.h", "rssieve_mlp.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(rs_[a-z0-9_]+)\s*\(", text))
    assert names, "no declarations found"
    L = R.lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing
    nm = subprocess.run(["nm", "-D", "--defined-only", R.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (rs_[a-z0-9_]+)$", nm, flags=re.M))
    assert names <= exported
    assert set(R.exported_symbols()) <= exported


def test_verify_factor_matches_oracle_and_python():
    rng = random.Random(11)
    L = O.lib()
    assert R.verify_factor(1000133, 91549, 54378870, 1, 2) == 1
    assert R.verify_factor(1000669, 131179, 72017282, -1, 2) == 0
    assert R.verify_factor(1000669, 131179, 72017282, 1, 2) == 1
    assert R.verify_factor(2047, 1, 11, -1, 2) == -1
    for _ in range(2000):
        p = rng.randrange(3, 2**64) | 1
        k = rng.randrange(1, 2**32)
        n = rng.randrange(1, 2**31)
        c = rng.choice([1, -1])
        b = rng.choice([2, 3, 5, 6, 10, 7, 2**31 + 11])
        assert R.verify_factor(p, k, n, c, b) == L.or_verify_factor(p, k, n, c, b)
    # true factors: p | k*b^n + c by construction
    for _ in range(300):
        p = rng.randrange(10**6, 10**9) | 1
        while not L.or_is_prime(p):
            p += 2
        b, n, c = 2, rng.randrange(1, 10**6), rng.choice([1, -1])
        k = (-c) * pow(pow(b, n, p), -1, p) % p
        if k == 0:
            continue
        assert R.verify_factor(p, k, n, c, b) == 1
        assert R.verify_factor(p, k + p, n, c, b) == 1
    for n in range(1, 5000):
        assert R.is_prime(n) == bool(L.or_is_prime(n)) or n < 3


def test_wheel_offset_matches_oracle():
    import ctypes as C
    L = O.lib()
    for s0 in list(range(0, 2000)) + [2**64 - 5000 + 3 * i for i in range(1000)]:
        s = C.c_uint64(s0)
        i = C.c_int32()
        L.or_find_wheel_offset(C.byref(s), C.byref(i))
        assert R.wheel_offset(s0) == (s.value, i.value)


def test_reference_params_match_oracle():
    for (nmin, nmax, K, lmem) in [(100000, 199999, 10, 65536), (3000000, 9999999, 100, 65536),
                                  (1000000, 2**24 - 1, 100, 65536), (0, 10**4, 1, 65536), (0, 10**8, 5, 16384),
                                  (0, 10**8, 5, 32768)]:
        prm = R.reference_params(nmin, nmax, K, lmem, 256)
        op = O.bsgs_params(nmin, nmax, K, lmem)
        M, QQ, MM = R.derive(prm, nmin, nmax)
        assert (prm.Q, prm.hsize, M, QQ, MM) == (op.Q, op.hsize, op.M, op.QQ, op.MM)
        assert prm.threads == 256 and prm.seg_primes == 65536


def test_factor_line_format_and_parse():
    line = R.format_factor(1000133, 91549, 2, 54378870, 1)
    assert line == "1000133 | 91549*2^54378870+1\n"
    assert R.format_factor(2**64 - 59, 2**32 - 1, 10, 1, -1) == "18446744073709551557 | 4294967295*10^1-1\n"
    assert R.parse_factor("1000133 | 91549*2^54378870+1\n", 2) == (1000133, 91549, "+", 54378870)
    assert R.parse_factor("  1000669 |131179 * 2 ^ 72017282 - 1", 2) == (1000669, 131179, "-", 72017282)
    for bad in ["1000133 | 91549*3^54378870+1", "1000133 | 91549*2^54378870+2", "no factors",
                "1000133 91549*2^54378870+1", "1000133 | 91549*2^54378870+1 x"]:
        with pytest.raises(R.RSError):
            R.parse_factor(bad, 2)


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)


def test_abcd_read_semantics(tmp_path):
    p = tmp_path / "a.abcd"
    _write(p, "# comment\n\nABCD 3*2^$a+1 [101] // Sieved to 1000\n2\n  4\nABCD 5*2^$a-1 [200]\n1\n   \n# x\n7\n")
    a = R.ABCD(str(p))
    assert (a.base, a.kcount, a.nmin, a.nmax) == (2, 2, 100, 208)
    assert list(a.k) == [3, 5] and list(a.c) == [1, -1]
    terms = bitmap_terms(a.bitmap(), 2, a.words_per_seq, a.nmin)
    assert terms == {(0, 101), (0, 103), (0, 107), (1, 200), (1, 201), (1, 208)}
    assert a.nterms == 6
    assert a.mark_used(3, "+", 103) == 1 and a.mark_used(3, "+", 103) == 0
    assert a.mark_used(3, "-", 101) == 0 and a.mark_used(99, "+", 101) == 0
    assert (0, 103) not in bitmap_terms(a.bitmap(), 2, a.words_per_seq, a.nmin)


@pytest.mark.parametrize("text,msg", [
    ("5\nABCD 3*2^$a+1 [10]\n", "delta line before first ABCD header"),
    ("ABCD 3*2^$a+1 [10]\nhello\n", "(seq 0) invalid or unrecognised line"),
    ("ABCD 0*2^$a+1 [10]\n", "invalid K (must be > 0)"),
    ("ABCD 3*2^$a+1 [0]\n", "invalid N0 (must be > 0)"),
    ("ABCD 3*1^$a+1 [10]\n", "invalid base in ABCD header (B must be >= 2)"),
    ("ABCD 3*2^$a+1 [10]\nABCD 5*3^$a+1 [10]\n", "mismatched base (B must match all sequences)"),
    ("ABCD 3*2^$a+3 [10]\n", "not standard form k*b^n+/-1 (c must be +1 or -1)"),
    ("ABCD 3*2^$a+1 [4294967290]\n10\n", "overflow while accumulating n += delta (uint32_t wrap)"),
    ("# only comments\n", "no sequences found in input"),
])
def test_abcd_errors(tmp_path, text, msg):
    p = tmp_path / "bad.abcd"
    _write(p, text)
    with pytest.raises(R.RSError) as e:
        R.ABCD(str(p))
    assert msg in str(e.value)


def test_abcd_too_many_sequences(tmp_path):
    p = tmp_path / "many.abcd"
    _write(p, "".join("ABCD %d*2^$a+1 [10]\n" % (2 * i + 3) for i in range(101)))
    with pytest.raises(R.RSError) as e:
        R.ABCD(str(p))
    assert "too many sequences (MAX_SEQUENCES=100)" in str(e.value)


def test_abcd_write_round_trip(tmp_path):
    ks, cs = [3, 7, 11], [1, -1, 1]
    nmin, nmax = 1000, 5000
    bm, wps = full_bitmap(3, nmin, nmax, density=0.05, seed=3)
    path = tmp_path / "rt.abcd"
    R.abcd_write(str(path), 2, ks, cs, nmin, nmax, bm, wps, sieved_to=65536)
    a = R.ABCD(str(path))
    assert list(a.k) == ks and list(a.c) == cs and a.base == 2
    t0 = bitmap_terms(bm, 3, wps, nmin)
    t1 = {(i, n) for (i, n) in bitmap_terms(a.bitmap(), 3, a.words_per_seq, a.nmin)}
    assert t0 == t1
    first = open(path).readline()
    assert re.match(r"ABCD 3\*2\^\$a\+1 \[\d+\] // Sieved to 65536$", first.strip())


@pytest.mark.parametrize("base,ks,cs,nmin,nmax,bound", [
    (2, [3, 5, 7, 1001], [1, -1, 1, -1], 100, 3000, 2000),
    (3, [2, 4, 10], [1, -1, 1], 50, 2000, 1000),
    (10, [3, 7], [-1, 1], 10, 1500, 3000),
    (6, [5, 7, 11], [1, 1, -1], 0, 600, 500),
])
def test_presieve_brute_force(base, ks, cs, nmin, nmax, bound):
    bm, wps = full_bitmap(len(ks), nmin, nmax)
    out = R.presieve(base, ks, cs, nmin, nmax, bm.copy(), wps, bound, nthreads=3)
    primes = [q for q in range(2, bound) if all(q % d for d in range(2, int(math.isqrt(q)) + 1))]
    keep = set()
    for i, (k, c) in enumerate(zip(ks, cs)):
        for n in range(nmin, nmax + 1):
            v = k * base**n + c
            if not any(v % q == 0 for q in primes):
                keep.add((i, n))
    assert bitmap_terms(out, len(ks), wps, nmin) == keep


def test_state_write_read_alternation(tmp_path):
    d = str(tmp_path)
    st = R.State()
    st.pmin, st.pmax, st.p, st.base, st.nmin, st.nmax, st.kcount = 10**9, 2 * 10**9, 10**9, 2, 100, 200, 3
    for i, v in enumerate([3, -5, 7]):
        st.klist[i] = v
    wa = __import__("ctypes").c_int(1)
    L = R.lib()
    assert L.rs_state_write(d.encode(), st, wa) == 0 and wa.value == 0
    assert os.path.getsize(os.path.join(d, "stateA.ckp")) == 488  # sizeof(workStatus), cl_sieve.h:21-26
    st.p = 10**9 + 5000
    st.factorcount = 4
    assert L.rs_state_write(d.encode(), st, wa) == 0 and wa.value == 1
    want = R.State()
    want.pmin, want.pmax, want.base, want.nmin, want.nmax, want.kcount = 10**9, 2 * 10**9, 2, 100, 200, 3
    w2 = __import__("ctypes").c_int(0)
    assert L.rs_state_read(d.encode(), want, w2) == 1
    assert want.p == 10**9 + 5000 and want.factorcount == 4 and w2.value == 1  # B was newest
    # corrupt B: A is used
    with open(os.path.join(d, "stateB.ckp"), "r+b") as f:
        f.seek(16)
        f.write(b"\x01")
    want2 = R.State()
    want2.pmin, want2.pmax, want2.base, want2.nmin, want2.nmax, want2.kcount = 10**9, 2 * 10**9, 2, 100, 200, 3
    assert L.rs_state_read(d.encode(), want2, w2) == 1 and want2.p == 10**9 and w2.value == 0
    assert "Checksum error in stateB.ckp" in L.rs_last_error().decode()
    # different search: no valid state
    want3 = R.State()
    want3.pmin, want3.pmax, want3.base, want3.nmin, want3.nmax, want3.kcount = 10**9, 3 * 10**9, 2, 100, 200, 3
    assert L.rs_state_read(d.encode(), want3, w2) == 0
