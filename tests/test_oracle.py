"""Pins the CPU oracle (oracle/rs_oracle.c) before it is trusted as the checker.

Parity status of each oracle function and the checks that pin it:
  or_invert, or_mmul, or_add       closed form a*b*2^-64 mod p (Python ints) and
                                   the reference's own verify_factor.c built
                                   under oracle/_ref (bit-exact)
  or_powmod*/or_basepowmod*        closed form pow() incl. the exp == 1 masked-
                                   shift case x^(2^64+1)
  or_strong_prp_two                independent Python strong test, OEIS A001262
                                   base-2 strong pseudoprimes, Fermat numbers
  or_find_wheel_offset             definition (first N >= start coprime to 30
                                   and its position in the 7,11,13,...,1 wheel)
  bit tables                       closed form, and the reference source text
                                   when the reference tree is present
  or_getsegprimes/_in_range        Eratosthenes + the pseudoprime definition
  or_bsgs_params_for               invariants of the split (power of two,
                                   coverage M*Q >= L, LDS fit)
  or_profile_range                 its defining equation x/ln x = 256*CU +
                                   start/ln start, the 1.5x buffer, the cap,
                                   and an actual prime count
  or_setup_prime                   brute-force discrete logs for small primes;
                                   for composite 2-PRPs (42799 = 127*337, ...)
                                   the closed form of setup.cl's decisions in
                                   plain modular arithmetic (skip test, power-
                                   residue tests, parity classes, targets)
  or_sieve                         brute-force divisibility of every term for
                                   small configurations; duplicate count from
                                   the brute-force hit count of each window;
                                   ranges with composite 2-PRPs: the terms a
                                   composite clears are the closed form's
  or_is_prime, or_verify_factor    reference verify_factor.c (oracle/_ref),
                                   Python MR, the reference's example lines
"""
import math
import os
import random

import numpy as np
import pytest

import oracle as O

R64 = 1 << 64
EDGE_P = [3, 5, 7, 127, 65521, 2**32 - 5, 2**32 + 15, 1000000007, 2**50 + 55, 2**61 - 1, 2**63 + 29, 2**64 - 59]


def py_is_prime(n):
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def py_sprp2(n):
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    x = pow(2, d, n)
    if x in (1, n - 1):
        return True
    for _ in range(s - 1):
        x = x * x % n
        if x == n - 1:
            return True
    return False


def mont(x, p):
    return x * R64 % p


def test_invert_and_mmul_closed_form():
    L = O.lib()
    rng = random.Random(1)
    for p in EDGE_P:
        q = L.or_invert(p)
        assert (p * q) % R64 == 1
        rinv = pow(R64, -1, p)
        for _ in range(200):
            a, b = rng.randrange(p), rng.randrange(p)
            assert L.or_mmul(a, b, p, q) == a * b * rinv % p
        for a, b in [(0, 0), (p - 1, p - 1), (p - 1, 1), (1, 1)]:
            assert L.or_mmul(a, b, p, q) == a * b * rinv % p
            assert L.or_add(a, b, p) == (a + b) % p


@pytest.mark.skipif(O.ref_lib() is None, reason="oracle/_ref not built (reference tree absent)")
def test_mmul_matches_reference_build():
    L, Rf = O.lib(), O.ref_lib()
    rng = random.Random(2)
    for p in EDGE_P:
        q = L.or_invert(p)
        assert Rf.ref_invert(p) == q
        for _ in range(500):
            a, b = rng.randrange(p), rng.randrange(p)
            assert L.or_mmul(a, b, p, q) == Rf.ref_m_mul(a, b, p, q)
            assert L.or_add(a, b, p) == Rf.ref_add(a, b, p)


def test_powmods_closed_form():
    L = O.lib()
    rng = random.Random(3)
    for p in EDGE_P[2:]:
        q = L.or_invert(p)
        one = mont(1, p)
        for _ in range(40):
            x = rng.randrange(1, p)
            mx = mont(x, p)
            e = rng.randrange(2, 2**63)
            assert L.or_powmodlg(mx, e, p, q) == mont(pow(x, e, p), p)
            e32 = rng.randrange(2, 2**32)
            assert L.or_powmodsm(mx, e32, p, q, one) == mont(pow(x, e32, p), p)
            for base in (2, 3, 5, 6, 10, 1009):
                mb = mont(base % p, p)
                assert L.or_basepowmodlg(mb, e, base, p, q) == mont(pow(base, e, p), p)
                assert L.or_basepowmodsm(mb, e32, base, p, q, one) == mont(pow(base, e32, p), p)
            # masked shift: exponent 1 squares 64 times, then multiplies once
            assert L.or_powmodlg(mx, 1, p, q) == mont(pow(x, 2**64 + 1, p), p)
            assert L.or_powmodsm(mx, 0, p, q, one) == one
            assert L.or_powmodsm(mx, 1, p, q, one) == mx


A001262 = [2047, 3277, 4033, 4681, 8321, 15841, 29341, 42799, 49141, 52633, 65281, 74665, 80581, 85489, 88357,
           90751, 104653, 130561, 196093, 220729, 233017, 252601, 253241, 256999, 271951, 280601, 314821, 357761,
           390937, 458989, 476971, 486737]


def test_strong_prp_two():
    L = O.lib()
    for n in A001262:
        assert L.or_strong_prp_two(n) == 1
        assert not py_is_prime(n)
    for n in range(5, 20001, 2):
        assert L.or_strong_prp_two(n) == int(py_sprp2(n)), n
    rng = random.Random(4)
    for _ in range(3000):
        n = rng.randrange(2**40, 2**64) | 1
        assert L.or_strong_prp_two(n) == int(py_sprp2(n))
    # N = 2^t + 1 takes the 64-squaring path; the result equals the closed form
    for t in (8, 16, 32, 40, 48):
        n = 2**t + 1
        a = pow(2, 2**64 + 1, n)
        ok = a in (1, n - 1)
        for _ in range(t - 1):
            if ok:
                break
            a = a * a % n
            ok = a == n - 1
        assert L.or_strong_prp_two(n) == int(ok)


def test_find_wheel_offset_definition():
    import ctypes as C
    L = O.lib()
    res = [7, 11, 13, 17, 19, 23, 29, 1]
    starts = list(range(0, 3000)) + [2**64 - 1000 + i for i in range(0, 900, 7)] + [10**9 + i for i in range(100)]
    for s0 in starts:
        s = C.c_uint64(s0)
        i = C.c_int32(0)
        L.or_find_wheel_offset(C.byref(s), C.byref(i))
        n = s0
        while math.gcd(n, 30) != 1:
            n += 1
        assert s.value == n
        assert res[i.value] == n % 30


def test_bit_tables_closed_form():
    L = O.lib()
    sp = [7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97, 101, 103, 107,
          109, 113]
    for s, p in enumerate(sp):
        for r in range(p):
            w = L.or_bitsieve_word(s, r)
            for i in range(32):
                assert ((w >> i) & 1) == int((r + 2 * i) % p == 0)


REF_GETSEG = "/root/reference/kernels/getsegprimes.cl"


@pytest.mark.skipif(not os.path.isfile(REF_GETSEG), reason="reference tree absent")
def test_bit_tables_match_reference_source():
    import re
    L = O.lib()
    text = open(REF_GETSEG).read()
    sp = [7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97, 101, 103, 107,
          109, 113]
    for s, p in enumerate(sp):
        m = re.search(r"__constant uint p%d\[%d\] = \{([^}]*)\}" % (p, p), text)
        assert m, p
        vals = [int(v) for v in m.group(1).split(",")]
        assert vals == [L.or_bitsieve_word(s, r) for r in range(p)]


def _eratosthenes(lo, hi):
    is_p = np.ones(hi - lo, dtype=bool)
    for q in range(2, int(math.isqrt(hi)) + 1):
        st = max(q * q, (lo + q - 1) // q * q)
        is_p[st - lo::q] = False
    out = np.nonzero(is_p)[0] + lo
    return out[out >= 2]


@pytest.mark.parametrize("lo,hi", [(3, 200000), (113, 114), (100, 1000), (10**9, 10**9 + 2 * 10**5),
                                   (4 * 10**6, 4 * 10**6 + 10**5)])
def test_primes_in_range(lo, hi):
    got = set(O.primes_in_range(lo, hi, seg=12345).tolist())
    want = set(_eratosthenes(lo, hi).tolist())
    psp = {n for n in A001262 if lo <= n < hi and all(n % int(q) for q in _eratosthenes(2, 114))}
    assert got == want | psp
    if lo <= 42799 < hi:
        assert 42799 in got  # 127 * 337: a 2-SPSP with no factor <= 113


def test_primes_near_top_of_range():
    lo, hi = 2**64 - 20000, 2**64 - 1
    got = set(O.primes_in_range(lo, hi).tolist())
    small = [int(q) for q in _eratosthenes(2, 114)]
    want = {n for n in range(lo | 1, hi, 2) if all(n % q for q in small) and py_sprp2(n)}
    assert got == want


def test_bsgs_params_invariants():
    for L in [1, 500, 10**5, 10**6, 7 * 10**6, 2**24, 10**8]:
        nmin = 1000
        nmax = nmin + L - 1
        prm = O.bsgs_params(nmin, nmax, 100)
        assert prm.Q in (1024, 2048, 4096) and prm.hsize == 2 * prm.Q
        assert prm.Q >= math.isqrt(L) or prm.Q == 4096
        assert prm.Q == 1024 or prm.Q // 2 < math.ceil(math.sqrt(L))
        assert prm.M * prm.Q >= L > (prm.M - 1) * prm.Q
        assert prm.QQ == 2 * prm.Q and prm.MM * prm.QQ >= L > (prm.MM - 1) * prm.QQ
    small = O.bsgs_params(0, 10**7, 10, lmemsize=16384)
    assert small.hsize * 4 + 16 <= 16384
    with pytest.raises(ValueError):
        O.bsgs_params(0, 10**7, 10, lmemsize=2048)


def _dlog_set(k, c, b, p):
    """n in [0, ord) with k*b^n + c == 0 (mod p), and ord."""
    sols, x, n = [], k % p, 0
    while True:
        if (x + c) % p == 0:
            sols.append(n)
        x = x * b % p
        n += 1
        if x == k % p:
            return sols, n


@pytest.mark.parametrize("base", [2, 3, 5, 6, 10, 7])
def test_setup_prefilter_brute_force(base):
    rng = random.Random(base)
    ks = []
    while len(ks) < 12:
        k = rng.randrange(3, 10**5)
        if math.gcd(k, base) == 1 and (k % 2) != (base % 2) and k not in ks:
            ks.append(k)
    cs = [1 if i % 2 else -1 for i in range(len(ks))]
    job = O.make_job(base, ks, cs, 1000, 5000)
    primes = [p for p in O.primes_in_range(127, 12000) if py_is_prime(int(p))]
    Q = 1024
    nclass = [0, 0, 0, 0]
    for p in primes[::3]:
        p = int(p)
        pr, kd = O.setup_prime(job, p, Q)
        if pr is None:
            continue
        assert pr.one == R64 % p and pr.mbase == mont(base % p, p)
        assert pr.gQ_inv == mont(pow(pow(base, Q, p), -1, p), p)
        got = {kidx: (hadj, par) for hadj, par, kidx in kd}
        for i, (k, c) in enumerate(zip(ks, cs)):
            sols, order = _dlog_set(k, c, base, p)
            if i not in got:
                assert not sols, (p, k, c)
                nclass[0] += 1
                continue
            hadj, par = got[i]
            assert hadj == mont((-c) * pow(k, -1, p) % p, p)
            nclass[par] += 1
            if par == 2:
                assert all(n % 2 == 0 for n in sols) and order % 2 == 0
            elif par == 3:
                assert all(n % 2 == 1 for n in sols) and order % 2 == 0
            if pow(base, (p - 1) // 2, p) == p - 1 and sols:
                assert par in (2, 3)
    assert all(nclass[c] > 0 for c in (0, 1, 2, 3)) or base in (7,)


def _brute_elim(base, ks, cs, nmin, nmax, bm, wps, primes):
    """Terms (kidx, n) with bit set and some p in primes dividing them."""
    from util import bitmap_terms
    terms = bitmap_terms(bm, len(ks), wps, nmin)
    out = set()
    pr = np.array(primes, dtype=object)
    for (i, n) in terms:
        k, c = ks[i], cs[i]
        for p in primes:
            if (k * pow(base, n, p) + c) % p == 0:
                out.add((i, n))
                break
    return out


@pytest.mark.parametrize("case", [
    (2, [3, 5, 7, 9, 11], [1, -1, 1, -1, 1], 100, 1900, 127, 6000),      # order path for many small p
    (2, [1001, 3003, 75], [-1, 1, 1], 4000, 9000, 10**6, 10**6 + 3000),
    (3, [2, 4, 8, 10], [1, -1, -1, 1], 200, 3000, 200, 4000),
    (5, [4, 6], [1, -1], 10, 2500, 130, 3000),
    (10, [3, 7, 9], [1, 1, -1], 100, 2100, 100, 3000),
])
def test_sieve_brute_force(case):
    from util import bitmap_terms, full_bitmap
    base, ks, cs, nmin, nmax, pmin, pmax = case
    bm, wps = full_bitmap(len(ks), nmin, nmax, density=0.7, seed=5)
    job = O.make_job(base, ks, cs, nmin, nmax)
    out = bm.copy()
    cnt, facs, nf = O.sieve(job, pmin, pmax, out)
    primes = [int(p) for p in O.primes_in_range(pmin, pmax)]
    elim = bitmap_terms(bm, len(ks), wps, nmin) - bitmap_terms(out, len(ks), wps, nmin)
    assert elim == _brute_elim(base, ks, cs, nmin, nmax, bm, wps, primes)
    assert cnt.claims == len(elim) == len(facs)
    assert cnt.primecount == len(primes)
    for p, n, i in facs:
        assert (ks[i] * pow(base, n, p) + cs[i]) % p == 0
    # duplicates: every hit of every search window, minus the successful claims
    prm = O.bsgs_params(nmin, nmax, len(ks))
    hits = 0
    for p in primes:
        pr, kd = O.setup_prime(job, p, prm.Q)
        if pr is None:
            continue
        for hadj, par, i in kd:
            Ql, Ml, stride = (prm.Q, prm.M, 1) if par == 1 else (prm.QQ, prm.MM, 2)
            nstart = nmin + (1 if par == 3 else 0)
            order = next((j for j in range(stride, Ql, stride) if pow(base, j, p) == 1), 0)
            top = nmax if order else nstart + Ml * Ql - 1
            for n in range(nstart, top + 1, stride):
                if (ks[i] * pow(base, n, p) + cs[i]) % p == 0:
                    hits += 1  # a claim: successful, or a duplicate (bit clear, or n > NMAX)
    assert cnt.claims + cnt.dups == hits
    if case[0] == 2 and case[5] == 127:
        assert cnt.order_lists > 0


def test_is_prime_and_verify_factor():
    L = O.lib()
    rng = random.Random(7)
    for n in list(range(3, 3000, 2)) + A001262 + [rng.randrange(2**40, 2**64) | 1 for _ in range(300)]:
        assert L.or_is_prime(n) == int(py_is_prime(n)), n
    # the reference's example lines (cl_sieve.cpp:826-827)
    assert L.or_verify_factor(1000133, 91549, 54378870, 1, 2) == 1
    assert L.or_verify_factor(1000669, 131179, 72017282, -1, 2) == 0
    assert L.or_verify_factor(1000669, 131179, 72017282, 1, 2) == 1
    for _ in range(300):
        p = rng.randrange(3, 2**62) | 1
        while not py_is_prime(p):
            p += 2
        b = rng.choice([2, 3, 5, 6, 10, 12345])
        k = rng.randrange(1, 2**32)
        n = rng.randrange(1, 2**31)
        c = rng.choice([1, -1])
        want = int((k * pow(b, n, p) + c) % p == 0)
        assert L.or_verify_factor(p, k, n, c, b) == want
        # construct a true factor: n with k*b^n == -c by brute force for small p
    assert L.or_verify_factor(2047, 1, 11, -1, 2) == -1


@pytest.mark.skipif(O.ref_lib() is None, reason="oracle/_ref not built (reference tree absent)")
def test_verify_matches_reference_build():
    L, Rf = O.lib(), O.ref_lib()
    rng = random.Random(8)
    cases = [(1000133, 91549, 54378870, 1, 2), (1000669, 131179, 72017282, -1, 2),
             (1000669, 131179, 72017282, 1, 2), (2047, 1, 11, -1, 2)]
    for _ in range(400):
        p = rng.randrange(3, 2**64) | 1
        cases.append((p, rng.randrange(1, 2**32), rng.randrange(1, 2**31), rng.choice([1, -1]),
                      rng.choice([2, 3, 5, 7, 10])))
    for n in A001262[:10]:
        cases.append((n, 3, 5, 1, 2))
    for (p, k, n, c, b) in cases:
        assert L.or_verify_factor(p, k, n, c, b) == Rf.ref_verify_factor(p, k, n, c, b)
    for n in list(range(3, 5000, 2)) + A001262:
        assert L.or_is_prime(n) == Rf.ref_is_prime(n)


# --- composite 2-PRPs ---------------------------------------------------------
#
# The generator emits base-2 strong pseudoprimes with no factor <= 113 (42799 =
# 127*337 is the first), and the setup runs on them as on primes.  For an odd
# modulus the Montgomery arithmetic is exact modular arithmetic, so setup.cl's
# decisions (setup.cl:137-208) have a closed form in plain integers:
#   skipped      iff (prod |k| * b^Q)^(p-1) != 1 (mod p), or every k is impossible
#   tests        d in {3,4,5,7,8,9,11,13,17,19,23,29} with d | p-1 and b^((p-1)/d) == 1
#   rg           b^((p-1)/2): +1, -1 or 0 (anything else)
#   class of k   h = -c * k^-1 mod p: 0 if some test has h^((p-1)/d) != 1; with
#                ra = h^((p-1)/2): rg == -1 -> 2 (ra == 1), 3 (ra == -1), else 0;
#                rg == 1 and ra == -1 -> 0; otherwise 1
# (Euler's criterion is not sound for a composite p: these decisions are the
# reference's, not the truth, and the sieve must reproduce them.)


def _composite_2prps(limit):
    out = []
    for n in range(115, limit, 2):
        if any(n % q == 0 for q in (3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79,
                                    83, 89, 97, 101, 103, 107, 109, 113)):
            continue
        d, s = n - 1, 0
        while d % 2 == 0:
            d //= 2
            s += 1
        x = pow(2, d, n)
        ok = x in (1, n - 1)
        for _ in range(s - 1):
            if ok:
                break
            x = x * x % n
            ok = x == n - 1
        if ok and not py_is_prime(n):
            out.append(n)
    return out


COMPOSITES = _composite_2prps(1000000)


def _setup_closed_form(base, ks, cs, p, Q):
    """None if the reference skips p, else {kidx: (hadj in Montgomery form, class)}."""
    X = pow(base, Q, p)
    for k in ks:
        X = X * k % p
    if pow(X, p - 1, p) != 1:
        return None
    pm = p - 1
    tests = [pm // d for d in (3, 4, 5, 7, 8, 9, 11, 13, 17, 19, 23, 29) if pm % d == 0 and pow(base, pm // d, p) == 1]
    g = pow(base, pm // 2, p)
    rg = 1 if g == 1 else -1 if g == p - 1 else 0
    out = {}
    for i, (k, c) in enumerate(zip(ks, cs)):
        h = (-c) * pow(k, -1, p) % p
        if any(pow(h, e, p) != 1 for e in tests):
            continue
        ra = pow(h, pm // 2, p)
        if rg == -1:
            cls = 2 if ra == 1 else 3 if ra == p - 1 else 0
        elif rg == 1 and ra == p - 1:
            cls = 0
        else:
            cls = 1
        if cls:
            out[i] = (mont(h, p), cls)
    return out or None


def test_composite_2prps_are_what_the_generator_emits():
    assert COMPOSITES[:4] == [42799, 49141, 88357, 90751]
    emitted = set(int(x) for x in O.primes_in_range(115, 1000000))
    assert set(COMPOSITES) <= emitted
    assert all(py_is_prime(int(x)) for x in emitted - set(COMPOSITES))


@pytest.mark.parametrize("base", [2, 3, 5, 6, 7, 10])
def test_setup_composite_2prps_closed_form(base):
    """or_setup_prime on composite 2-PRPs against the closed form above: the
    skip decision, every k's class and its target, for k's coprime to p and k's
    sharing a factor with p (the invertibility test must skip those)."""
    rng = random.Random(100 + base)
    checked = skipped = 0
    classes = set()
    for p in COMPOSITES:
        facs = [q for q in range(127, int(math.isqrt(p)) + 1, 2) if p % q == 0]
        for trial in range(6):
            ks = []
            while len(ks) < 8:
                k = rng.randrange(3, 10**6)
                if math.gcd(k, base) == 1 and k not in ks:
                    ks.append(k)
            if trial == 5:
                ks[3] = facs[0] * rng.randrange(1, 50)  # not invertible mod p
            cs = [1 if i % 2 else -1 for i in range(len(ks))]
            job = O.make_job(base, ks, cs, 1000, 5000)
            for Q in (1024, 4096):
                pr, kd = O.setup_prime(job, p, Q)
                want = _setup_closed_form(base, ks, cs, p, Q)
                if want is None:
                    assert pr is None or not kd, (p, ks, Q)
                    skipped += 1
                    continue
                assert pr is not None
                got = {kidx: (hadj, par) for hadj, par, kidx in kd}
                assert got == want, (p, base, ks, Q)
                checked += 1
                classes |= {par for (_, par) in want.values()}
    assert checked >= 20 and skipped >= 50
    if base in (2, 6):  # b a quadratic non-residue mod some composites: the parity classes occur
        assert classes == {1, 2, 3}


@pytest.mark.parametrize("base,pmin,pmax", [(2, 42000, 50000), (3, 88000, 92000), (10, 42000, 43000)])
def test_sieve_composite_2prps_brute_force(base, pmin, pmax):
    """The sieve over ranges holding composite 2-PRPs: every claim's modulus
    divides its term, and the terms cleared by a composite p are exactly the
    ones its closed-form classes admit (the solutions of k*b^n + c == 0 mod p
    whose k the reference's tests keep, of the allowed parity)."""
    from util import bitmap_terms, full_bitmap
    ks = [3, 5, 7, 11, 13, 17] if base % 2 == 0 else [2, 4, 8, 10, 14, 16]
    ks = [k for k in ks if math.gcd(k, base) == 1]
    cs = [1 if i % 2 else -1 for i in range(len(ks))]
    nmin, nmax = 1000, 9000
    bm, wps = full_bitmap(len(ks), nmin, nmax)
    comps = [p for p in COMPOSITES if pmin <= p < pmax]
    assert comps
    for p in comps:
        # the composite alone: its claims are the closed form's solutions
        job = O.make_job(base, ks, cs, nmin, nmax)
        out = bm.copy()
        cnt, facs, nf = O.sieve(job, p, p + 1, out)
        assert cnt.primecount == 1
        prm = O.bsgs_params(nmin, nmax, len(ks))
        want = _setup_closed_form(base, ks, cs, p, prm.Q)
        sols = set()
        for i, (hm, cls) in (want or {}).items():
            for n in range(nmin, nmax + 1):
                if (ks[i] * pow(base, n, p) + cs[i]) % p == 0 and (cls == 1 or n % 2 == cls - 2):
                    sols.add((i, n))
        assert {(i, n) for (_, n, i) in facs} == sols, (p, base)
        for q, n, i in facs:
            assert q == p and (ks[i] * pow(base, n, p) + cs[i]) % p == 0
    # the whole range: every claim's modulus divides its term
    job = O.make_job(base, ks, cs, nmin, nmax)
    out = bm.copy()
    cnt, facs, nf = O.sieve(job, pmin, pmax, out)
    assert facs
    for q, n, i in facs:
        assert (ks[i] * pow(base, n, q) + cs[i]) % q == 0
    assert bitmap_terms(bm, len(ks), wps, nmin) - bitmap_terms(out, len(ks), wps, nmin) == {(i, n) for (_, n, i) in facs}


def test_profile_range_solves_its_equation():
    """or_profile_range (cl_sieve.cpp:693-735): x = start + range solves
    x / ln x = 256 * CU + start / ln start, so the segment holds about 256 * CU
    primes by the prime number theorem; psize = 1.5 * 256 * CU; the range is
    capped at 4294900000.  Checked against the equation and against an actual
    prime count."""
    import ctypes as C
    L = O.lib()
    for start in (10**9, 2**40, 2**50, 2**63 - 2**40):
        for cu in (8, 104, 256):
            r, ps = C.c_uint64(), C.c_uint32()
            L.or_profile_range(start, cu, C.byref(r), C.byref(ps))
            x = start + r.value
            target = 256 * cu + start / math.log(start)
            assert abs(x / math.log(x) - target) < 1e-6 * target + 2.0
            assert ps.value == int(1.5 * 256 * cu)
    r, ps = C.c_uint64(), C.c_uint32()
    L.or_profile_range(10**9, 8, C.byref(r), C.byref(ps))
    n = len(O.primes_in_range(10**9, 10**9 + r.value))
    assert abs(n - 2048) < 0.1 * 2048
    L.or_profile_range(10**18, 10**6, C.byref(r), C.byref(ps))
    assert r.value == 4294900000
