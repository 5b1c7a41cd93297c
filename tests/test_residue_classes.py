"""The residue-class rule of the setup's class tables (RS_ALGO_CLASSES),
checked by brute force.

For L = gcd(p - 1, 720720) * Ld and z = b^((p-1)/L), every n with b^n = h
(mod p) satisfies z^n = y with y = h^((p-1)/L), so n = x (mod s) for the
x < s = ord(z) with z^x = y, and there is no n when y is not a power of z.
Ld is one prime r in [17, 1021] dividing p - 1 with b^((p-1)/r) != 1 (the
largest whose stride 2 * gcd(p - 1, 720720) * r stays below 2^24), or 1; its
order is Ld when (b^((p-1)/Ld))^Ld = 1, else the rule keeps the parity split.
The setup keeps the reference's own tests (setup.cl:59-87: the parity classes and the
power-residue tests) and only narrows the n searched to that class, so the
solutions found in [nmin, nmax] must be exactly those of the reference's
class.  Composite base-2 strong pseudoprimes are included: there the
reference's Euler-criterion tests are not sound, and the rule must still
leave its result unchanged.  Restated here in Python as the kernel states it
(rs_kernels.hip, setup_prime); the GPU parity tests check the kernel itself.
"""
import math
import random

import pytest

PRD = [3, 4, 5, 7, 8, 9, 11, 13, 17, 19, 23, 29]  # setup.cl:22-48
RD = [r for r in range(17, 1022) if all(r % q for q in range(2, math.isqrt(r) + 1))]


def _sprp2(n):
    d, t = n - 1, 0
    while d % 2 == 0:
        d //= 2
        t += 1
    x = pow(2, d, n)
    if x in (1, n - 1):
        return True
    for _ in range(t - 1):
        x = x * x % n
        if x == n - 1:
            return True
    return False


def _sign(a, p):
    return 1 if a == 1 else (-1 if a == p - 1 else 0)


def _reference_classes(p, b, hs):
    """cls per target: 0 no search, 1 all n, 2 even n, 3 odd n (setup.cl:59-87, 180-208)."""
    pm = p - 1
    resg = _sign(pow(b, pm // 2, p), p)
    dpass = [d for d in PRD if pm % d == 0 and pow(b, pm // d, p) == 1]
    out = []
    for h in hs:
        cls = 1
        for d in dpass:
            if _sign(pow(h, pm // d, p), p) != 1:
                cls = 0
        t = _sign(pow(h, pm // 2, p), p)
        if cls:
            if resg == -1:
                cls = 2 if t == 1 else (3 if t == -1 else 0)
            elif resg == 1 and t == -1:
                cls = 0
        out.append(cls)
    return out


def _class_residues(p, b, hs, clss):
    """(S, residues n mod S per target) as the setup computes them."""
    pm = p - 1
    v2 = (pm & -pm).bit_length() - 1
    La = (1 << min(v2, 4)) * (5 if pm % 5 == 0 else 1)                       # gcd(p-1, 80)
    Lb = (9 if pm % 9 == 0 else 3 if pm % 3 == 0 else 1) * (7 if pm % 7 == 0 else 1)  # gcd(p-1, 63)
    Lc = (11 if pm % 11 == 0 else 1) * (13 if pm % 13 == 0 else 1)                    # gcd(p-1, 143)
    L0 = La * Lb * Lc                                                          # gcd(p-1, 720720)
    # the fourth part: the primes of RD dividing p - 1 (17..29 always, the
    # others while their product stays below 2^32), the largest r with
    # b^((p-1)/r) != 1 and 2 * L0 * r < 2^24
    D, Ld = 1, 1
    for j, r in enumerate(RD):
        if pm % r == 0 and (j < 4 or D * r < 2**32):
            D *= r
    for r in RD:
        if D % r == 0 and pow(b, pm // r, p) != 1 and 2 * L0 * r < 2**24:
            Ld = r
    L = L0 * Ld
    z = pow(b, pm // L, p)
    # n mod the orders of z^(L/La), z^(L/Lb), z^(L/Lc) and z^(L/Ld) (each
    # exact: the first i with z_a^i = 1), combined by CRT when they are
    # pairwise coprime (always for a prime p); otherwise (a composite p) only
    # the parity split
    za, zb, zc, zd = (pow(z, L // La, p), pow(z, L // Lb, p), pow(z, L // Lc, p), pow(z, L // Ld, p))
    sa = next((i for i in range(1, La + 1) if pow(za, i, p) == 1), 0)
    sb = next((i for i in range(1, Lb + 1) if pow(zb, i, p) == 1), 0)
    # Lc = 1: no constraint mod sc = 1 and no third table (as the kernel); Ld likewise
    sc = 1 if Lc == 1 else next((i for i in range(1, Lc + 1) if pow(zc, i, p) == 1), 0)
    sd = 1 if Ld == 1 else (Ld if pow(zd, Ld, p) == 1 else 0)
    coprime = math.gcd(sa, sb) == 1 and math.gcd(sa * sb, sc) == 1 and math.gcd(sa * sb * sc, sd) == 1
    s = sa * sb * sc * sd if sa and sb and sc and sd and coprime else 0
    any_par = any(c >= 2 for c in clss)
    S = 2 if s == 0 else (2 * s if (s & 1) and any_par else s)
    ta = {pow(za, i, p): i for i in range(sa)} if s else {}
    tb = {pow(zb, i, p): i for i in range(sb)} if s else {}
    tc = {pow(zc, i, p): i for i in range(sc)} if s and Lc > 1 else {}
    td = {pow(zd, i, p): i for i in range(sd)} if s and Ld > 1 else {}

    def dlog(h):
        """x < s with z^x = y (its residues mod sa, sb, sc and sd), or None."""
        y = pow(h, pm // L, p)
        xa, xb = ta.get(pow(y, L // La, p)), tb.get(pow(y, L // Lb, p))
        xc = tc.get(pow(y, L // Lc, p)) if Lc > 1 else 0
        xd = td.get(pow(y, L // Ld, p)) if Ld > 1 else 0
        if xa is None or xb is None or xc is None or xd is None:
            return None
        return next(x for x in range(s) if x % sa == xa and x % sb == xb and x % sc == xc and x % sd == xd)
    res = []
    for h, cls in zip(hs, clss):
        rr = []
        if cls and s == 0:
            rr = [0, 1] if cls == 1 else [cls - 2]
        elif cls:
            x = dlog(h)
            if x is not None:
                if S == s:
                    if not (cls >= 2 and (x & 1) != cls - 2):
                        rr = [x]
                elif cls == 1:
                    rr = [x, x + s]
                else:
                    rr = [x if (x & 1) == cls - 2 else x + s]
        res.append(rr)
    return S, res


def _cases(limit, nprimes, seed):
    rng = random.Random(seed)
    cands = [p for p in range(1025, limit, 2) if _sprp2(p)]
    comp = [p for p in cands if any(p % q == 0 for q in range(3, math.isqrt(p) + 1, 2))]
    prime = [p for p in cands if p not in comp]
    return comp, rng.sample(prime, nprimes)


def test_pseudoprimes_are_present():
    comp, _ = _cases(60000, 1, 1)
    assert comp[:4] == [2047, 3277, 4033, 4681]


@pytest.mark.parametrize("base", [2, 3, 5, 6, 7, 10])
def test_residue_classes_keep_the_reference_solutions(base):
    comp, prime = _cases(60000, 40, base)
    rng = random.Random(base)
    narrowed = checked = fourth = 0
    for p in comp + prime:
        if math.gcd(base, p) != 1:
            continue
        ks = [k for k in (rng.randrange(2, 10**6) for _ in range(8)) if math.gcd(k, p) == 1]
        cs = [rng.choice((1, -1)) for _ in ks]
        hs = [(-c * pow(k, -1, p)) % p for k, c in zip(ks, cs)]
        clss = _reference_classes(p, base, hs)
        S, res = _class_residues(p, base, hs, clss)
        nmin, nmax = 1000, 1000 + p
        powb = [pow(base, nmin, p)]
        for _ in range(nmax - nmin):
            powb.append(powb[-1] * base % p)
        for h, cls, rr in zip(hs, clss, res):
            if not cls:
                continue
            sols = [nmin + i for i, v in enumerate(powb) if v == h]
            ref = {n for n in sols if cls == 1 or n % 2 == cls - 2}
            ours = {n for n in sols if n % S in rr}
            assert ours == ref, (p, base, h, cls, S, rr)
            checked += 1
            narrowed += len(rr) * 2 < S * (2 if cls == 1 else 1)
            fourth += any(S % r == 0 for r in RD)
    assert checked > 100 and narrowed > checked // 4 and fourth > checked // 10
