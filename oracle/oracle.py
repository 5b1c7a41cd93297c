"""ctypes binding of the CPU oracle (oracle/rs_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the timed CPU baseline,
never by the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libref_verify.so")
MAX_SEQ = 100


class OrJob(C.Structure):
    _fields_ = [
        ("base", C.c_uint32),
        ("kcount", C.c_int32),
        ("k", C.c_uint32 * MAX_SEQ),
        ("c", C.c_int32 * MAX_SEQ),
        ("nmin", C.c_uint32),
        ("nmax", C.c_uint32),
        ("words_per_seq", C.c_uint64),
        ("k31_compat", C.c_int32),
        ("lmemsize", C.c_uint32),
    ]


class OrFactor(C.Structure):
    _fields_ = [("p", C.c_uint64), ("n", C.c_uint32), ("kidx", C.c_uint32)]


class OrCounts(C.Structure):
    _fields_ = [
        ("primecount", C.c_uint64),
        ("claims", C.c_uint64),
        ("dups", C.c_uint64),
        ("skipped", C.c_uint64),
        ("n_full", C.c_uint64),
        ("n_even", C.c_uint64),
        ("n_odd", C.c_uint64),
        ("order_lists", C.c_uint64),
        ("mmul_work", C.c_uint64),
    ]


class OrKdata(C.Structure):
    _fields_ = [("hadj", C.c_uint64), ("parity", C.c_int32), ("kidx", C.c_int32)]


class OrPrime(C.Structure):
    _fields_ = [("p", C.c_uint64), ("q", C.c_uint64), ("one", C.c_uint64),
                ("mbase", C.c_uint64), ("nmo", C.c_uint64), ("gQ_inv", C.c_uint64)]


class OrParams(C.Structure):
    _fields_ = [("Q", C.c_uint32), ("hsize", C.c_uint32), ("M", C.c_uint32),
                ("QQ", C.c_uint32), ("MM", C.c_uint32), ("cache_k", C.c_uint32)]


def build():
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if os.path.isfile("/root/reference/verify_factor.c"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.isfile(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        u64, u32, i32 = C.c_uint64, C.c_uint32, C.c_int32
        sig = {
            "or_invert": (u64, [u64]),
            "or_mmul": (u64, [u64, u64, u64, u64]),
            "or_add": (u64, [u64, u64, u64]),
            "or_powmodlg": (u64, [u64, u64, u64, u64]),
            "or_basepowmodlg": (u64, [u64, u64, u32, u64, u64]),
            "or_powmodsm": (u64, [u64, u32, u64, u64, u64]),
            "or_basepowmodsm": (u64, [u64, u32, u32, u64, u64, u64]),
            "or_strong_prp_two": (C.c_int, [u64]),
            "or_find_wheel_offset": (None, [C.POINTER(u64), C.POINTER(i32)]),
            "or_bitsieve_word": (u32, [u32, u32]),
            "or_getsegprimes": (C.c_size_t, [u64, u64, i32, C.POINTER(u64), C.c_size_t]),
            "or_addsmallprimes": (C.c_size_t, [u64, u64, C.POINTER(u64), C.c_size_t]),
            "or_primes_in_range": (C.c_size_t, [u64, u64, u64, C.POINTER(u64), C.c_size_t]),
            "or_bsgs_params_for": (C.c_int, [u32, u32, i32, u32, C.POINTER(OrParams)]),
            "or_profile_range": (None, [u64, u32, C.POINTER(u64), C.POINTER(u32)]),
            "or_setup_prime": (C.c_int, [C.POINTER(OrJob), u64, u32, C.POINTER(OrPrime), C.POINTER(OrKdata)]),
            "or_sieve": (C.c_int64, [C.POINTER(OrJob), u64, u64, C.POINTER(u32), C.POINTER(OrFactor), u64,
                                     C.POINTER(OrCounts), C.c_int, C.POINTER(u64)]),
            "or_is_prime": (C.c_int, [u64]),
            "or_verify_factor": (C.c_int, [u64, u64, u32, i32, u32]),
            "or_sizeof_counts": (u32, []),
            "or_sizeof_job": (u32, []),
            "or_sizeof_factor": (u32, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        assert L.or_sizeof_counts() == C.sizeof(OrCounts)
        assert L.or_sizeof_job() == C.sizeof(OrJob)
        assert L.or_sizeof_factor() == C.sizeof(OrFactor)
        _lib = L
    return _lib


def ref_lib():
    """The reference's verify_factor.c built under oracle/_ref, or None."""
    global _ref
    if _ref is None and os.path.isfile(REF_LIB_PATH):
        R = C.CDLL(REF_LIB_PATH)
        u64, u32, i32 = C.c_uint64, C.c_uint32, C.c_int32
        for name, res, args in [("ref_invert", u64, [u64]), ("ref_m_mul", u64, [u64, u64, u64, u64]),
                                ("ref_add", u64, [u64, u64, u64]),
                                ("ref_verify_factor", C.c_int, [u64, u64, u32, i32, u32]),
                                ("ref_is_prime", C.c_int, [u64])]:
            f = getattr(R, name)
            f.restype = res
            f.argtypes = args
        _ref = R
    return _ref


def make_job(base, ks, cs, nmin, nmax, k31_compat=False, lmemsize=65536):
    """An or_job for sequences (ks[i], cs[i]); nmin is the bitmap NMIN (even)."""
    assert len(ks) == len(cs) and 1 <= len(ks) <= MAX_SEQ
    j = OrJob()
    j.base = base
    j.kcount = len(ks)
    for i, (k, c) in enumerate(zip(ks, cs)):
        j.k[i] = k
        j.c[i] = c
    j.nmin = nmin
    j.nmax = nmax
    j.words_per_seq = (nmax - nmin + 1 + 31) // 32
    j.k31_compat = 1 if k31_compat else 0
    j.lmemsize = lmemsize
    return j


def primes_in_range(pmin, pmax, seg=1 << 20):
    L = lib()
    n = L.or_primes_in_range(pmin, pmax, seg, None, 0)
    out = np.zeros(max(n, 1), dtype=np.uint64)
    L.or_primes_in_range(pmin, pmax, seg, out.ctypes.data_as(C.POINTER(C.c_uint64)), n)
    return out[:n]


def setup_prime(job, p, Q):
    L = lib()
    kd = (OrKdata * MAX_SEQ)()
    pr = OrPrime()
    n = L.or_setup_prime(C.byref(job), p, Q, C.byref(pr), kd)
    if n < 0:
        return None, []
    return pr, [(kd[i].hadj, kd[i].parity, kd[i].kidx) for i in range(n)]


def bsgs_params(nmin, nmax, kcount, lmemsize=65536):
    prm = OrParams()
    rc = lib().or_bsgs_params_for(nmin, nmax, kcount, lmemsize, C.byref(prm))
    if rc:
        raise ValueError("insufficient local memory")
    return prm


def sieve(job, pmin, pmax, bitmap, fac_cap=1 << 20, nthreads=1):
    """Run the oracle sieve; bitmap (np.uint32) is cleared in place.
    Returns (counts, factors as list of (p, n, kidx), nfactors_total)."""
    L = lib()
    assert bitmap.dtype == np.uint32 and bitmap.flags.c_contiguous
    fac = (OrFactor * max(fac_cap, 1))()
    cnt = OrCounts()
    nfac = C.c_uint64(0)
    rc = L.or_sieve(C.byref(job), pmin, pmax, bitmap.ctypes.data_as(C.POINTER(C.c_uint32)), fac, fac_cap,
                    C.byref(cnt), nthreads, C.byref(nfac))
    if rc < 0:
        raise ValueError("oracle sieve parameter error")
    n = min(nfac.value, fac_cap)
    return cnt, [(fac[i].p, fac[i].n, fac[i].kidx) for i in range(n)], nfac.value


def counts_dict(cnt):
    return {name: getattr(cnt, name) for name, _ in OrCounts._fields_}
