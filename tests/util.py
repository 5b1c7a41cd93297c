"""Shared helpers of the test suite: seeded job data and comparisons."""
import numpy as np

import oracle as O


def full_bitmap(kcount, nmin, nmax, density=1.0, seed=0, nlo=None):
    """Bitmap with every n in [nlo or nmin, nmax] set (or a seeded fraction)."""
    wps = (nmax - nmin + 1 + 31) // 32
    bits = np.zeros((kcount, wps * 32), dtype=bool)
    lo = nmin if nlo is None else nlo
    rng = np.random.default_rng(seed)
    for i in range(kcount):
        sel = np.ones(nmax - lo + 1, dtype=bool) if density >= 1.0 else rng.random(nmax - lo + 1) < density
        bits[i, lo - nmin:nmax - nmin + 1] = sel
    return np.packbits(bits.reshape(kcount, wps, 32)[:, :, ::-1], axis=-1, bitorder="big").view(">u4").astype(
        np.uint32).reshape(-1), wps


def bitmap_terms(bm, kcount, wps, nmin):
    """Set of (kidx, n) whose bit is set."""
    b = np.unpackbits(bm.view(np.uint8).reshape(kcount, wps * 4), axis=1, bitorder="little")
    ks, offs = np.nonzero(b)
    return set(zip(ks.tolist(), (offs + nmin).tolist()))


def cleared_terms(bm_in, bm_out, wps, nmin):
    """Set of (kidx, n) whose bit is set in bm_in and clear in bm_out, found
    from the changed words only (fast for production-size bitmaps)."""
    diff = np.asarray(bm_in, np.uint32) & ~np.asarray(bm_out, np.uint32)
    out = set()
    for w in np.nonzero(diff)[0].tolist():
        v = int(diff[w])
        kidx, off = divmod(w, wps)
        while v:
            b = (v & -v).bit_length() - 1
            out.add((kidx, nmin + 32 * off + b))
            v &= v - 1
    return out


def oracle_run(base, ks, cs, nmin, nmax, bm, pmin, pmax, k31=False, nthreads=1, fac_cap=1 << 22):
    job = O.make_job(base, list(ks), list(cs), nmin, nmax, k31_compat=k31)
    b = bm.copy()
    cnt, facs, nf = O.sieve(job, pmin, pmax, b, fac_cap=fac_cap, nthreads=nthreads)
    return O.counts_dict(cnt), facs, b
