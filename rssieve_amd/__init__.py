"""rssieve_amd: Python binding of librssieve.so (include/rssieve.h,
include/rssieve_io.h), the HIP/gfx950 sieve for k*b^n+-1.

The library is the product; this module only loads it with ctypes and moves
numpy arrays across the C ABI.  Every compute call goes to the HIP kernels:
there is no CPU fallback, and loading fails loudly when the library is
missing.
"""
import atexit
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librssieve.so")
CLI_PATH = os.path.join(HERE, "rssieve")
MAX_SEQ = 100

RS_OK, RS_E_ARG, RS_E_HIP, RS_E_NOMEM, RS_E_OVERFLOW, RS_E_IO, RS_E_VERIFY = 0, -1, -2, -3, -4, -5, -6
ALGO_THREE_LISTS, ALGO_MERGED, ALGO_CLASSES = 0, 1, 2


class RSError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("rssieve error %d: %s" % (code, msg))
        self.code = code


class Factor(C.Structure):
    _fields_ = [("p", C.c_uint64), ("n", C.c_uint32), ("kidx", C.c_uint32)]


FACTOR_DTYPE = np.dtype([("p", np.uint64), ("n", np.uint32), ("kidx", np.uint32)])


class Counters(C.Structure):
    _fields_ = [
        ("primecount", C.c_uint64),
        ("claims", C.c_uint64),
        ("dups", C.c_uint64),
        ("nfactors", C.c_uint64),
        ("skipped", C.c_uint64),
        ("lists", C.c_uint64 * 3),
        ("order_lists", C.c_uint64),
        ("work_mmul", C.c_uint64),
        ("work_probe", C.c_uint64),
        ("max_seg_primes", C.c_uint64),
        ("seg_overflow", C.c_uint64),
    ]

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if name == "lists" else v
        return d


class Params(C.Structure):
    _fields_ = [
        ("algo", C.c_int32),
        ("Q", C.c_uint32),
        ("hsize", C.c_uint32),
        ("threads", C.c_uint32),
        ("seg_primes", C.c_uint32),
        ("blocks_per_cu", C.c_uint32),
        ("k31_compat", C.c_int32),
        ("record_factors", C.c_int32),
        ("factor_cap", C.c_uint64),
        ("queue_cap", C.c_uint32),
        ("reserved", C.c_uint32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class State(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("pmin", "pmax", "p", "checksum", "primecount", "factorcount",
                                          "last_trickle", "state_sum", "dupcount")] + [
        ("nmin", C.c_uint32), ("nmax", C.c_uint32), ("base", C.c_uint32), ("kcount", C.c_int32),
        ("klist", C.c_int32 * MAX_SEQ)]


_lib = None

_SIGS = {
    # rssieve.h
    "rs_open": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "rs_close": (None, [C.c_void_p]),
    "rs_last_error": (C.c_char_p, []),
    "rs_device_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_char_p, C.c_size_t]),
    "rs_params_reference": (C.c_int, [C.c_uint32, C.c_uint32, C.c_int32, C.c_uint32, C.c_uint32, C.POINTER(Params)]),
    "rs_params_derive": (C.c_int, [C.POINTER(Params), C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32),
                                   C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "rs_job_create": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int32, C.POINTER(C.c_uint32), C.POINTER(C.c_int32),
                                C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32), C.c_uint64, C.POINTER(Params),
                                C.POINTER(C.c_void_p)]),
    "rs_job_destroy": (None, [C.c_void_p]),
    "rs_job_view": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "rs_job_sequence": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_int32)]),
    "rs_sieve": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64]),
    "rs_sieve_noalloc": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64]),
    "rs_sieve_batch": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "rs_sync": (C.c_int, [C.c_void_p]),
    "rs_counters_read": (C.c_int, [C.c_void_p, C.POINTER(Counters)]),
    "rs_factors_drain": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "rs_bitmap_read": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.c_uint64]),
    "rs_bitmap_write": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.c_uint64]),
    "rs_job_kernel_time": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64),
                                     C.POINTER(C.c_double)]),
    "rs_job_timer_reset": (C.c_int, [C.c_void_p]),
    "rs_primegen": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64), C.c_uint64,
                              C.POINTER(C.c_uint64)]),
    "rs_peak_mad_u64_u32": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "rs_verify_factor": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint32, C.c_int32, C.c_uint32]),
    "rs_is_prime": (C.c_int, [C.c_uint64]),
    "rs_wheel_offset": (None, [C.POINTER(C.c_uint64), C.POINTER(C.c_int32)]),
    # rssieve_io.h
    "rs_abcd_read": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "rs_abcd_free": (None, [C.c_void_p]),
    "rs_abcd_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_int32), C.POINTER(C.c_uint32),
                               C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "rs_abcd_sequences": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_int32)]),
    "rs_abcd_bitmap": (C.POINTER(C.c_uint32), [C.c_void_p]),
    "rs_abcd_mark_used": (C.c_int, [C.c_void_p, C.c_uint32, C.c_char, C.c_uint32, C.c_int32]),
    "rs_abcd_write": (C.c_int, [C.c_char_p, C.c_uint32, C.c_int32, C.POINTER(C.c_uint32), C.POINTER(C.c_int32),
                                C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32), C.c_uint64, C.c_uint64]),
    "rs_presieve": (C.c_int, [C.c_uint32, C.c_int32, C.POINTER(C.c_uint32), C.POINTER(C.c_int32), C.c_uint32,
                              C.c_uint32, C.POINTER(C.c_uint32), C.c_uint64, C.c_uint64, C.c_int]),
    "rs_format_factor": (C.c_int, [C.c_char_p, C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32]),
    "rs_parse_factor": (C.c_int, [C.c_char_p, C.c_uint32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32),
                                  C.POINTER(C.c_char), C.POINTER(C.c_uint32)]),
    "rs_state_write": (C.c_int, [C.c_char_p, C.POINTER(State), C.POINTER(C.c_int)]),
    "rs_state_read": (C.c_int, [C.c_char_p, C.POINTER(State), C.POINTER(C.c_int)]),
    "rs_context_stream": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    # rssieve_mlp.h
    "rs_tune_candidates": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_int32, C.POINTER(Params), C.c_int,
                                     C.POINTER(C.c_int)]),
    "rs_model_load": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    "rs_model_free": (None, [C.c_void_p]),
    "rs_model_choose": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int32, C.c_uint64, C.c_uint32, C.c_double,
                                  C.POINTER(Params), C.c_int, C.c_uint32, C.POINTER(C.c_int)]),
    "rs_model_rank": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int32, C.c_uint64, C.c_uint32, C.c_double,
                                C.POINTER(Params), C.c_int, C.c_uint32, C.c_int, C.POINTER(C.c_int),
                                C.POINTER(C.c_int)]),
    "rs_model_split": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int32, C.c_uint64, C.c_uint32, C.c_double,
                                 C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "rs_autotune_split": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_uint32), C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_int)]),
    "rs_autotune": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(Params), C.c_int,
                              C.POINTER(C.c_double), C.POINTER(C.c_uint64), C.POINTER(C.c_int)]),
    "rs_job_reconfigure": (C.c_int, [C.c_void_p, C.POINTER(Params)]),
    "rs_mlp_create": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]),
    "rs_mlp_destroy": (None, [C.c_void_p]),
    "rs_mlp_num_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "rs_mlp_get_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "rs_mlp_set_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "rs_mlp_forward": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float)]),
    "rs_mlp_loss_grad": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                   C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "rs_mlp_adam_step": (C.c_int, [C.c_void_p, C.c_float]),
    "rs_mlp_train": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int,
                               C.c_int, C.c_int, C.c_float, C.c_uint64, C.POINTER(C.c_float)]),
}


def lib():
    """Load librssieve.so (built by rssieve_amd.build); raises if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.isfile(LIB_PATH):
            raise RuntimeError("librssieve.so not built: run python -m rssieve_amd.build (%s)" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return sorted(_SIGS)


def _check(rc):
    if rc != RS_OK:
        raise RSError(rc, lib().rs_last_error().decode(errors="replace"))
    return rc


def _u32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _i32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _u64p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


# --- host-side helpers (no device needed) ---------------------------------

def verify_factor(p, k, n, c, b):
    return lib().rs_verify_factor(p, k, n, c, b)


def is_prime(p):
    return bool(lib().rs_is_prime(p))


def wheel_offset(start):
    s = C.c_uint64(start)
    i = C.c_int32(0)
    lib().rs_wheel_offset(C.byref(s), C.byref(i))
    return s.value, i.value


def reference_params(nmin, nmax, kcount, lmemsize=65536, compute_units=256):
    p = Params()
    _check(lib().rs_params_reference(nmin, nmax, kcount, lmemsize, compute_units, C.byref(p)))
    return p


def derive(prm, nmin, nmax):
    M, QQ, MM = C.c_uint32(), C.c_uint32(), C.c_uint32()
    _check(lib().rs_params_derive(C.byref(prm), nmin, nmax, C.byref(M), C.byref(QQ), C.byref(MM)))
    return M.value, QQ.value, MM.value


def format_factor(p, k, b, n, c):
    buf = C.create_string_buffer(128)
    r = lib().rs_format_factor(buf, 128, p, k, b, n, c)
    if r < 0:
        _check(r)
    return buf.value.decode()


def parse_factor(line, base):
    p, k, n = C.c_uint64(), C.c_uint32(), C.c_uint32()
    s = C.c_char()
    _check(lib().rs_parse_factor(line.encode(), base, C.byref(p), C.byref(k), C.byref(s), C.byref(n)))
    return p.value, k.value, s.value.decode(), n.value


class ABCD:
    """A parsed ABCD file (rs_abcd_read)."""

    def __init__(self, path):
        h = C.c_void_p()
        _check(lib().rs_abcd_read(os.fsencode(path), C.byref(h)))
        self._h = h
        b, kc, nmin, nmax, wps, nt = (C.c_uint32(), C.c_int32(), C.c_uint32(), C.c_uint32(), C.c_uint64(),
                                      C.c_uint64())
        _check(lib().rs_abcd_info(h, C.byref(b), C.byref(kc), C.byref(nmin), C.byref(nmax), C.byref(wps),
                                  C.byref(nt)))
        self.base, self.kcount, self.nmin, self.nmax = b.value, kc.value, nmin.value, nmax.value
        self.words_per_seq, self.nterms = wps.value, nt.value
        self.k = np.zeros(self.kcount, np.uint32)
        self.c = np.zeros(self.kcount, np.int32)
        _check(lib().rs_abcd_sequences(h, _u32p(self.k), _i32p(self.c)))

    def bitmap(self):
        ptr = lib().rs_abcd_bitmap(self._h)
        n = self.kcount * self.words_per_seq
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy()

    def mark_used(self, k, sign, n, compat=False):
        return lib().rs_abcd_mark_used(self._h, k, sign.encode(), n, 1 if compat else 0)

    def close(self):
        if self._h:
            lib().rs_abcd_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def abcd_write(path, base, k, c, nmin, nmax, bitmap, words_per_seq, sieved_to=0):
    k = np.ascontiguousarray(k, np.uint32)
    c = np.ascontiguousarray(c, np.int32)
    bitmap = np.ascontiguousarray(bitmap, np.uint32)
    _check(lib().rs_abcd_write(os.fsencode(path), base, len(k), _u32p(k), _i32p(c), nmin, nmax, _u32p(bitmap),
                               words_per_seq, sieved_to))


def presieve(base, k, c, nmin, nmax, bitmap, words_per_seq, bound, nthreads=8):
    """Clear, in place, terms with a prime factor below 'bound' (CPU, exact)."""
    k = np.ascontiguousarray(k, np.uint32)
    c = np.ascontiguousarray(c, np.int32)
    assert bitmap.dtype == np.uint32 and bitmap.flags.c_contiguous
    _check(lib().rs_presieve(base, len(k), _u32p(k), _i32p(c), nmin, nmax, _u32p(bitmap), words_per_seq, bound,
                             nthreads))
    return bitmap


# --- device ---------------------------------------------------------------

_live_contexts = weakref.WeakSet()


@atexit.register
def _close_live_contexts():
    # contexts (and their jobs and MLPs) still open at exit are closed while
    # the interpreter and the HIP runtime are both intact
    for ctx in list(_live_contexts):
        try:
            ctx.close()
        except Exception:
            pass


class Context:
    """A HIP device context (rs_open)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().rs_open(device, C.byref(h)))
        self._h = h
        _live_contexts.add(self)
        # jobs and MLPs live on the context's stream and device memory: they
        # are closed before it
        self._children = weakref.WeakSet()
        cu, lds = C.c_uint32(), C.c_uint32()
        name = C.create_string_buffer(256)
        _check(lib().rs_device_info(h, C.byref(cu), C.byref(lds), name, 256))
        self.compute_units, self.lds_per_block, self.name = cu.value, lds.value, name.value.decode()

    def sync(self):
        _check(lib().rs_sync(self._h))

    def primegen(self, pmin, pmax, cap=None):
        if cap is None:
            cap = int((pmax - pmin) * 8 // 30 + 64)
        out = np.zeros(max(cap, 1), np.uint64)
        cnt = C.c_uint64()
        _check(lib().rs_primegen(self._h, pmin, pmax, _u64p(out), cap, C.byref(cnt)))
        return out[:min(cnt.value, cap)]

    def peak_mad_u64_u32(self):
        """Measured v_mad_u64_u32 lane operations per second (rs_peak_mad_u64_u32)."""
        v = C.c_double()
        _check(lib().rs_peak_mad_u64_u32(self._h, C.byref(v)))
        return v.value

    def tune_candidates(self, nmin, nmax, kcount):
        cap = 128
        arr = (Params * cap)()
        n = C.c_int()
        _check(lib().rs_tune_candidates(self._h, nmin, nmax, kcount, arr, cap, C.byref(n)))
        if n.value > cap:
            raise RSError(RS_E_ARG, "rs_tune_candidates: %d candidates, room for %d" % (n.value, cap))
        return [arr[i] for i in range(n.value)]

    def mlp(self, n_in, n_hidden, n_out, seed=1):
        return MLP(self, n_in, n_hidden, n_out, seed)

    def job(self, base, k, c, nmin, nmax, bitmap, words_per_seq, params):
        return Job(self, base, k, c, nmin, nmax, bitmap, words_per_seq, params)

    def close(self):
        if self._h:
            for ch in list(self._children):
                ch.close()
            lib().rs_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Job:
    """One sieve job on a context (rs_job_create)."""

    def __init__(self, ctx, base, k, c, nmin, nmax, bitmap, words_per_seq, params):
        self.ctx = ctx
        self.k = np.ascontiguousarray(k, np.uint32)
        self.c = np.ascontiguousarray(c, np.int32)
        self.base, self.nmin, self.nmax, self.words_per_seq = base, nmin, nmax, words_per_seq
        self.nwords = len(self.k) * words_per_seq
        bitmap = np.ascontiguousarray(bitmap, np.uint32)
        assert bitmap.size == self.nwords
        self.params = params
        h = C.c_void_p()
        _check(lib().rs_job_create(ctx._h, base, len(self.k), _u32p(self.k), _i32p(self.c), nmin, nmax,
                                   _u32p(bitmap), words_per_seq, C.byref(params), C.byref(h)))
        self._h = h
        ctx._children.add(self)

    def sieve(self, pmin, pmax):
        _check(lib().rs_sieve(self._h, pmin, pmax))

    def sieve_noalloc(self, pmin, pmax):
        """rs_sieve_noalloc: the same sieve with no device allocation (static
        recovery scratch, no timing events)."""
        _check(lib().rs_sieve_noalloc(self._h, pmin, pmax))

    def counters(self):
        c = Counters()
        _check(lib().rs_counters_read(self._h, C.byref(c)))
        return c

    def drain(self, cap=1 << 22):
        out = np.zeros(cap, FACTOR_DTYPE)
        n = C.c_uint64()
        _check(lib().rs_factors_drain(self._h, out.ctypes.data_as(C.c_void_p), cap, C.byref(n)))
        return out[:n.value]

    def bitmap(self):
        out = np.zeros(self.nwords, np.uint32)
        _check(lib().rs_bitmap_read(self._h, _u32p(out), self.nwords))
        return out

    def set_bitmap(self, bm):
        bm = np.ascontiguousarray(bm, np.uint32)
        _check(lib().rs_bitmap_write(self._h, _u32p(bm), self.nwords))

    def sequence(self, kidx):
        k, c = C.c_uint32(), C.c_int32()
        _check(lib().rs_job_sequence(self._h, kidx, C.byref(k), C.byref(c)))
        return k.value, c.value

    def kernel_time(self):
        t, n, tot = C.c_double(), C.c_uint64(), C.c_double()
        _check(lib().rs_job_kernel_time(self._h, C.byref(t), C.byref(n), C.byref(tot)))
        return t.value, n.value, tot.value

    def view(self):
        """A handle sharing this job's state, with its own segment buffers."""
        v = Job.__new__(Job)
        v.ctx, v.k, v.c, v.base = self.ctx, self.k, self.c, self.base
        v.nmin, v.nmax, v.words_per_seq, v.nwords, v.params = (self.nmin, self.nmax, self.words_per_seq,
                                                                self.nwords, self.params)
        h = C.c_void_p()
        _check(lib().rs_job_view(self._h, C.byref(h)))
        v._h = h
        self.ctx._children.add(v)
        return v

    def reconfigure(self, params):
        _check(lib().rs_job_reconfigure(self._h, C.byref(params)))
        self.params = params

    def autotune(self, pmin, width, candidates):
        n = len(candidates)
        arr = (Params * n)(*candidates)
        secs = (C.c_double * n)()
        pr = (C.c_uint64 * n)()
        best = C.c_int()
        _check(lib().rs_autotune(self._h, pmin, width, arr, n, secs, pr, C.byref(best)))
        return best.value, list(secs), list(pr)

    def autotune_split(self, pmin, width, views, seg_primes):
        """rs_autotune_split: (best option, seconds, primes) of each (views,
        seg_primes) option over the same sample."""
        n = len(views)
        assert len(seg_primes) == n
        v = (C.c_uint32 * n)(*views)
        sp = (C.c_uint32 * n)(*seg_primes)
        secs = (C.c_double * n)()
        pr = (C.c_uint64 * n)()
        best = C.c_int()
        _check(lib().rs_autotune_split(self._h, pmin, width, v, sp, n, secs, pr, C.byref(best)))
        return best.value, list(secs), list(pr)

    def timer_reset(self):
        _check(lib().rs_job_timer_reset(self._h))

    def close(self):
        if self._h:
            lib().rs_job_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sieve_batch(jobs, pmins, pmaxs):
    arr = (C.c_void_p * len(jobs))(*[j._h for j in jobs])
    lo = (C.c_uint64 * len(jobs))(*pmins)
    hi = (C.c_uint64 * len(jobs))(*pmaxs)
    _check(lib().rs_sieve_batch(arr, len(jobs), lo, hi))


def _f32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class MLP:
    """The cost-model MLP (rssieve_mlp.h): n_in -> H -> H -> n_out, ReLU, fp32,
    trained and evaluated by the library's HIP kernels."""

    def __init__(self, ctx, n_in, n_hidden, n_out, seed=1):
        h = C.c_void_p()
        _check(lib().rs_mlp_create(ctx._h, n_in, n_hidden, n_out, seed, C.byref(h)))
        self._h = h
        ctx._children.add(self)
        self.n_in, self.H, self.n_out = n_in, n_hidden, n_out
        n = C.c_uint64()
        _check(lib().rs_mlp_num_params(h, C.byref(n)))
        self.nparam = n.value

    def shapes(self):
        a, H, C_ = self.n_in, self.H, self.n_out
        return [(a, H), (H,), (H, H), (H,), (H, C_), (C_,)]

    def get_params(self):
        out = np.zeros(self.nparam, np.float32)
        _check(lib().rs_mlp_get_params(self._h, _f32p(out)))
        res, o = [], 0
        for sh in self.shapes():
            k = int(np.prod(sh))
            res.append(out[o:o + k].reshape(sh).copy())
            o += k
        return res

    def set_params(self, params):
        flat = np.ascontiguousarray(np.concatenate([np.asarray(p, np.float32).reshape(-1) for p in params]))
        assert flat.size == self.nparam
        _check(lib().rs_mlp_set_params(self._h, _f32p(flat)))

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((X.shape[0], self.n_out), np.float32)
        _check(lib().rs_mlp_forward(self._h, _f32p(X), X.shape[0], _f32p(Y)))
        return Y

    def loss_grad(self, X, T, M=None):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        Mp = None
        if M is not None:
            M = np.ascontiguousarray(M, np.float32)
            Mp = _f32p(M)
        g = np.zeros(self.nparam, np.float32)
        loss = C.c_float()
        _check(lib().rs_mlp_loss_grad(self._h, _f32p(X), _f32p(T), Mp, X.shape[0], C.byref(loss), _f32p(g)))
        res, o = [], 0
        for sh in self.shapes():
            k = int(np.prod(sh))
            res.append(g[o:o + k].reshape(sh).copy())
            o += k
        return loss.value, res

    def adam_step(self, lr):
        _check(lib().rs_mlp_adam_step(self._h, lr))

    def train(self, X, T, M=None, epochs=50, batch=256, lr=1e-3, seed=1):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        Mp = None
        if M is not None:
            M = np.ascontiguousarray(M, np.float32)
            Mp = _f32p(M)
        loss = C.c_float()
        _check(lib().rs_mlp_train(self._h, _f32p(X), _f32p(T), Mp, X.shape[0], epochs, batch, lr, seed,
                                  C.byref(loss)))
        return loss.value

    def close(self):
        if self._h:
            lib().rs_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
