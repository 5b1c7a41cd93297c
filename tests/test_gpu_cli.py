"""The rssieve command line on the GPU: ABCD in, factor file out, same
terms eliminated as the oracle; checkpoint resume; self test."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import rssieve_amd as R
from util import bitmap_terms, full_bitmap

pytestmark = pytest.mark.gpu


def _make_abcd(path, base, ks, cs, nmin, nmax, density, seed):
    bm, wps = full_bitmap(len(ks), nmin, nmax, density=density, seed=seed)
    R.abcd_write(str(path), base, ks, cs, nmin, nmax, bm, wps, sieved_to=3)
    a = R.ABCD(str(path))
    return a


def _run(args, cwd, timeout=120):
    print("rssieve", " ".join(args), flush=True)
    return subprocess.run([R.CLI_PATH] + args, cwd=cwd, capture_output=True, text=True, timeout=timeout)


def _factor_lines(path):
    return [l for l in open(path).read().splitlines() if l.strip()]


def test_cli_matches_oracle(tmp_path):
    base, ks, cs = 2, [3, 5, 7, 11, 13, 15, 17, 19], [1, -1, 1, -1, 1, -1, 1, -1]
    nmin, nmax = 20000, 60000
    a = _make_abcd(tmp_path / "in.abcd", base, ks, cs, nmin, nmax, 0.6, 1)
    r = _run(["-p", "1e8", "-P", "100500000", "-i", "in.abcd", "-f", "out.txt"], str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Sequences read: 8" in r.stdout and "Base = 2" in r.stdout
    lines = _factor_lines(tmp_path / "out.txt")
    job = O.make_job(base, [int(x) for x in a.k], [int(x) for x in a.c], a.nmin, a.nmax)
    bm = a.bitmap()
    cnt, facs, nf = O.sieve(job, 10**8, 100500000, bm.copy())
    want = {(int(a.k[i]), int(a.c[i]), n) for (p, n, i) in facs}
    got = set()
    for l in lines:
        p, k, s, n = R.parse_factor(l, 2)
        c = 1 if s == "+" else -1
        assert R.verify_factor(p, k, n, c, 2) == 1
        got.add((k, c, n))
    assert got == want
    # lines sorted by (p, n) within the batch, reference format
    ps = [int(l.split(" | ")[0]) for l in lines]
    assert ps == sorted(ps)
    summary = [l for l in r.stdout.splitlines() if l.startswith("factors ")][-1]
    assert summary.startswith("factors %d, prime count %d, checksum " % (len(want), cnt.primecount))
    assert summary.endswith("duplicate factors %d" % cnt.dups)


def test_cli_no_factors_and_resume(tmp_path):
    base, ks, cs = 2, [1001, 3003], [1, -1]
    _make_abcd(tmp_path / "in.abcd", base, ks, cs, 1000000, 1001000, 0.5, 2)
    r = _run(["-p", "1000000007", "-P", "1000000100", "-i", "in.abcd", "-f", "f.txt"], str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    assert open(tmp_path / "f.txt").read() == "no factors\n"
    # a completed run leaves a checkpoint at p == P: running again reports completion
    r2 = _run(["-p", "1000000007", "-P", "1000000100", "-i", "in.abcd", "-f", "f.txt"], str(tmp_path))
    assert r2.returncode == 0 and "Workunit complete." in r2.stdout


def test_cli_errors(tmp_path):
    (tmp_path / "bad.abcd").write_text("ABCD 3*4^$a+1 [10]\n")
    r = _run(["-p", "1e6", "-P", "2e6", "-i", "bad.abcd"], str(tmp_path))
    assert r.returncode != 0 and "perfect power" in r.stdout
    (tmp_path / "ok.abcd").write_text("ABCD 3*2^$a+1 [10]\n")
    r = _run(["-i", "ok.abcd"], str(tmp_path))
    assert r.returncode != 0 and "-p and -P arguments are required" in r.stdout
    r = _run(["-p", "1e6", "-P", "2e6"], str(tmp_path))
    assert r.returncode != 0 and "no input file" in r.stdout


def test_cli_selftest(tmp_path):
    r = _run(["-s"], str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "1000133 | 91549*2^54378870+1 found" in r.stdout


def test_cli_tuned_same_terms(tmp_path):
    """-t (autotuned merged table) eliminates exactly the terms the default
    reference split does."""
    base, ks, cs = 2, [3, 5, 7, 11, 13, 15], [1, -1, 1, -1, 1, -1]
    _make_abcd(tmp_path / "in.abcd", base, ks, cs, 40000, 90000, 0.5, 7)
    outs = []
    for flags, name in [([], "plain"), (["-t"], "tuned")]:
        d = tmp_path / name
        d.mkdir()
        (d / "in.abcd").write_text((tmp_path / "in.abcd").read_text())
        r = _run(["-p", "5e8", "-P", "503000000", "-i", "in.abcd", "-f", "out.txt"] + flags, str(d), timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        if flags:
            assert "Autotuned: candidate" in r.stdout
        terms = set()
        for l in _factor_lines(d / "out.txt"):
            if l == "no factors":
                continue
            p, k, s, n = R.parse_factor(l, 2)
            terms.add((k, s, n))
        outs.append(terms)
    assert outs[0] == outs[1]


def _terms(path, base):
    out = []
    for l in _factor_lines(path):
        if l == "no factors":
            continue
        p, k, s, n = R.parse_factor(l, base)
        out.append((p, k, 1 if s == "+" else -1, n))
    return out


def test_cli_learned_batched_same_terms_as_plain_and_oracle(tmp_path):
    """-l (the committed cost model picks the configuration and the split,
    through the C ABI) with -b 4 (four views per batched round): the same
    eliminated terms as the plain run (reference split, one view) and the
    oracle."""
    base, ks, cs = 2, [3, 5, 7, 11, 13, 15, 17, 19], [1, -1, 1, -1, 1, -1, 1, -1]
    a = _make_abcd(tmp_path / "in.abcd", base, ks, cs, 20000, 60000, 0.6, 3)
    p0, p1 = 10**9, 10**9 + 4 * 10**6
    job = O.make_job(base, [int(x) for x in a.k], [int(x) for x in a.c], a.nmin, a.nmax)
    cnt, facs, nf = O.sieve(job, p0, p1, a.bitmap().copy())
    want = {(int(a.k[i]), int(a.c[i]), n) for (p, n, i) in facs}
    got = {}
    for flags, name in [([], "plain"), (["-l", "-b", "4"], "learned")]:
        d = tmp_path / name
        d.mkdir()
        (d / "in.abcd").write_text((tmp_path / "in.abcd").read_text())
        r = _run(["-p", str(p0), "-P", str(p1), "-i", "in.abcd", "-f", "out.txt"] + flags, str(d), timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        if flags:
            assert "Cost model: candidate" in r.stdout and "Batched: 4 views per round" in r.stdout
        terms = _terms(d / "out.txt", base)
        for (p, k, c, n) in terms:
            assert R.verify_factor(p, k, n, c, base) == 1
        got[name] = {(k, c, n) for (p, k, c, n) in terms}
        summary = [l for l in r.stdout.splitlines() if l.startswith("factors ")][-1]
        assert ("prime count %d," % cnt.primecount) in summary
        assert "Primes tested per second:" in r.stdout
    assert got["plain"] == got["learned"] == want


def test_cli_resume_mid_range(tmp_path):
    """A run resumed from a checkpoint at mid-range (cl_sieve.cpp:797-863,
    1047-1063): the factor file's lines clear their terms from the bitmap,
    the rest of the range is sieved, and the file ends with the same term set
    as an uninterrupted run, each term once.  A factor file with fewer lines
    than the checkpoint counts stops with "Missing factors"."""
    base, ks, cs = 2, [3, 5, 7, 11, 13, 15, 17, 19], [1, -1, 1, -1, 1, -1, 1, -1]
    a = _make_abcd(tmp_path / "in.abcd", base, ks, cs, 20000, 60000, 0.6, 4)
    p0, p1 = 10**8, 10**8 + 10**6
    full = tmp_path / "full"
    full.mkdir()
    (full / "in.abcd").write_text((tmp_path / "in.abcd").read_text())
    args = ["-p", str(p0), "-P", str(p1), "-i", "in.abcd", "-f", "out.txt"]
    r = _run(args, str(full))
    assert r.returncode == 0, r.stdout + r.stderr
    lines = _factor_lines(full / "out.txt")
    all_terms = {(k, c, n) for (p, k, c, n) in _terms(full / "out.txt", base)}
    mid = p0 + (p1 - p0) // 2
    before = [l for l in lines if int(l.split(" | ")[0]) < mid]
    assert before and len(before) < len(lines)
    job = O.make_job(base, [int(x) for x in a.k], [int(x) for x in a.c], a.nmin, a.nmax)
    cnt_before, _, _ = O.sieve(job, p0, mid, a.bitmap().copy())
    cnt_full, _, _ = O.sieve(job, p0, p1, a.bitmap().copy())

    def checkpoint(d, nlines):
        st = R.State()
        st.pmin, st.pmax, st.p = p0, p1, mid
        st.base, st.nmin, st.nmax, st.kcount = base, a.nmin, a.nmax, len(ks)
        for i in range(len(ks)):
            st.klist[i] = int(a.k[i]) if int(a.c[i]) > 0 else -int(a.k[i])
        st.factorcount = len(before)
        st.primecount = cnt_before.primecount
        wa = C.c_int(1)
        assert R.lib().rs_state_write(str(d).encode(), st, wa) == 0
        (d / "in.abcd").write_text((tmp_path / "in.abcd").read_text())
        (d / "out.txt").write_text("".join(l + "\n" for l in before[:nlines]))

    import ctypes as C
    res = tmp_path / "resumed"
    res.mkdir()
    checkpoint(res, len(before))
    r = _run(args, str(res))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Resuming from checkpoint" in r.stdout and ("Updated bitmap with %d factors" % len(before)) in r.stdout
    got = _terms(res / "out.txt", base)
    keys = [(k, c, n) for (p, k, c, n) in got]
    assert len(keys) == len(set(keys))
    assert set(keys) == all_terms
    summary = [l for l in r.stdout.splitlines() if l.startswith("factors ")][-1]
    assert summary.startswith("factors %d, prime count %d," % (len(keys), cnt_full.primecount))

    short = tmp_path / "short"
    short.mkdir()
    checkpoint(short, len(before) - 1)
    r = _run(args, str(short))
    assert r.returncode != 0 and "Missing factors" in r.stdout


def test_cli_boinc_shim(tmp_path):
    """The BOINC shim (cl_sieve.cpp:40-62, 1406-1414; main.cpp:207-229):
    without -d the device comes from the slot's init_data.xml, -F leaves the
    fraction done 1 at the end, and -T writes the trickle-up message at a
    checkpoint once a day has passed since the checkpoint's last_trickle (a
    resumed checkpoint from long ago); the factor file is the plain run's.  A
    malformed device tag stops the run before the GPU is opened."""
    import ctypes as C
    base, ks, cs = 2, [3, 5, 7, 11], [1, -1, 1, -1]
    a = _make_abcd(tmp_path / "in.abcd", base, ks, cs, 20000, 60000, 0.6, 5)
    p0, p1 = 10**8, 10**8 + 10**6
    args = ["-p", str(p0), "-P", str(p1), "-i", "in.abcd", "-f", "out.txt"]
    plain = tmp_path / "plain"
    plain.mkdir()
    (plain / "in.abcd").write_text((tmp_path / "in.abcd").read_text())
    r = _run(args, str(plain))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "init_data.xml not found, using device 0." in r.stdout

    d = tmp_path / "boinc"
    d.mkdir()
    (d / "in.abcd").write_text((tmp_path / "in.abcd").read_text())
    (d / "init_data.xml").write_text("<app_init_data>\n<gpu_type>ATI</gpu_type>\n<gpu_device_num>0</gpu_device_num>\n"
                                     "</app_init_data>\n")
    st = R.State()
    st.pmin, st.pmax, st.p = p0, p1, p0
    st.base, st.nmin, st.nmax, st.kcount = base, a.nmin, a.nmax, len(ks)
    for i in range(len(ks)):
        st.klist[i] = int(a.k[i]) if int(a.c[i]) > 0 else -int(a.k[i])
    st.last_trickle = 1  # the checkpoint's last trickle, long ago
    assert R.lib().rs_state_write(str(d).encode(), st, C.c_int(1)) == 0
    (d / "out.txt").write_text("")
    r = _run(args + ["-F", "frac.txt", "-T", "trickle.xml"], str(d))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Using device 0 from init_data.xml" in r.stdout
    assert (d / "frac.txt").read_text() == "1.000000\n"
    msg = (d / "trickle.xml").read_text()
    assert msg.startswith("<trickle_up>\n   <progress>") and msg.endswith("</runtime>\n</trickle_up>\n")
    assert "<cputime>" in msg
    assert sorted(_factor_lines(d / "out.txt")) == sorted(_factor_lines(plain / "out.txt"))

    bad = tmp_path / "bad"
    bad.mkdir()
    (bad / "in.abcd").write_text((tmp_path / "in.abcd").read_text())
    (bad / "init_data.xml").write_text("<gpu_device_num>gpu0</gpu_device_num>\n")
    r = _run(args, str(bad))
    assert r.returncode != 0 and "init_data.xml holds no valid gpu_device_num" in r.stdout
