// rs_io.cpp -- ABCD input, factor file, checkpoint state and the CPU
// pre-sieve.  Formats are byte-compatible with the reference
// (inputfile.cpp, cl_sieve.cpp:185-328, 597-643, 738-863).
#include <cctype>
#include <cinttypes>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/rssieve.h"
#include "../../include/rssieve_io.h"

namespace rs {
void set_error(const std::string& msg);  // rs_capi.cpp: the text rs_last_error() returns
}

namespace {

int io_fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  rs::set_error(buf);
  return code;
}

}  // namespace

struct rs_abcd {
  uint32_t base = 0;
  int32_t kcount = 0;
  uint32_t k[RS_MAX_SEQUENCES] = {0};
  int32_t c[RS_MAX_SEQUENCES] = {0};
  uint32_t nmin = 0, nmax = 0;
  uint64_t words_per_seq = 0;
  uint64_t nterms = 0;
  std::vector<uint32_t> bitmap;
};

// sr5sieve-style header (inputfile.cpp:188-227): with or without the tail
static int parse_header(const char* line, uint32_t* k, uint32_t* b, int32_t* c, uint32_t* n) {
  {
    uint64_t pd = 0;
    uint32_t kk = 0, bb = 0, nn = 0;
    int32_t cc = 0;
    if (sscanf(line, "ABCD %" SCNu32 "*%" SCNu32 "^$a%" SCNd32 " [%" SCNu32 "] // Sieved to %" SCNu64, &kk, &bb, &cc,
               &nn, &pd) == 5) {
      *k = kk, *b = bb, *c = cc, *n = nn;
      return 1;
    }
  }
  {
    uint32_t kk = 0, bb = 0, nn = 0;
    int32_t cc = 0;
    if (sscanf(line, "ABCD %" SCNu32 "*%" SCNu32 "^$a%" SCNd32 " [%" SCNu32 "]", &kk, &bb, &cc, &nn) == 4) {
      *k = kk, *b = bb, *c = cc, *n = nn;
      return 1;
    }
  }
  return 0;
}

// "file:line: (seq i) msg\n  >> line" (inputfile.cpp:129-149)
static int line_error(const char* file, uint64_t line_no, const char* msg, const char* line, int seq) {
  std::string s;
  char buf[512];
  if (seq >= 0)
    snprintf(buf, sizeof(buf), "%s:%" PRIu64 ": (seq %d) %s", file, line_no, seq, msg);
  else
    snprintf(buf, sizeof(buf), "%s:%" PRIu64 ": %s", file, line_no, msg);
  s = buf;
  if (line && *line) {
    s += "\n  >> ";
    s += line;
    while (!s.empty() && (s.back() == '\n' || s.back() == '\r')) s.pop_back();
  }
  rs::set_error(s);
  return RS_E_IO;
}

extern "C" int rs_abcd_read(const char* path, rs_abcd** out) {
  if (!path || !*path || !out) return io_fail(RS_E_ARG, "Error: input_file is null/empty");
  FILE* fp = fopen(path, "r");
  if (!fp) return io_fail(RS_E_IO, "Error: unable to open input file %s", path);
  rs_abcd* a = new rs_abcd;
  std::vector<std::vector<uint32_t>> nl(RS_MAX_SEQUENCES);
  uint32_t firstN[RS_MAX_SEQUENCES] = {0}, lastN[RS_MAX_SEQUENCES] = {0};
  char line[2048];
  uint64_t line_no = 0;
  int cur = -1;
  uint32_t running = 0;
  int64_t gB = -1;
  uint32_t NMIN = UINT32_MAX, NMAX = 0;
  int rc = RS_OK;
  while (fgets(line, sizeof(line), fp)) {
    line_no++;
    const char* s = line;
    while (*s && isspace((unsigned char)*s)) s++;
    if (*s == '\0' || *s == '#') continue;
    uint32_t delta = 0;
    if (sscanf(line, "%" SCNu32, &delta) == 1) {  // inputfile.cpp:303-318
      if (cur < 0) {
        rc = line_error(path, line_no, "delta line before first ABCD header", line, cur);
        break;
      }
      uint32_t r = running + delta;
      if (r < running) {
        rc = line_error(path, line_no, "overflow while accumulating n += delta (uint32_t wrap)", line, cur);
        break;
      }
      running = r;
      lastN[cur] = running;
      nl[cur].push_back(running);
      continue;
    }
    uint32_t hk = 0, hb = 0, hn = 0;
    int32_t hc = 0;
    if (!parse_header(line, &hk, &hb, &hc, &hn)) {
      rc = line_error(path, line_no, "invalid or unrecognised line", line, cur);
      break;
    }
    if (hk == 0) { rc = line_error(path, line_no, "invalid K (must be > 0)", line, cur); break; }
    if (hn == 0) { rc = line_error(path, line_no, "invalid N0 (must be > 0)", line, cur); break; }
    if (hb < 2) { rc = line_error(path, line_no, "invalid base in ABCD header (B must be >= 2)", line, cur); break; }
    if (gB < 0) gB = hb;
    else if ((uint32_t)gB != hb) {
      rc = line_error(path, line_no, "mismatched base (B must match all sequences)", line, cur);
      break;
    }
    if (hc != 1 && hc != -1) {
      rc = line_error(path, line_no, "not standard form k*b^n+/-1 (c must be +1 or -1)", line, cur);
      break;
    }
    if (a->kcount >= RS_MAX_SEQUENCES) {
      char buf[128];
      snprintf(buf, sizeof(buf), "too many sequences (MAX_SEQUENCES=%d)", RS_MAX_SEQUENCES);
      rc = line_error(path, line_no, buf, line, cur);
      break;
    }
    cur = a->kcount++;
    a->k[cur] = hk;
    a->c[cur] = hc;
    firstN[cur] = lastN[cur] = hn;
    running = hn;
    nl[cur].push_back(running);
  }
  if (rc == RS_OK && ferror(fp)) rc = io_fail(RS_E_IO, "%s:%" PRIu64 ": read error while reading ABCD file", path, line_no);
  fclose(fp);
  if (rc == RS_OK && (a->kcount == 0 || gB < 0)) rc = io_fail(RS_E_IO, "%s: no sequences found in input", path);
  if (rc != RS_OK) {
    delete a;
    return rc;
  }
  for (int i = 0; i < a->kcount; i++) {
    if (firstN[i] < NMIN) NMIN = firstN[i];
    if (lastN[i] > NMAX) NMAX = lastN[i];
  }
  if (NMIN & 1u) NMIN--;  // inputfile.cpp:405-407
  a->base = (uint32_t)gB;
  a->nmin = NMIN;
  a->nmax = NMAX;
  const uint64_t nbits = (uint64_t)NMAX - NMIN + 1;
  a->words_per_seq = (nbits + 31) >> 5;
  a->bitmap.assign((size_t)a->kcount * a->words_per_seq, 0u);
  for (int i = 0; i < a->kcount; i++)
    for (uint32_t n : nl[i]) {
      const uint64_t off = (uint64_t)n - NMIN;
      uint32_t& w = a->bitmap[(size_t)i * a->words_per_seq + (off >> 5)];
      const uint32_t m = 1u << (off & 31);
      if (!(w & m)) a->nterms++;
      w |= m;
    }
  *out = a;
  return RS_OK;
}

extern "C" void rs_abcd_free(rs_abcd* a) { delete a; }

extern "C" int rs_abcd_info(const rs_abcd* a, uint32_t* base, int32_t* kcount, uint32_t* nmin, uint32_t* nmax,
                            uint64_t* wps, uint64_t* nterms) {
  if (!a) return io_fail(RS_E_ARG, "rs_abcd_info: null");
  if (base) *base = a->base;
  if (kcount) *kcount = a->kcount;
  if (nmin) *nmin = a->nmin;
  if (nmax) *nmax = a->nmax;
  if (wps) *wps = a->words_per_seq;
  if (nterms) *nterms = a->nterms;
  return RS_OK;
}

extern "C" int rs_abcd_sequences(const rs_abcd* a, uint32_t* k, int32_t* c) {
  if (!a || !k || !c) return io_fail(RS_E_ARG, "rs_abcd_sequences: null");
  for (int i = 0; i < a->kcount; i++) {
    k[i] = a->k[i];
    c[i] = a->c[i];
  }
  return RS_OK;
}

extern "C" uint32_t* rs_abcd_bitmap(rs_abcd* a) { return a ? a->bitmap.data() : nullptr; }

extern "C" int rs_abcd_mark_used(rs_abcd* a, uint32_t K, char sign, uint32_t n, int32_t compat) {
  if (!a) return io_fail(RS_E_ARG, "rs_abcd_mark_used: null");
  int idx = -1;
  for (int i = 0; i < a->kcount && idx < 0; i++) {
    if (compat) {
      // inputfile.cpp:89-96 on the signed encoding
      const int32_t want = (sign == '+') ? (int32_t)K : (int32_t)(0u - K);
      const int32_t have = (a->c[i] > 0) ? (int32_t)a->k[i] : (int32_t)(0u - a->k[i]);
      if (want == have) idx = i;
    } else if (a->k[i] == K && ((sign == '+') == (a->c[i] > 0))) {
      idx = i;
    }
  }
  if (idx < 0 || n < a->nmin || n > a->nmax) return 0;
  const uint64_t off = (uint64_t)n - a->nmin;
  uint32_t& w = a->bitmap[(size_t)idx * a->words_per_seq + (off >> 5)];
  const uint32_t m = 1u << (off & 31);
  const int was = (w & m) != 0;
  w &= ~m;
  return was;
}

extern "C" int rs_abcd_write(const char* path, uint32_t base, int32_t kcount, const uint32_t* k, const int32_t* c,
                             uint32_t nmin, uint32_t nmax, const uint32_t* bitmap, uint64_t wps, uint64_t sieved_to) {
  if (!path || !k || !c || !bitmap || kcount < 1) return io_fail(RS_E_ARG, "rs_abcd_write: bad args");
  FILE* f = fopen(path, "w");
  if (!f) return io_fail(RS_E_IO, "Cannot open %s !!!", path);
  for (int i = 0; i < kcount; i++) {
    const uint32_t* row = bitmap + (size_t)i * wps;
    uint32_t prev = 0;
    bool first = true;
    for (uint64_t n = nmin; n <= nmax; n++) {
      const uint64_t off = n - nmin;
      if (!((row[off >> 5] >> (off & 31)) & 1u)) continue;
      if (first) {
        fprintf(f, "ABCD %u*%u^$a%+d [%u] // Sieved to %" PRIu64 "\n", k[i], base, c[i], (uint32_t)n, sieved_to);
        first = false;
      } else {
        fprintf(f, "%u\n", (uint32_t)n - prev);
      }
      prev = (uint32_t)n;
    }
  }
  if (fclose(f) != 0) return io_fail(RS_E_IO, "Cannot write to %s !!!", path);
  return RS_OK;
}

// ---------------------------------------------------------------------------
// pre-sieve (discrete logarithms modulo small p)
// ---------------------------------------------------------------------------
namespace {

inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((unsigned __int128)a * b % m); }
inline uint64_t powmod(uint64_t a, uint64_t e, uint64_t m) {
  uint64_t r = 1 % m;
  a %= m;
  while (e) {
    if (e & 1) r = mulmod(r, a, m);
    a = mulmod(a, a, m);
    e >>= 1;
  }
  return r;
}
inline uint64_t invmod(uint64_t a, uint64_t m) {  // m prime, a != 0
  return powmod(a, m - 2, m);
}

// multiplicative order of b modulo prime p (p does not divide b)
uint64_t order_of(uint64_t b, uint64_t p) {
  uint64_t n = p - 1, ord = p - 1;
  for (uint64_t f = 2; f * f <= n; f++) {
    if (n % f) continue;
    while (n % f == 0) n /= f;
    while (ord % f == 0 && powmod(b, ord / f, p) == 1) ord /= f;
  }
  if (n > 1)
    while (ord % n == 0 && powmod(b, ord / n, p) == 1) ord /= n;
  return ord;
}

struct Sol {
  uint32_t n0;   // smallest solution >= 0
  uint32_t ord;  // period
};

// open-addressing map from value (< 2^32) to exponent
struct SmallMap {
  std::vector<uint64_t> key;
  std::vector<uint32_t> val;
  uint64_t mask = 0;
  void reset(size_t n) {
    size_t sz = 16;
    while (sz < 2 * n + 2) sz <<= 1;
    key.assign(sz, ~0ull);
    val.assign(sz, 0);
    mask = sz - 1;
  }
  static uint64_t h(uint64_t x) { return (x * 0x9E3779B97F4A7C15ull) >> 29; }
  void put(uint64_t x, uint32_t v) {
    uint64_t s = h(x) & mask;
    while (key[s] != ~0ull) {
      if (key[s] == x) return;  // keep the smallest exponent
      s = (s + 1) & mask;
    }
    key[s] = x;
    val[s] = v;
  }
  bool get(uint64_t x, uint32_t* v) const {
    uint64_t s = h(x) & mask;
    while (key[s] != ~0ull) {
      if (key[s] == x) {
        *v = val[s];
        return true;
      }
      s = (s + 1) & mask;
    }
    return false;
  }
};

}  // namespace

extern "C" int rs_presieve(uint32_t base, int32_t kcount, const uint32_t* k, const int32_t* c, uint32_t nmin,
                           uint32_t nmax, uint32_t* bitmap, uint64_t wps, uint64_t bound, int nthreads) {
  if (!k || !c || !bitmap || kcount < 1 || kcount > RS_MAX_SEQUENCES || nmax < nmin || base < 2 ||
      bound > (1ull << 32))
    return io_fail(RS_E_ARG, "rs_presieve: bad args");
  if (bound <= 2) return RS_OK;
  // primes below bound
  std::vector<uint8_t> comp(bound, 0);
  std::vector<uint32_t> primes;
  for (uint64_t i = 2; i < bound; i++) {
    if (comp[i]) continue;
    primes.push_back((uint32_t)i);
    for (uint64_t j = i * i; j < bound; j += i) comp[j] = 1;
  }
  comp.clear();
  comp.shrink_to_fit();
  if (nthreads < 1) nthreads = 1;
  // phase 1: threads over primes find (n0, ord) per (p, k)
  std::vector<std::vector<std::vector<Sol>>> per(nthreads, std::vector<std::vector<Sol>>(kcount));
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; t++) {
    th.emplace_back([&, t]() {
      SmallMap map;
      for (size_t pi = t; pi < primes.size(); pi += nthreads) {
        const uint64_t p = primes[pi];
        const uint64_t bm = base % p;
        if (bm == 0) {
          // k*b^n + c == c (mod p) for n >= 1; only n = 0 can be divisible
          if (nmin == 0)
            for (int i = 0; i < kcount; i++)
              if ((k[i] + (int64_t)c[i] + (int64_t)p) % (int64_t)p == 0) per[t][i].push_back({0, 0});
          continue;
        }
        const uint64_t ord = order_of(bm, p);
        const uint64_t m = (uint64_t)ceil(sqrt((double)ord * (kcount > 4 ? 4.0 : 1.0)));
        const uint64_t mm = m < ord ? m : ord;
        map.reset(mm);
        uint64_t v = 1;
        for (uint64_t j = 0; j < mm; j++) {
          map.put(v, (uint32_t)j);
          v = mulmod(v, bm, p);
        }
        const uint64_t gstep = invmod(powmod(bm, mm, p), p);  // b^-m
        const uint64_t ngiant = (ord + mm - 1) / mm;
        for (int i = 0; i < kcount; i++) {
          const uint64_t kk = k[i] % p;
          if (kk == 0) continue;
          // target b^n == -c / k
          uint64_t tgt = mulmod((c[i] > 0) ? p - 1 : 1 % p, invmod(kk, p), p);
          uint64_t g = tgt;
          for (uint64_t qq = 0; qq < ngiant; qq++) {
            uint32_t j;
            if (map.get(g, &j)) {
              per[t][i].push_back({(uint32_t)((qq * mm + j) % ord), (uint32_t)ord});
              break;
            }
            g = mulmod(g, gstep, p);
          }
        }
      }
    });
  }
  for (auto& x : th) x.join();
  th.clear();
  // phase 2: threads over sequences clear bits
  std::vector<uint64_t> cleared(kcount, 0);
  for (int t = 0; t < nthreads; t++) {
    th.emplace_back([&, t]() {
      for (int i = t; i < kcount; i += nthreads) {
        uint32_t* row = bitmap + (size_t)i * wps;
        for (int s = 0; s < nthreads; s++)
          for (const Sol& so : per[s][i]) {
            if (so.ord == 0) {  // only n = 0
              if (nmin == 0) row[0] &= ~1u;
              continue;
            }
            // first n >= nmin with n == n0 (mod ord)
            uint64_t n = so.n0;
            if (n < nmin) n += ((nmin - n + so.ord - 1) / so.ord) * (uint64_t)so.ord;
            for (; n <= nmax; n += so.ord) {
              const uint64_t off = n - nmin;
              row[off >> 5] &= ~(1u << (off & 31));
            }
          }
      }
    });
  }
  for (auto& x : th) x.join();
  return RS_OK;
}

// ---------------------------------------------------------------------------
// factor file lines
// ---------------------------------------------------------------------------
extern "C" int rs_format_factor(char* buf, int buflen, uint64_t p, uint32_t k, uint32_t b, uint32_t n, int32_t c) {
  if (!buf || buflen < 2) return io_fail(RS_E_ARG, "rs_format_factor: buffer");
  const int r = snprintf(buf, (size_t)buflen, "%" PRIu64 " | %u*%u^%u%+d\n", p, k, b, n, c);  // cl_sieve.cpp:624
  if (r < 0 || r >= buflen) return io_fail(RS_E_ARG, "rs_format_factor: buffer too small");
  return r;
}

extern "C" int rs_parse_factor(const char* line, uint32_t base, uint64_t* p, uint32_t* k, char* sign, uint32_t* n) {
  if (!line) return io_fail(RS_E_ARG, "rs_parse_factor: null");
  uint64_t pp;
  uint32_t kk, bb, nn, cc;
  char bar, sg, extra;
  // cl_sieve.cpp:830-845
  const int fields = sscanf(line,
                            " %" SCNu64 " %c "
                            "%" SCNu32 " * "
                            "%" SCNu32 " ^ "
                            "%" SCNu32 " %c "
                            "%" SCNu32 " %c",
                            &pp, &bar, &kk, &bb, &nn, &sg, &cc, &extra);
  if (fields != 7 || bar != '|' || (sg != '+' && sg != '-') || cc != 1 || bb != base)
    return io_fail(RS_E_IO, "Invalid line read in factors file: %s", line);
  if (p) *p = pp;
  if (k) *k = kk;
  if (sign) *sign = sg;
  if (n) *n = nn;
  return RS_OK;
}

// ---------------------------------------------------------------------------
// checkpoint state (cl_sieve.cpp:185-328)
// ---------------------------------------------------------------------------
static uint64_t state_sum(const rs_state* s) {
  uint64_t v = s->base + s->pmin + s->pmax + s->p + s->checksum + s->primecount + s->factorcount + s->last_trickle +
               s->nmin + s->nmax + s->kcount + s->dupcount;
  for (int i = 0; i < s->kcount && i < RS_MAX_SEQUENCES; ++i) v += s->klist[i];
  return v;
}

static std::string state_path(const char* dir, bool a) {
  std::string d = (dir && *dir) ? dir : ".";
  return d + (a ? "/stateA.ckp" : "/stateB.ckp");
}

extern "C" int rs_state_write(const char* dir, rs_state* st, int* write_a_next) {
  if (!st || !write_a_next) return io_fail(RS_E_ARG, "rs_state_write: null");
  st->state_sum = state_sum(st);
  const std::string path = state_path(dir, *write_a_next != 0);
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return io_fail(RS_E_IO, "Cannot open %s !!!", path.c_str() + path.rfind('/') + 1);
  if (fwrite(st, sizeof(rs_state), 1, f) != 1) {
    fclose(f);
    return io_fail(RS_E_IO, "Cannot write checkpoint to file. Continuing...");
  }
  if (fclose(f) == 0) *write_a_next = !*write_a_next;
  return RS_OK;
}

static int load_state(const std::string& path, const rs_state* want, rs_state* got, std::string* msg) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return 0;
  const char* name = path.c_str() + path.rfind('/') + 1;
  int ok = 1;
  char buf[256];
  if (fread(got, sizeof(rs_state), 1, f) != 1) {
    snprintf(buf, sizeof(buf), "Cannot parse %s !!!", name);
    ok = 0;
  } else if (got->base != want->base || got->pmin != want->pmin || got->pmax != want->pmax ||
             got->nmin != want->nmin || got->nmax != want->nmax || got->kcount != want->kcount) {
    snprintf(buf, sizeof(buf), "Invalid checkpoint file %s !!!", name);
    ok = 0;
  } else if (got->kcount < 0 || got->kcount > RS_MAX_SEQUENCES || state_sum(got) != got->state_sum) {
    snprintf(buf, sizeof(buf), "Checksum error in %s !!!", name);
    ok = 0;
  }
  fclose(f);
  if (!ok) *msg += std::string(msg->empty() ? "" : "\n") + buf;
  return ok;
}

extern "C" int rs_state_read(const char* dir, rs_state* st, int* write_a_next) {
  if (!st || !write_a_next) return io_fail(RS_E_ARG, "rs_state_read: null");
  rs_state a, b;
  std::string msg;
  const int ga = load_state(state_path(dir, true), st, &a, &msg);
  const int gb = load_state(state_path(dir, false), st, &b, &msg);
  rs::set_error(msg);
  bool use_a = ga && (!gb || a.p > b.p);
  bool use_b = gb && !use_a;
  if (use_a) {
    *st = a;
    *write_a_next = 0;
    return 1;
  }
  if (use_b) {
    *st = b;
    *write_a_next = 1;
    return 1;
  }
  return 0;
}
