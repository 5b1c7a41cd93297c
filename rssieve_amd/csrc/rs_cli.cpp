// rs_cli.cpp -- the 'rssieve' command line: the reference's -p/-P/-i/-f
// interface (main.cpp:25-134) and sieve driver (cl_sieve.cpp:975-1451) on top
// of the C ABI of librssieve.so.  Same ABCD input, factor-file lines,
// checkpoint files and summary line as the reference.
#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include <unistd.h>

#include "rs_boinc.h"
#include "rssieve.h"
#include "rssieve_io.h"
#include "rssieve_mlp.h"

namespace {

struct Options {
  uint64_t pmin = 0, pmax = 0;
  std::string input, factor_file;
  bool selftest = false;
  int device = 0;
  bool device_set = false;  // -d given (else init_data.xml's device, else 0)
  std::string frac_file;    // -F: fraction done for a BOINC wrapper
  std::string trickle_file; // -T: daily trickle-up message for a BOINC wrapper
  bool autotune = false;  // -t: time every candidate configuration on a sample, sieve with the fastest
  bool learned = false;   // -l / -m: the cost model picks the configuration and the batch split
  std::string model_path; // -m: model file (default: model/cost_model.bin beside the executable)
  uint32_t views = 0;     // -b: views per batched round (0: 1, or the model's split with -l)
  bool compat = false;    // -c: reference signed-int k encoding
};

void help() {
  printf("rssieve: a HIP (gfx950) Riesel Sierpinski sieve for multiple sequences k*b^n+-1 using the BSGS algorithm.\n");
  printf("Program usage:\n");
  printf("-p #\t\t\tStarting prime factor p\n");
  printf("-P #\t\t\tEnd prime factor P\n");
  printf("\t\t\tP range is 3 <= -p < -P < 2^64, [-p, -P) exclusive\n");
  printf("-i inputfile\t\tUse specified sr2sieve ABCD input file with a maximum of 100 sequences\n");
  printf("-f factorfile\t\tOverride default factor file name (psp_sr2sieve.out) with specified file name.\n");
  printf("-d #\t\t\tHIP device index (default 0)\n");
  printf("-t\t\t\tAutotune: time every candidate BSGS configuration on a sample and use the fastest\n");
  printf("-l\t\t\tLearned: the committed cost model ranks the BSGS configurations (its three fastest\n");
  printf("\t\t\tare timed on a sample) and picks the batch split\n");
  printf("-m modelfile\t\tLike -l with the given cost model file\n");
  printf("-b #\t\t\tSieve as # views of the range per batched round (one persistent launch for all)\n");
  printf("-c\t\t\tEncode k as the reference's signed int (k >= 2^31 wraps)\n");
  printf("-F file\t\t\tWrite the fraction done to file every 2 seconds (a BOINC wrapper's fraction_done_filename)\n");
  printf("-T file\t\t\tWrite the daily trickle-up message to file (a BOINC wrapper's trickle_up_filename)\n");
  printf("\t\t\tWithout -d the device is init_data.xml's <gpu_device_num> when that file is present\n");
  printf("-s\t\t\tRun the built-in self test\n");
  printf("-h\t\t\tPrint this help\n");
  exit(EXIT_FAILURE);
}

// main.cpp:39-70: unsigned 64-bit value with an optional e<exp> suffix
int parse_uint64(uint64_t* result, const char* str, uint64_t lo, uint64_t hi) {
  if (!str || !*str) return -1;
  char* tail;
  errno = 0;
  uint64_t num = strtoull(str, &tail, 10);
  if (errno != 0) return -2;
  if (*tail == 'e' || *tail == 'E') {
    tail++;
    errno = 0;
    uint64_t exp = strtoul(tail, &tail, 10);
    if (errno != 0) return -2;
    if (*tail != '\0') return -1;
    while (exp-- > 0) {
      if (num > hi / 10) return -2;
      num *= 10;
    }
  } else if (*tail != '\0') {
    return -1;
  }
  if (num < lo || num > hi) return -2;
  *result = num;
  return 0;
}

// main.cpp:83-134: whitespace-split tokens; unknown tokens are ignored
void parse_cmdline(int argc, char** argv, Options* o) {
  std::string joined;
  for (int i = 1; i < argc; i++) {
    joined += argv[i];
    if (i < argc - 1) joined += " ";
  }
  std::vector<char> buf(joined.begin(), joined.end());
  buf.push_back('\0');
  char* save = nullptr;
  char* tok = strtok_r(buf.data(), " \t", &save);
  while (tok) {
    if (!strcmp(tok, "-p")) {
      tok = strtok_r(nullptr, " \t", &save);
      if (tok && parse_uint64(&o->pmin, tok, 3, 0xFFFFFFFFFFFFFFFFull) != 0) {
        fprintf(stderr, "Invalid value for -p: %s\n", tok);
        printf("Invalid value for -p: %s\n", tok);
      }
    } else if (!strcmp(tok, "-P")) {
      tok = strtok_r(nullptr, " \t", &save);
      if (tok && parse_uint64(&o->pmax, tok, 4, 0xFFFFFFFFFFFFFFFFull) != 0) {
        fprintf(stderr, "Invalid value for -P: %s\n", tok);
        printf("Invalid value for -P: %s\n", tok);
      }
    } else if (!strcmp(tok, "-i")) {
      tok = strtok_r(nullptr, " \t", &save);
      if (tok) o->input = tok;
    } else if (!strcmp(tok, "-f")) {
      tok = strtok_r(nullptr, " \t", &save);
      if (tok) o->factor_file = tok;
    } else if (!strcmp(tok, "-d")) {
      tok = strtok_r(nullptr, " \t", &save);
      if (tok) {
        o->device = atoi(tok);
        o->device_set = true;
      }
    } else if (!strcmp(tok, "-s")) {
      o->selftest = true;
    } else if (!strcmp(tok, "-t")) {
      o->autotune = true;
    } else if (!strcmp(tok, "-l")) {
      o->learned = true;
    } else if (!strcmp(tok, "-m")) {
      tok = strtok_r(nullptr, " \t", &save);
      if (tok) {
        o->model_path = tok;
        o->learned = true;
      }
    } else if (!strcmp(tok, "-b")) {
      tok = strtok_r(nullptr, " \t", &save);
      uint64_t v = 0;
      if (tok && parse_uint64(&v, tok, 1, 256) == 0) o->views = (uint32_t)v;
      else if (tok) {
        fprintf(stderr, "Invalid value for -b: %s\n", tok);
        printf("Invalid value for -b: %s\n", tok);
      }
    } else if (!strcmp(tok, "-c")) {
      o->compat = true;
    } else if (!strcmp(tok, "-F")) {
      tok = strtok_r(nullptr, " \t", &save);
      if (tok) o->frac_file = tok;
    } else if (!strcmp(tok, "-T")) {
      tok = strtok_r(nullptr, " \t", &save);
      if (tok) o->trickle_file = tok;
    } else if (!strcmp(tok, "-h")) {
      help();
    }
    if (!tok) break;
    tok = strtok_r(nullptr, " \t", &save);
  }
}

// cl_sieve.cpp:646-659
bool is_perfect_power(uint64_t b, uint32_t* base) {
  for (uint64_t a = 2; a * a <= b; ++a) {
    uint64_t p = a * a;
    while (p < b) p *= a;
    if (p == b) {
      *base = (uint32_t)a;
      return true;
    }
  }
  return false;
}

[[noreturn]] void die(const char* what) {
  fprintf(stderr, "%s\n", what);
  printf("%s\n", what);
  exit(EXIT_FAILURE);
}

[[noreturn]] void die_lib(const char* ctx_msg) {
  std::string m = std::string(ctx_msg) + ": " + rs_last_error();
  die(m.c_str());
}

void format_rate(double val, char* buf) {
  const char* units[] = {"", "k", "M", "G", "T", "P"};
  int u = 0;
  while (val >= 1000.0 && u < 5) {
    val /= 1000.0;
    u++;
  }
  sprintf(buf, "%6.2f %sp/s", val, units[u]);
}

void format_eta(double seconds, char* buf) {
  int h = (int)(seconds / 3600);
  int m = ((int)seconds % 3600) / 60;
  int s = (int)seconds % 60;
  sprintf(buf, "%02d:%02d:%02d", h, m, s);
}

// cl_sieve.cpp:86-144: progress bar with an exponentially smoothed p/s
void print_progress(const rs_state& st, double* smooth, time_t start_time, uint64_t run_start_p) {
  const int bar_width = 40;
  const uint64_t total = st.pmax - st.pmin, done = st.p - st.pmin;
  double progress = total ? (double)done / (double)total : 0.0;
  progress = std::min(1.0, std::max(0.0, progress));
  const int pos = (int)(progress * bar_width);
  const double elapsed = difftime(time(nullptr), start_time);
  const uint64_t done_run = st.p > run_start_p ? st.p - run_start_p : 0;
  const double inst = (double)done_run / (elapsed > 0 ? elapsed : 1.0);
  *smooth = (*smooth == 0.0) ? inst : 0.9 * (*smooth) + 0.1 * inst;
  const uint64_t remaining = st.pmax > st.p ? st.pmax - st.p : 0;
  const double remain = (double)remaining / (*smooth > 0.0 ? *smooth : 1.0);
  char rate[32], eta[32];
  format_rate(*smooth, rate);
  format_eta(remain, eta);
  printf("\r[");
  for (int i = 0; i < bar_width; i++) putchar(i < pos ? '#' : '-');
  printf("] %6.2f%% | %s | ETA %s", progress * 100.0, rate, eta);
  fflush(stdout);
}

// model/cost_model.bin in the directory of this executable
std::string default_model_path() {
  char buf[4096];
  const ssize_t n = readlink("/proc/self/exe", buf, sizeof(buf) - 1);
  if (n <= 0) return "model/cost_model.bin";
  buf[n] = '\0';
  std::string d(buf);
  const size_t slash = d.rfind('/');
  return (slash == std::string::npos ? std::string(".") : d.substr(0, slash)) + "/model/cost_model.bin";
}

const char* algo_name(int32_t a) {
  return a == RS_ALGO_CLASSES ? "residue-class tables" : a == RS_ALGO_MERGED ? "merged table" : "three lists";
}

struct Run {
  Options o;
  rs_abcd* abcd = nullptr;
  rs_ctx* ctx = nullptr;
  rs_job* job = nullptr;
  std::vector<rs_job*> views;  // the job's extra views of a batched run
  rs_state st{};
  int write_a_next = 1;
  uint64_t factor_cap = 0;
  uint64_t base_primecount = 0, base_dups = 0;  // totals restored from a checkpoint
  std::vector<rs_factor> fbuf;
  std::chrono::steady_clock::time_point t_start = std::chrono::steady_clock::now();  // this run's start
};

// cl_sieve.cpp:530-643: drain, sort by (p, n), verify on the CPU, append
void get_results(Run& r, bool print) {
  if (print) {
    printf("\r                                                                                \r");
    fflush(stdout);
  }
  uint64_t n = 0;
  int rc = rs_factors_drain(r.job, r.fbuf.data(), r.fbuf.size(), &n);
  if (rc == RS_E_OVERFLOW) {
    fprintf(stderr, "Error: %s\n", rs_last_error());
    exit(EXIT_FAILURE);
  }
  if (rc) die_lib("rs_factors_drain");
  rs_counters c;
  if (rs_counters_read(r.job, &c)) die_lib("rs_counters_read");
  if (c.seg_overflow) die("error: gpu prime array overflow");
  r.st.primecount = r.base_primecount + c.primecount;
  r.st.dupcount = r.base_dups + c.dups;
  if (n) {
    std::sort(r.fbuf.begin(), r.fbuf.begin() + n, [](const rs_factor& a, const rs_factor& b) {
      if (a.p != b.p) return a.p < b.p;
      if (a.n != b.n) return a.n < b.n;
      return a.kidx < b.kidx;
    });
    if (print) printf("Verifying %" PRIu64 " factors on CPU...\n", n);
    uint32_t prp = 0;
    FILE* res = nullptr;
    for (uint64_t i = 0; i < n; i++) {
      const rs_factor& f = r.fbuf[i];
      uint32_t k;
      int32_t cc;
      rs_job_sequence(r.job, f.kidx, &k, &cc);
      const int v = rs_verify_factor(f.p, k, f.n, cc, r.st.base);
      if (v == 0) {
        fprintf(stderr, "CPU factor verification failed!  %" PRIu64 " is not a factor of %u*%u^%u%+d\n", f.p, k,
                r.st.base, f.n, cc);
        printf("CPU factor verification failed!  %" PRIu64 " is not a factor of %u*%u^%u%+d\n", f.p, k, r.st.base,
               f.n, cc);
        exit(EXIT_FAILURE);
      } else if (v == -1) {
        ++prp;
      } else {
        ++r.st.factorcount;
        if (!res) {
          res = fopen(r.o.factor_file.c_str(), "a");
          if (!res) {
            fprintf(stderr, "Cannot open %s !!!\n", r.o.factor_file.c_str());
            exit(EXIT_FAILURE);
          }
        }
        char line[128];
        rs_format_factor(line, sizeof(line), f.p, k, r.st.base, f.n, cc);
        if (fputs(line, res) < 0) {
          fprintf(stderr, "Cannot write to %s !!!\n", r.o.factor_file.c_str());
          exit(EXIT_FAILURE);
        }
        r.st.checksum += f.p + k + f.n + (uint64_t)(int64_t)cc;
      }
    }
    if (res) fclose(res);
    fprintf(stderr, "Verified %" PRIu64 " factors.\n", n - prp);
    if (print) printf("Verified %" PRIu64 " factors.\n", n - prp);
  }
  // checkpoint (cl_sieve.cpp:321-328), with the daily trickle-up message
  // (cl_sieve.cpp:40-62) when -T names its file: progress, this process's
  // CPU time and this run's elapsed time
  if (!r.o.trickle_file.empty() && rs_boinc::trickle_due((uint64_t)time(nullptr), &r.st.last_trickle)) {
    const double run = std::chrono::duration<double>(std::chrono::steady_clock::now() - r.t_start).count();
    const std::string msg = rs_boinc::trickle_message(rs_boinc::fraction_done(r.st.pmin, r.st.pmax, r.st.p),
                                                      (double)clock() / CLOCKS_PER_SEC, run);
    if (!rs_boinc::write_file_atomic(r.o.trickle_file, msg))
      fprintf(stderr, "Cannot write %s\n", r.o.trickle_file.c_str());
  }
  if (rs_state_write(".", &r.st, &r.write_a_next)) fprintf(stderr, "%s\n", rs_last_error());
  if (print) printf("Checkpoint, current p: %" PRIu64 "\n", r.st.p);
}

// cl_sieve.cpp:797-863
void read_factors_update_bitmap(Run& r) {
  FILE* fp = fopen(r.o.factor_file.c_str(), "r");
  if (!fp) {
    fprintf(stderr, "Cannot open %s !!!\n", r.o.factor_file.c_str());
    printf("Cannot open %s !!!\n", r.o.factor_file.c_str());
    exit(EXIT_FAILURE);
  }
  char line[512];
  uint64_t ln = 0;
  while (fgets(line, sizeof(line), fp)) {
    ln++;
    uint64_t p;
    uint32_t k, n;
    char sign;
    if (rs_parse_factor(line, r.st.base, &p, &k, &sign, &n)) {
      fprintf(stderr, "Invalid line read in factors file %" PRIu64 ": %s", ln, line);
      printf("Invalid line read in factors file %" PRIu64 ": %s", ln, line);
      exit(EXIT_FAILURE);
    }
    rs_abcd_mark_used(r.abcd, k, sign, n, r.o.compat ? 1 : 0);
  }
  fclose(fp);
  if (ln < r.st.factorcount) {
    fprintf(stderr, "ERROR: Missing factors in %s !!!\n", r.o.factor_file.c_str());
    printf("ERROR: Missing factors in %s !!!\n", r.o.factor_file.c_str());
    exit(EXIT_FAILURE);
  }
  printf("Updated bitmap with %" PRIu64 " factors from %s\n", ln, r.o.factor_file.c_str());
}

// cl_sieve.cpp:738-795
void finalize(Run& r) {
  if (r.st.factorcount) {
    FILE* f = fopen(r.o.factor_file.c_str(), "r");
    if (!f) {
      fprintf(stderr, "Cannot open %s !!!\n", r.o.factor_file.c_str());
      exit(EXIT_FAILURE);
    }
    char line[256];
    uint64_t lc = 0;
    while (fgets(line, sizeof(line), f)) ++lc;
    fclose(f);
    if (lc < r.st.factorcount) {
      fprintf(stderr, "ERROR: Missing factors in %s !!!\n", r.o.factor_file.c_str());
      printf("ERROR: Missing factors in %s !!!\n", r.o.factor_file.c_str());
      exit(EXIT_FAILURE);
    }
  }
  FILE* f = fopen(r.o.factor_file.c_str(), "a");
  if (!f) {
    fprintf(stderr, "Cannot open %s !!!\n", r.o.factor_file.c_str());
    exit(EXIT_FAILURE);
  }
  if (!r.st.factorcount && fprintf(f, "no factors\n") < 0) {
    fprintf(stderr, "Cannot write to %s !!!\n", r.o.factor_file.c_str());
    exit(EXIT_FAILURE);
  }
  fclose(f);
}

int sieve(Options o) {
  Run r;
  r.o = o;
  if (o.input.empty()) die("ERROR: no input file specified!");
  if (rs_abcd_read(o.input.c_str(), &r.abcd)) {
    fprintf(stderr, "%s\n", rs_last_error());
    exit(EXIT_FAILURE);
  }
  uint32_t base, nmin, nmax;
  int32_t kcount;
  uint64_t wps, nterms;
  rs_abcd_info(r.abcd, &base, &kcount, &nmin, &nmax, &wps, &nterms);
  printf("Sequences read: %d\n", kcount);
  fprintf(stderr, "Sequences read: %d\n", kcount);
  printf("Base = %" PRIu32 "\n", base);
  fprintf(stderr, "Base = %" PRIu32 "\n", base);
  if (r.o.factor_file.empty()) r.o.factor_file = "psp_sr2sieve.out";  // cl_sieve.cpp:992-994

  // setupSearch (cl_sieve.cpp:661-689)
  uint32_t bp;
  if (is_perfect_power(base, &bp)) {
    printf("\nerror: selected base %u is a perfect power of base %u!\n", base, bp);
    fprintf(stderr, "error: selected base %u is a perfect power of base %u!\n", base, bp);
    exit(EXIT_FAILURE);
  }
  if (!o.pmin || !o.pmax) {
    printf("\n-p and -P arguments are required\nuse -h for help\n");
    fprintf(stderr, "-p and -P arguments are required\n");
    exit(EXIT_FAILURE);
  }
  if (o.pmin > o.pmax) {
    printf("pmin <= pmax is required\nuse -h for help\n");
    fprintf(stderr, "pmin <= pmax is required\n");
    exit(EXIT_FAILURE);
  }
  rs_state& st = r.st;
  st.pmin = o.pmin;
  st.pmax = o.pmax;
  st.p = o.pmin;
  st.base = base;
  st.nmin = nmin;
  st.nmax = nmax;
  st.kcount = kcount;
  std::vector<uint32_t> ks(kcount);
  std::vector<int32_t> cs(kcount);
  rs_abcd_sequences(r.abcd, ks.data(), cs.data());
  for (int i = 0; i < kcount; i++) st.klist[i] = (cs[i] > 0) ? (int32_t)ks[i] : (int32_t)(0u - ks[i]);
  fprintf(stderr, "Starting sieve at p: %" PRIu64 " n: %u\nStopping sieve at P: %" PRIu64 " N: %u\n", st.pmin,
          st.nmin, st.pmax, st.nmax);
  printf("Starting sieve at p: %" PRIu64 " n: %u\nStopping sieve at P: %" PRIu64 " N: %u\n", st.pmin, st.nmin,
         st.pmax, st.nmax);

  // the device: -d, else the BOINC slot's init_data.xml (main.cpp:207-229)
  int device = o.device;
  if (!o.device_set) {
    int d = 0;
    const int got = rs_boinc::device_from_init_data("init_data.xml", &d);
    if (got > 0) {
      device = d;
      printf("Using device %d from init_data.xml\n", device);
    } else if (got < 0) {
      fprintf(stderr, "Error: init_data.xml holds no valid gpu_device_num\n");
      printf("Error: init_data.xml holds no valid gpu_device_num\n");
      exit(EXIT_FAILURE);
    } else {
      printf("init_data.xml not found, using device 0.\n");
    }
  }
  if (rs_open(device, &r.ctx)) die_lib("rs_open");
  uint32_t cus, lds;
  char name[256];
  rs_device_info(r.ctx, &cus, &lds, name, sizeof(name));
  printf("GPU Info:\n  Name: \t\t%s\n  Compute Units: \t%u\n  Local Mem Size: \t%u bytes\n", name, cus, lds);
  fprintf(stderr, "GPU Info:\n  Name: \t\t%s\n  Compute Units: \t%u\n  Local Mem Size: \t%u bytes\n", name, cus, lds);

  // resume or start (cl_sieve.cpp:1035-1074)
  const int resumed = rs_state_read(".", &st, &r.write_a_next);
  if (resumed > 0) {
    printf("Resuming from checkpoint in %s\n", r.write_a_next ? "stateB.ckp" : "stateA.ckp");
    printf("Current p: %" PRIu64 "\n", st.p);
    fprintf(stderr, "Resuming from checkpoint, current p: %" PRIu64 "\n", st.p);
    if (st.p == st.pmax) {
      printf("Workunit complete.\n");
      fprintf(stderr, "Workunit complete.\n");
      return EXIT_SUCCESS;
    }
    read_factors_update_bitmap(r);
  } else {
    FILE* t = fopen(r.o.factor_file.c_str(), "w");
    if (!t) {
      fprintf(stderr, "Cannot open %s !!!\n", r.o.factor_file.c_str());
      exit(EXIT_FAILURE);
    }
    fclose(t);
    st.last_trickle = (uint64_t)time(nullptr);
  }
  r.base_primecount = st.primecount;
  r.base_dups = st.dupcount;

  rs_params prm;
  if (rs_params_reference(nmin, nmax, kcount, 65536, cus, &prm)) die_lib("rs_params_reference");
  prm.k31_compat = o.compat ? 1 : 0;
  prm.factor_cap = 10000000ull;                    // main.cpp:167
  if (st.p < UINT32_MAX) prm.factor_cap *= 10;     // cl_sieve.cpp:1081-1083
  r.factor_cap = prm.factor_cap;
  if (rs_job_create(r.ctx, base, kcount, ks.data(), cs.data(), nmin, nmax, rs_abcd_bitmap(r.abcd), wps, &prm,
                    &r.job))
    die_lib("rs_job_create");
  uint32_t nviews = o.views;
  if (o.learned) {
    // the cost model's configuration and batch split (cl_sieve.cpp:885-973's
    // fixed rule stays the fallback)
    const std::string path = o.model_path.empty() ? default_model_path() : o.model_path;
    rs_model* model = nullptr;
    if (rs_model_load(r.ctx, path.c_str(), &model)) {
      fprintf(stderr, "warning: %s; using the reference configuration\n", rs_last_error());
      printf("warning: %s; using the reference configuration\n", rs_last_error());
    } else {
      std::vector<rs_params> cand(128);
      int nc = 0;
      if (rs_tune_candidates(r.ctx, nmin, nmax, kcount, cand.data(), (int)cand.size(), &nc)) die_lib("rs_tune_candidates");
      nc = std::min<int>(nc, (int)cand.size());
      const uint32_t span = nmax - nmin + 1;
      const double log2p = std::log2((double)st.p);
      // the model's three fastest candidates, timed on a sample at the start
      // of the range when it is long enough (about 2^20 primes; the job's
      // bitmap and counters are restored), the fastest kept
      int top[3] = {0, 0, 0}, ntop = 1, idx = 0;
      if (rs_model_rank(model, base, kcount, nterms, span, log2p, cand.data(), nc, 0, 3, top, &ntop)) {
        fprintf(stderr, "warning: %s; using the reference configuration\n", rs_last_error());
        ntop = 1;
        top[0] = 0;
      }
      idx = top[0];
      const uint64_t width = (uint64_t)(1048576.0 * std::log((double)st.p));
      if (ntop > 1 && st.pmax - st.p > 4 * width) {
        std::vector<rs_params> tc(ntop);
        for (int i = 0; i < ntop; i++) {
          tc[i] = cand[top[i]];
          tc[i].k31_compat = prm.k31_compat;
          tc[i].factor_cap = prm.factor_cap;
        }
        std::vector<double> secs(ntop);
        std::vector<uint64_t> pr(ntop);
        int best = 0;
        if (rs_autotune(r.job, st.p, width, tc.data(), ntop, secs.data(), pr.data(), &best)) die_lib("rs_autotune");
        idx = top[best];
        printf("Cost model: candidates %d, %d, %d timed on a sample: %.3g, %.3g, %.3g ns per prime\n", top[0],
               ntop > 1 ? top[1] : -1, ntop > 2 ? top[2] : -1, 1e9 * secs[0] / (double)std::max<uint64_t>(pr[0], 1),
               ntop > 1 ? 1e9 * secs[1] / (double)std::max<uint64_t>(pr[1], 1) : 0.0,
               ntop > 2 ? 1e9 * secs[2] / (double)std::max<uint64_t>(pr[2], 1) : 0.0);
      }
      rs_params pick = cand[idx];
      pick.k31_compat = prm.k31_compat;
      pick.factor_cap = prm.factor_cap;
      uint32_t mv = 0, msp = 0;
      if (rs_model_split(model, base, kcount, nterms, span, log2p, &mv, &msp) == RS_OK) {
        pick.seg_primes = msp;
        if (!nviews) nviews = mv;
      } else if (!nviews) {
        nviews = 32;  // no split head: the benchmark's batch
      }
      if (rs_job_reconfigure(r.job, &pick)) die_lib("rs_job_reconfigure");
      prm = pick;
      printf("Cost model: candidate %d of %d (%s), %u target primes per segment\n", idx, nc, algo_name(prm.algo),
             prm.seg_primes);
      rs_model_free(model);
    }
  }
  if (o.autotune && st.pmax - st.p > 4096) {
    // -t: time every candidate configuration on a short sample at the start of
    // the range (the job's bitmap and counters are restored afterwards)
    std::vector<rs_params> cand(128);
    int nc = 0;
    if (rs_tune_candidates(r.ctx, nmin, nmax, kcount, cand.data(), (int)cand.size(), &nc)) die_lib("rs_tune_candidates");
    nc = std::min<int>(nc, (int)cand.size());
    for (int i = 0; i < nc; i++) {
      cand[i].k31_compat = prm.k31_compat;
      cand[i].factor_cap = prm.factor_cap;
    }
    // about 2^16 primes: width ~ 2^16 * ln p
    const double lnp = std::log((double)st.p);
    uint64_t width = (uint64_t)(65536.0 * lnp);
    if (width > st.pmax - st.p) width = st.pmax - st.p;
    std::vector<double> secs(nc);
    std::vector<uint64_t> pr(nc);
    int best = 0;
    if (rs_autotune(r.job, st.p, width, cand.data(), nc, secs.data(), pr.data(), &best)) die_lib("rs_autotune");
    prm = cand[best];
    printf("Autotuned: candidate %d of %d (%s)\n", best, nc, algo_name(prm.algo));
  }
  if (!nviews) nviews = 1;
  // a batched run: the range of every round cut into nviews parts, one per
  // view of the job (they share its bitmap, counters and factor list)
  for (uint32_t v = 1; v < nviews; v++) {
    rs_job* h = nullptr;
    if (rs_job_view(r.job, &h)) die_lib("rs_job_view");
    r.views.push_back(h);
  }
  std::vector<rs_job*> handles(1, r.job);
  handles.insert(handles.end(), r.views.begin(), r.views.end());
  std::vector<uint64_t> vlo(nviews), vhi(nviews);
  if (nviews > 1) printf("Batched: %u views per round\n", nviews);
  uint32_t M, QQ, MM;
  rs_params_derive(&prm, nmin, nmax, &M, &QQ, &MM);
  printf("Q:%u M:%u QQ:%u MM:%u HT:%u LS:%u\n", prm.Q, M, QQ, MM, prm.hsize, prm.threads);
  fprintf(stderr, "Q:%u M:%u QQ:%u MM:%u HT:%u LS:%u\n", prm.Q, M, QQ, MM, prm.hsize, prm.threads);
  r.fbuf.resize(r.factor_cap);

  const time_t start_time = time(nullptr);
  time_t last_prog = start_time, last_ckpt = start_time;
  double smooth = 0;
  const uint64_t run_start_p = st.p;
  fprintf(stderr, "Starting Sieve...\n");
  printf("Starting Sieve...\n");
  const auto t0 = std::chrono::steady_clock::now();
  // Chunks of the p-range between host checks; the width adapts so that a
  // chunk takes about a second and cannot fill the factor list.
  uint64_t chunk = 1u << 20;
  while (st.p < st.pmax) {
    uint64_t stop = st.p + chunk;
    if (stop > st.pmax || stop < st.p) stop = st.pmax;
    const auto c0 = std::chrono::steady_clock::now();
    if (nviews == 1) {
      if (rs_sieve(r.job, st.p, stop)) die_lib("rs_sieve");
    } else {
      const uint64_t part = (stop - st.p) / nviews;
      for (uint32_t v = 0; v < nviews; v++) {
        vlo[v] = st.p + v * part;
        vhi[v] = v + 1 < nviews ? st.p + (v + 1) * part : stop;
      }
      if (rs_sieve_batch(handles.data(), (int)nviews, vlo.data(), vhi.data())) die_lib("rs_sieve_batch");
    }
    rs_counters c;
    if (rs_counters_read(r.job, &c)) die_lib("rs_counters_read");
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
    st.p = stop;
    if (dt < 0.5 && c.nfactors < r.factor_cap / 8 && chunk < (1ull << 62)) chunk *= 2;
    else if (dt > 2.0 && chunk > (1u << 16)) chunk /= 2;
    const time_t now = time(nullptr);
    if (now - last_prog > 1) {  // every 2 seconds (cl_sieve.cpp:1406-1414)
      print_progress(st, &smooth, start_time, run_start_p);
      if (!o.frac_file.empty()) rs_boinc::write_fraction_done(o.frac_file, rs_boinc::fraction_done(st.pmin, st.pmax, st.p));
      last_prog = now;
    }
    if (now - last_ckpt > 60 || c.nfactors > r.factor_cap / 2) {
      get_results(r, true);
      last_ckpt = now;
    }
  }
  st.p = st.pmax;
  get_results(r, true);
  finalize(r);
  if (!o.frac_file.empty() && !rs_boinc::write_fraction_done(o.frac_file, 1.0))
    fprintf(stderr, "Cannot write %s\n", o.frac_file.c_str());
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  fprintf(stderr, "Sieve complete.\nfactors %" PRIu64 ", prime count %" PRIu64 ", duplicate factors %" PRIu64 "\n",
          st.factorcount, st.primecount, st.dupcount);
  printf("Sieve finished in %d sec.\n", (int)secs);
  printf("Primes tested per second: %.0f\n", secs > 0 ? (double)(st.primecount - r.base_primecount) / secs : 0.0);
  printf("factors %" PRIu64 ", prime count %" PRIu64 ", checksum %016" PRIX64 ", duplicate factors %" PRIu64 "\n",
         st.factorcount, st.primecount, st.checksum, st.dupcount);
  for (rs_job* v : r.views) rs_job_destroy(v);
  rs_job_destroy(r.job);
  rs_close(r.ctx);
  rs_abcd_free(r.abcd);
  return EXIT_SUCCESS;
}

// Built-in self test (the reference's -s is a stub, cl_sieve.cpp:1462-1469):
// the known factor 1000133 | 91549*2^54378870+1 (cl_sieve.cpp:826-827) must be
// found, with every reported factor exact.
int selftest(int device) {
  rs_ctx* ctx;
  if (rs_open(device, &ctx)) die_lib("rs_open");
  const uint32_t k = 91549;
  const int32_t c = 1;
  const uint32_t n0 = 54378870, nmin = n0 - 1000, nmax = n0 + 1000;  // nmin even
  const uint64_t wps = ((uint64_t)nmax - nmin + 1 + 31) / 32;
  std::vector<uint32_t> bm(wps, 0xFFFFFFFFu);
  rs_params prm;
  rs_params_reference(nmin, nmax, 1, 65536, 1, &prm);
  rs_job* job;
  if (rs_job_create(ctx, 2, 1, &k, &c, nmin, nmax, bm.data(), wps, &prm, &job)) die_lib("rs_job_create");
  if (rs_sieve(job, 1000000, 1001000)) die_lib("rs_sieve");
  std::vector<rs_factor> f(1 << 16);
  uint64_t n = 0;
  if (rs_factors_drain(job, f.data(), f.size(), &n)) die_lib("rs_factors_drain");
  bool found = false, exact = true;
  for (uint64_t i = 0; i < n; i++) {
    if (f[i].p == 1000133 && f[i].n == n0) found = true;
    if (rs_verify_factor(f[i].p, k, f[i].n, c, 2) != 1) exact = false;
  }
  rs_job_destroy(job);
  rs_close(ctx);
  printf("self test: %" PRIu64 " factors, 1000133 | 91549*2^54378870+1 %s, all exact: %s\n", n,
         found ? "found" : "MISSING", exact ? "yes" : "NO");
  return (found && exact) ? EXIT_SUCCESS : EXIT_FAILURE;
}

}  // namespace

int main(int argc, char** argv) {
  fprintf(stderr, "\nrssieve (HIP, gfx950)\n");
  printf("\nrssieve (HIP, gfx950)\n");
  fprintf(stderr, "Command line: ");
  for (int i = 0; i < argc; i++) fprintf(stderr, "%s ", argv[i]);
  fprintf(stderr, "\n");
  Options o;
  parse_cmdline(argc, argv, &o);
  if (o.selftest) return selftest(o.device);
  return sieve(o);
}
