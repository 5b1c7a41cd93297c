// Host checks of rs_boinc.h (the BOINC shim of the command line): the
// init_data.xml device tag, the fraction done and its file, the trickle-up
// message and its once-per-day cadence.  Prints "ok" and exits 0 when every
// check passes.
#include <cstdio>
#include <string>

#include "../../rssieve_amd/csrc/rs_boinc.h"

static int fails = 0;
#define CHECK(c)                                             \
  do {                                                       \
    if (!(c)) {                                              \
      fprintf(stderr, "FAIL %s:%d %s\n", __FILE__, __LINE__, #c); \
      fails++;                                               \
    }                                                        \
  } while (0)

static std::string slurp(const std::string& path) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return "<missing>";
  std::string s;
  char b[256];
  size_t n;
  while ((n = fread(b, 1, sizeof(b), f)) > 0) s.append(b, n);
  fclose(f);
  return s;
}

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  int d = -7;
  // the sample file of the reference (init_data.xml:1-4) and variants
  CHECK(rs_boinc::device_from_init_data_text("<app_init_data>\n<gpu_type>ATI</gpu_type>\n<gpu_device_num>0</gpu_device_num>\n</app_init_data>\n", &d) == 1 && d == 0);
  CHECK(rs_boinc::device_from_init_data_text("<gpu_device_num> 3 </gpu_device_num>", &d) == 1 && d == 3);
  CHECK(rs_boinc::device_from_init_data_text("<gpu_device_num>\n12\n</gpu_device_num>", &d) == 1 && d == 12);
  d = -7;
  CHECK(rs_boinc::device_from_init_data_text("<app_init_data><gpu_type>ATI</gpu_type></app_init_data>", &d) == 0 && d == -7);
  CHECK(rs_boinc::device_from_init_data_text("<gpu_device_num>x1</gpu_device_num>", &d) == -1);
  CHECK(rs_boinc::device_from_init_data_text("<gpu_device_num>-1</gpu_device_num>", &d) == -1);
  CHECK(rs_boinc::device_from_init_data_text("<gpu_device_num></gpu_device_num>", &d) == -1);
  CHECK(rs_boinc::device_from_init_data_text("<gpu_device_num>1", &d) == -1);
  CHECK(rs_boinc::device_from_init_data_text("<gpu_device_num>12345678901</gpu_device_num>", &d) == -1);
  CHECK(rs_boinc::device_from_init_data((dir + "/no_such_init_data.xml").c_str(), &d) == 0);
  {
    const std::string p = dir + "/init_data.xml";
    FILE* f = fopen(p.c_str(), "wb");
    fputs("<app_init_data>\n<gpu_device_num>5</gpu_device_num>\n</app_init_data>\n", f);
    fclose(f);
    CHECK(rs_boinc::device_from_init_data(p.c_str(), &d) == 1 && d == 5);
  }
  // fraction done of [pmin, pmax) at p
  CHECK(rs_boinc::fraction_done(100, 200, 100) == 0.0);
  CHECK(rs_boinc::fraction_done(100, 200, 50) == 0.0);
  CHECK(rs_boinc::fraction_done(100, 200, 150) == 0.5);
  CHECK(rs_boinc::fraction_done(100, 200, 200) == 1.0);
  CHECK(rs_boinc::fraction_done(100, 200, 300) == 1.0);
  CHECK(rs_boinc::fraction_done(200, 100, 150) == 0.0);
  CHECK(rs_boinc::write_fraction_done(dir + "/frac.txt", 0.25) && slurp(dir + "/frac.txt") == "0.250000\n");
  CHECK(rs_boinc::write_fraction_done(dir + "/frac.txt", 1.0) && slurp(dir + "/frac.txt") == "1.000000\n");
  CHECK(slurp(dir + "/frac.txt.tmp") == "<missing>");
  CHECK(!rs_boinc::write_fraction_done(dir + "/no_such_dir/frac.txt", 0.5));
  // the trickle-up message (the reference's variety rssieve_progress)
  CHECK(rs_boinc::trickle_message(0.5, 12.25, 3600.0) ==
        "<trickle_up>\n   <progress>0.500000</progress>\n   <cputime>12.250000</cputime>\n"
        "   <runtime>3600.000000</runtime>\n</trickle_up>\n");
  // once per day: strictly more than 86400 s since the last one
  uint64_t last = 1000;
  CHECK(!rs_boinc::trickle_due(1000, &last) && last == 1000);
  CHECK(!rs_boinc::trickle_due(1000 + 86400, &last) && last == 1000);
  CHECK(rs_boinc::trickle_due(1000 + 86401, &last) && last == 1000 + 86401);
  CHECK(!rs_boinc::trickle_due(1000 + 86402, &last));
  last = 5000;  // ahead of the clock: the unsigned difference is large, one is sent
  CHECK(rs_boinc::trickle_due(4000, &last) && last == 4000);
  last = 0;
  CHECK(rs_boinc::trickle_due(86401, &last) && last == 86401);
  CHECK(!rs_boinc::trickle_due(86401 + 50, &last, 50) && rs_boinc::trickle_due(86401 + 51, &last, 50));
  if (fails) return 1;
  printf("ok\n");
  return 0;
}
