// rs_boinc.h -- what the reference's host takes from the BOINC runtime, for
// running the rssieve command line inside a BOINC client without the BOINC
// SDK (which this build does not link):
//
//  * the GPU of the task: <gpu_device_num> of the slot's init_data.xml, as
//    boinc_get_opencl_ids reads it (main.cpp:207-229), when the command line
//    names no device;
//  * the fraction done, every two seconds and at the end
//    (cl_sieve.cpp:1406-1414), written as one number to a file: the form in
//    which the BOINC wrapper's fraction_done_filename hands it to the client;
//  * the daily trickle-up message with progress, CPU time and run time
//    (cl_sieve.cpp:40-62, sent at each checkpoint when a day has passed since
//    the checkpoint's last_trickle), written to a file: the form of the
//    wrapper's trickle_up_filename.
//
// Header-only, C++17, no dependency beyond the C library.
#pragma once

#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

namespace rs_boinc {

// <gpu_device_num>N</gpu_device_num> in the text of an init_data.xml: 1 and
// *dev = N when the tag holds a non-negative decimal integer, 0 when the tag
// is absent, -1 when its value is malformed
inline int device_from_init_data_text(const std::string& xml, int* dev) {
  static const char kOpen[] = "<gpu_device_num>", kClose[] = "</gpu_device_num>";
  const size_t a = xml.find(kOpen);
  if (a == std::string::npos) return 0;
  const size_t v = a + sizeof(kOpen) - 1;
  const size_t b = xml.find(kClose, v);
  if (b == std::string::npos) return -1;
  size_t i = v, j = b;
  while (i < j && (xml[i] == ' ' || xml[i] == '\t' || xml[i] == '\n' || xml[i] == '\r')) i++;
  while (j > i && (xml[j - 1] == ' ' || xml[j - 1] == '\t' || xml[j - 1] == '\n' || xml[j - 1] == '\r')) j--;
  if (i == j || j - i > 9) return -1;
  int n = 0;
  for (size_t k = i; k < j; k++) {
    if (xml[k] < '0' || xml[k] > '9') return -1;
    n = 10 * n + (xml[k] - '0');
  }
  *dev = n;
  return 1;
}

// the same for a file: 0 also when the file cannot be read
inline int device_from_init_data(const char* path, int* dev) {
  FILE* f = fopen(path, "rb");
  if (!f) return 0;
  std::string xml;
  char buf[4096];
  size_t got;
  while ((got = fread(buf, 1, sizeof(buf), f)) > 0 && xml.size() < (1u << 20)) xml.append(buf, got);
  fclose(f);
  return device_from_init_data_text(xml, dev);
}

// the fraction done of [pmin, pmax) at p, clamped to [0, 1]
inline double fraction_done(uint64_t pmin, uint64_t pmax, uint64_t p) {
  if (pmax <= pmin || p <= pmin) return 0.0;
  if (p >= pmax) return 1.0;
  return (double)(p - pmin) / (double)(pmax - pmin);
}

// Replace path's content by text: written to path + ".tmp", then renamed, so
// a reader sees the old or the new content whole.  False on an I/O error.
inline bool write_file_atomic(const std::string& path, const std::string& text) {
  const std::string tmp = path + ".tmp";
  FILE* f = fopen(tmp.c_str(), "wb");
  if (!f) return false;
  const bool ok = fwrite(text.data(), 1, text.size(), f) == text.size();
  if (fclose(f) != 0 || !ok) {
    remove(tmp.c_str());
    return false;
  }
  return rename(tmp.c_str(), path.c_str()) == 0;
}

inline bool write_fraction_done(const std::string& path, double fd) {
  char buf[32];
  snprintf(buf, sizeof(buf), "%.6f\n", fd);
  return write_file_atomic(path, buf);
}

// the reference's trickle-up message (cl_sieve.cpp:51-56), variety
// "rssieve_progress"
inline std::string trickle_message(double progress, double cputime, double runtime) {
  char msg[512];
  snprintf(msg, sizeof(msg),
           "<trickle_up>\n"
           "   <progress>%lf</progress>\n"
           "   <cputime>%lf</cputime>\n"
           "   <runtime>%lf</runtime>\n"
           "</trickle_up>\n",
           progress, cputime, runtime);
  return msg;
}

// At a checkpoint: true, and *last = now, when more than period seconds have
// passed since *last (cl_sieve.cpp:43-44: once per day; the difference is
// unsigned, as there, so a last_trickle ahead of the clock also sends one)
inline bool trickle_due(uint64_t now, uint64_t* last, uint64_t period = 86400) {
  if (now - *last <= period) return false;
  *last = now;
  return true;
}

}  // namespace rs_boinc
