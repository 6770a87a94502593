// Tuner of batched-tracking plan shapes (include/lsdg_tune.h): the class key
// of a batch, the process-wide table of tuned shapes with its JSON file, and
// the search lsdg_batch_tune runs.
//
// Search.  Shapes differ only in the order of the FP64 sums (the LM path is
// the same), so the search only times plans.  It starts from the default shape and walks one
// knob at a time, keeping a candidate only when its median run beats the
// best so far: the schedule (cooperative or steps); for the cooperative
// schedule the XCD grouping (a job's blocks on one XCD), the block size x
// blocks per job, then the split threshold; for the steps schedule the
// whole-job threshold, then the chains, then the points per thread of a
// reduction tile.  Candidates that the device caps turn into an already
// measured plan are not run twice.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "batch_internal.h"
#include "capi_common.h"

using namespace lsdg;

namespace {

std::mutex g_table_mutex;
std::map<std::string, lsdg_batch_config> g_table;

bool key_ok(const char* key) {
  if (!key) return false;
  size_t n = std::strlen(key);
  if (n == 0 || n >= LSDG_TUNE_KEY_BYTES) return false;
  for (size_t i = 0; i < n; ++i) {
    const char ch = key[i];
    if (!((ch >= 'a' && ch <= 'z') || (ch >= 'A' && ch <= 'Z') || (ch >= '0' && ch <= '9') || ch == '/' ||
          ch == '_' || ch == '-' || ch == '.'))
      return false;
  }
  return true;
}

bool config_ok(const lsdg_batch_config& c) {
  if (c.schedule != LSDG_SCHEDULE_COOP && c.schedule != LSDG_SCHEDULE_STEPS) return false;
  if (c.schedule == LSDG_SCHEDULE_COOP && !batch_coop_block_size_ok(c.coop_block)) return false;
  return c.coop_k >= 0 && c.solo_max >= 0 && c.job_max_points >= 0 && c.chains >= 1 && c.chains <= 8 &&
         c.points_per_thread >= 1 && c.points_per_thread <= 64 && c.xcd_groups >= 1 && c.xcd_groups <= 64;
}

int ceil_log2(long long v) {
  int k = 0;
  while ((1LL << k) < v) ++k;
  return k;
}

// ------------------------------------------------------------------ JSON
constexpr int kNumFields = 8;
const char* kFields[kNumFields] = {"schedule", "coop_block", "coop_k", "solo_max", "job_max_points", "chains",
                                   "points_per_thread", "xcd_groups"};

int32_t* field(lsdg_batch_config& c, int i) {
  int32_t* f[kNumFields] = {&c.schedule, &c.coop_block, &c.coop_k, &c.solo_max, &c.job_max_points, &c.chains,
                            &c.points_per_thread, &c.xcd_groups};
  return f[i];
}

// A minimal reader for the table file: objects, arrays, strings without
// escapes beyond \" and \\, integers.  Every object that holds a "key"
// string and all the shape fields becomes an entry.
struct Reader {
  const std::string& t;
  size_t i = 0;
  bool bad = false;
  std::vector<std::pair<std::string, lsdg_batch_config>> out;
  explicit Reader(const std::string& s) : t(s) {}

  void ws() {
    while (i < t.size() && (t[i] == ' ' || t[i] == '\n' || t[i] == '\r' || t[i] == '\t')) ++i;
  }
  bool str(std::string& s) {
    ws();
    if (i >= t.size() || t[i] != '"') return false;
    ++i;
    s.clear();
    while (i < t.size() && t[i] != '"') {
      if (t[i] == '\\' && i + 1 < t.size()) ++i;
      s.push_back(t[i++]);
    }
    if (i >= t.size()) return false;
    ++i;
    return true;
  }
  bool integer(long long& v) {
    ws();
    size_t j = i;
    if (j < t.size() && t[j] == '-') ++j;
    if (j >= t.size() || t[j] < '0' || t[j] > '9') return false;
    v = std::strtoll(t.c_str() + i, nullptr, 10);
    while (j < t.size() && t[j] >= '0' && t[j] <= '9') ++j;
    i = j;
    return true;
  }
  void value(int depth) {
    ws();
    if (bad || i >= t.size() || depth > 16) {
      bad = true;
      return;
    }
    if (t[i] == '{') {
      object(depth + 1);
    } else if (t[i] == '[') {
      ++i;
      ws();
      if (i < t.size() && t[i] == ']') {
        ++i;
        return;
      }
      while (!bad) {
        value(depth + 1);
        ws();
        if (i < t.size() && t[i] == ',') {
          ++i;
          continue;
        }
        if (i < t.size() && t[i] == ']') {
          ++i;
          return;
        }
        bad = true;
      }
    } else if (t[i] == '"') {
      std::string s;
      if (!str(s)) bad = true;
    } else {
      long long v;
      if (!integer(v)) bad = true;
    }
  }
  void object(int depth) {
    ++i;  // '{'
    std::string key;
    bool hasKey = false;
    lsdg_batch_config c{};
    int seen = 0;
    ws();
    if (i < t.size() && t[i] == '}') {
      ++i;
      return;
    }
    while (!bad) {
      std::string name;
      if (!str(name)) {
        bad = true;
        return;
      }
      ws();
      if (i >= t.size() || t[i] != ':') {
        bad = true;
        return;
      }
      ++i;
      ws();
      int fi = -1;
      for (int k = 0; k < kNumFields; ++k)
        if (name == kFields[k]) fi = k;
      if (name == "key") {
        hasKey = str(key);
        if (!hasKey) bad = true;
      } else if (fi >= 0) {
        long long v;
        if (!integer(v) || v < -(1LL << 31) || v >= (1LL << 31)) {
          bad = true;
          return;
        }
        *field(c, fi) = (int32_t)v;
        seen |= 1 << fi;
      } else {
        value(depth);
      }
      ws();
      if (i < t.size() && t[i] == ',') {
        ++i;
        continue;
      }
      if (i < t.size() && t[i] == '}') {
        ++i;
        break;
      }
      bad = true;
      return;
    }
    if (!bad && hasKey && seen == (1 << kNumFields) - 1) out.emplace_back(key, c);
  }
};

}  // namespace

bool lsdg::tune_table_lookup(const char* key, lsdg_batch_config* out) {
  std::lock_guard<std::mutex> lock(g_table_mutex);
  auto it = g_table.find(key);
  if (it == g_table.end()) return false;
  *out = it->second;
  return true;
}

extern "C" lsdg_status lsdg_tune_table_put(const char* key, const lsdg_batch_config* config) {
  if (!key_ok(key) || !config || !config_ok(*config)) return LSDG_EINVAL;
  std::lock_guard<std::mutex> lock(g_table_mutex);
  g_table[key] = *config;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_tune_table_get(const char* key, lsdg_batch_config* out) {
  if (!key || !out) return LSDG_EINVAL;
  return tune_table_lookup(key, out) ? LSDG_OK : LSDG_ESTATE;
}

extern "C" lsdg_status lsdg_tune_table_clear(void) {
  std::lock_guard<std::mutex> lock(g_table_mutex);
  g_table.clear();
  return LSDG_OK;
}

extern "C" int32_t lsdg_tune_table_size(void) {
  std::lock_guard<std::mutex> lock(g_table_mutex);
  return (int32_t)g_table.size();
}

extern "C" lsdg_status lsdg_tune_table_save(const char* path) {
  if (!path) return LSDG_EINVAL;
  std::string s = "{\"lsdg_tune\": 1, \"entries\": [";
  {
    std::lock_guard<std::mutex> lock(g_table_mutex);
    bool first = true;
    for (auto& kv : g_table) {
      s += first ? "\n  {" : ",\n  {";
      first = false;
      s += "\"key\": \"" + kv.first + "\"";
      lsdg_batch_config c = kv.second;
      for (int k = 0; k < kNumFields; ++k)
        s += ", \"" + std::string(kFields[k]) + "\": " + std::to_string(*field(c, k));
      s += "}";
    }
  }
  s += "\n]}\n";
  FILE* f = std::fopen(path, "wb");
  if (!f) return LSDG_EINVAL;
  const bool ok = std::fwrite(s.data(), 1, s.size(), f) == s.size();
  return (std::fclose(f) == 0 && ok) ? LSDG_OK : LSDG_EINVAL;
}

extern "C" lsdg_status lsdg_tune_table_load(const char* path) {
  if (!path) return LSDG_EINVAL;
  FILE* f = std::fopen(path, "rb");
  if (!f) return LSDG_EINVAL;
  std::string s;
  char buf[4096];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof buf, f)) > 0 && s.size() < (1u << 24)) s.append(buf, got);
  std::fclose(f);
  Reader r(s);
  r.value(0);
  if (r.bad) return LSDG_EINVAL;
  for (auto& e : r.out)
    if (!key_ok(e.first.c_str()) || !config_ok(e.second)) return LSDG_EINVAL;
  std::lock_guard<std::mutex> lock(g_table_mutex);
  for (auto& e : r.out) g_table[e.first] = e.second;
  return LSDG_OK;
}

// ------------------------------------------------------------------ key
extern "C" lsdg_status lsdg_batch_tune_key(lsdg_ctx c, const lsdg_job* jobs, int32_t n, char key[LSDG_TUNE_KEY_BYTES]) {
  if (!c || !jobs || !key || n < 1) return LSDG_EINVAL;
  long long maxP = 0;
  bool mixed = false;
  for (int j = 0; j < n; ++j) {
    if (!jobs[j].frame) return LSDG_EINVAL;
    const lsdg_frame f = jobs[j].frame;
    maxP = std::max(maxP, (long long)f->w[0] * f->h[0]);
    mixed |= f->w[0] != jobs[0].frame->w[0] || f->h[0] != jobs[0].frame->h[0];
  }
  static const char* kPrec[] = {"fp32", "fp16", "bf16"};
  const int prec = jobs[0].frame->precision;
  // compensated sums run different kernels: their own class
  std::snprintf(key, LSDG_TUNE_KEY_BYTES, "%s/%s%s/p%d/%s/n%d", jobs[0].kind == LSDG_JOB_SIM3 ? "sim3" : "se3",
                kPrec[prec >= 0 && prec <= 2 ? prec : 0], c->sums == LSDG_SUMS_KAHAN ? "-kahan" : "",
                ceil_log2(maxP), mixed ? "mixed" : "single", ceil_log2(n));
  return LSDG_OK;
}

// ------------------------------------------------------------------ search
namespace {

bool same_plan(const lsdg_batch_config& a, const lsdg_batch_config& b) {
  if (a.schedule != b.schedule) return false;
  if (a.schedule == LSDG_SCHEDULE_COOP)
    return a.coop_block == b.coop_block && a.coop_k == b.coop_k && a.solo_max == b.solo_max &&
           a.xcd_groups == b.xcd_groups;
  return a.job_max_points == b.job_max_points && a.chains == b.chains && a.points_per_thread == b.points_per_thread;
}

// Round-robin dispatch groups of the device: the XCDs of a gfx94x / gfx950
// part (8 on MI300X and MI355X), read from the device's CU count and the
// 32 CUs an XCD holds; 1 when that does not divide.
int device_xcds(lsdg_ctx c) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c->device) != hipSuccess) return 1;
  const std::string arch = prop.gcnArchName;
  if (arch.rfind("gfx950", 0) != 0 && arch.rfind("gfx942", 0) != 0) return 1;
  const int cus = prop.multiProcessorCount;
  return cus % 32 == 0 && cus / 32 > 1 ? cus / 32 : 1;
}

struct Search {
  lsdg_ctx c;
  const lsdg_job* jobs;
  int n;
  int runs;
  std::vector<std::pair<lsdg_batch_config, double>> seen;  // plans as taken, median ms

  // Median ms of `runs` runs of the plan the shape gives, after one warm-up
  // (which also sizes the plan's step graphs); *taken = the plan's shape.
  lsdg_status measure(const lsdg_batch_config& want, lsdg_batch_config* taken, double* ms) {
    lsdg_batch b = nullptr;
    lsdg_status st = batch_create_shaped(c, jobs, n, want, &b);
    if (st != LSDG_OK) return st;
    lsdg_batch_get_config(b, taken);
    for (auto& s : seen)
      if (same_plan(s.first, *taken)) {
        lsdg_batch_destroy(b);
        *ms = s.second;
        return LSDG_OK;
      }
    st = lsdg_batch_run(b, nullptr);
    std::vector<double> t;
    for (int r = 0; r < runs && st == LSDG_OK; ++r) {
      auto t0 = std::chrono::steady_clock::now();
      st = lsdg_batch_run(b, nullptr);
      t.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    lsdg_batch_destroy(b);
    if (st != LSDG_OK) return st;
    std::sort(t.begin(), t.end());
    *ms = t[t.size() / 2];
    seen.emplace_back(*taken, *ms);
    return LSDG_OK;
  }
};

}  // namespace

extern "C" lsdg_status lsdg_batch_tune(lsdg_ctx c, const lsdg_job* jobs, int32_t n, int32_t runs,
                                       lsdg_batch_config* best, double* default_ms, double* best_ms) {
  if (!c || !jobs || n < 1) return LSDG_EINVAL;
  char key[LSDG_TUNE_KEY_BYTES];
  lsdg_status st = lsdg_batch_tune_key(c, jobs, n, key);
  if (st != LSDG_OK) return st;
  const bool prof = c->prof_on;
  c->prof_on = false;  // time the graphs, not the profiled replay
  Search S{c, jobs, n, std::max(1, std::min(50, runs)), {}};
  lsdg_batch_config cur{}, taken{};
  double curMs = 0.0, ms = 0.0;
  auto done = [&](lsdg_status s) {
    c->prof_on = prof;
    return s;
  };
  const lsdg_batch_config def = batch_default_shape(n);
  if ((st = S.measure(def, &cur, &curMs)) != LSDG_OK) return done(st);
  if (default_ms) *default_ms = curMs;
  auto consider = [&](const lsdg_batch_config& want) -> lsdg_status {
    lsdg_status s = S.measure(want, &taken, &ms);
    if (s == LSDG_OK && ms < curMs) {
      cur = taken;
      curMs = ms;
    }
    return s;
  };
  // schedule
  {
    lsdg_batch_config s = def;
    s.schedule = def.schedule == LSDG_SCHEDULE_COOP ? LSDG_SCHEDULE_STEPS : LSDG_SCHEDULE_COOP;
    if ((st = consider(s)) != LSDG_OK) return done(st);
  }
  if (cur.schedule == LSDG_SCHEDULE_COOP) {
    {
      lsdg_batch_config s = cur;
      s.xcd_groups = device_xcds(c);
      if (s.xcd_groups > 1 && (st = consider(s)) != LSDG_OK) return done(st);
    }
    const lsdg_batch_config base = cur;
    for (int bs : {512, 256})
      for (int k : {2, 4, 8, 16}) {
        lsdg_batch_config s = base;
        s.coop_block = bs;
        s.coop_k = k;
        if ((st = consider(s)) != LSDG_OK) return done(st);
      }
    const lsdg_batch_config base2 = cur;
    for (int solo : {0, 512, 1024, 2048, 4096, 8192, 1 << 30}) {
      lsdg_batch_config s = base2;
      s.solo_max = solo;
      if ((st = consider(s)) != LSDG_OK) return done(st);
    }
  } else {
    for (int jmp : {0, 3072, 6144, 12288, 1 << 30}) {
      lsdg_batch_config s = cur;
      s.job_max_points = jmp;
      if ((st = consider(s)) != LSDG_OK) return done(st);
    }
    for (int ch : {1, 2, 3, 4}) {
      lsdg_batch_config s = cur;
      s.chains = ch;
      if ((st = consider(s)) != LSDG_OK) return done(st);
    }
    for (int ppt : {2, 4, 8}) {
      lsdg_batch_config s = cur;
      s.points_per_thread = ppt;
      if ((st = consider(s)) != LSDG_OK) return done(st);
    }
  }
  if ((st = lsdg_tune_table_put(key, &cur)) != LSDG_OK) return done(st);
  if (best) *best = cur;
  if (best_ms) *best_ms = curMs;
  return done(LSDG_OK);
}
