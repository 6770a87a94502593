// lsdg_ctx_profile / lsdg_ctx_kernel_stats: hipEvent pairs per timed launch
// (single kernels: the kernel's own start / stop events through
// hipExtLaunchKernel; multi-kernel sequences: events recorded on the stream
// around them), folded into per-kind totals.
#include <cstring>

#include "capi_common.h"

using namespace lsdg;

static hipEvent_t take_event(lsdg_ctx c) {
  if (!c->prof_events.empty()) {
    hipEvent_t e = c->prof_events.back();
    c->prof_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

ProfMark lsdg::prof_begin(lsdg_ctx c, int kind) {
  ProfMark m;
  if (!c->prof_on) return m;
  m.kind = kind;
  m.a = take_event(c);
  if (m.a && hipEventRecord(m.a, c->stream) != hipSuccess) m.kind = -1;
  return m;
}

void lsdg::prof_end(lsdg_ctx c, ProfMark& m, double bytes) {
  if (!c->prof_on || m.kind < 0 || !m.a) return;
  hipEvent_t b = take_event(c);
  if (!b || hipEventRecord(b, c->stream) != hipSuccess) return;
  c->prof_pending.push_back({m.kind, m.a, b, bytes});
  if (c->prof_pending.size() > 4096) (void)prof_resolve(c);
}

KTimer lsdg::prof_timer(lsdg_ctx c) {
  KTimer t;
  if (!c->prof_on) return t;
  t.a = take_event(c);
  t.b = t.a ? take_event(c) : nullptr;
  if (!t.b) {
    if (t.a) c->prof_events.push_back(t.a);
    t.a = nullptr;
  }
  return t;
}

void lsdg::prof_timer_done(lsdg_ctx c, int kind, const KTimer& t, double bytes) {
  if (!t.a || !t.b) return;
  c->prof_pending.push_back({kind, t.a, t.b, bytes});
  if (c->prof_pending.size() > 4096) (void)prof_resolve(c);
}

lsdg_status lsdg::prof_resolve(lsdg_ctx c) {
  if (c->prof_pending.empty()) return LSDG_OK;
  CK(c, hipStreamSynchronize(c->stream));
  for (auto& r : c->prof_pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      c->prof[r.kind].launches++;
      c->prof[r.kind].ms += ms;
      c->prof[r.kind].bytes += r.bytes;
    }
    c->prof_events.push_back(r.a);
    c->prof_events.push_back(r.b);
  }
  c->prof_pending.clear();
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_ctx_profile(lsdg_ctx c, int32_t enable) {
  if (!c) return LSDG_EINVAL;
  lsdg_status st = prof_resolve(c);
  if (st != LSDG_OK) return st;
  if (enable) std::memset(c->prof, 0, sizeof c->prof);
  c->prof_on = enable != 0;
  return LSDG_OK;
}

extern "C" lsdg_status lsdg_ctx_kernel_stats(lsdg_ctx c, int32_t kernel, lsdg_kernel_stats* out) {
  if (!c || !out || kernel < 0 || kernel >= LSDG_KERNEL_KINDS) return LSDG_EINVAL;
  lsdg_status st = prof_resolve(c);
  if (st != LSDG_OK) return st;
  *out = c->prof[kernel];
  return LSDG_OK;
}
