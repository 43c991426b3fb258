// p3x_server.hip -- the served form of include/p3x.h: continuous batching of
// P3 step requests from concurrent ensemble members.
//
// One scheduler thread owns every device call after creation.  It repeatedly
//   1. retires finished slots (their completion event has fired): each request
//      of the batch gets its latency and its own set's error status, its member
//      becomes free;
//   2. fills every free slot with a new batch: the queued requests, in arrival
//      order, whose member is neither in flight nor already in this batch and
//      whose scratch still fits the slot's share of the scratch budget, up to
//      max_batch of them; the batch's set descriptors, block offsets and one
//      zeroed error word per set go through the slot's pinned staging buffer to
//      its device buffer (one asynchronous copy on the slot's stream), followed
//      by the launches of the batched kernels, a read-back of the error words
//      and an event.
// Slots are independent non-blocking streams, so batches admitted while
// others run overlap on the device.  Members may register at any time; each
// is admitted against the HBM budget (the bytes its resident arrays span plus
// the slots' scratch).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "p3x_internal.hpp"

using namespace p3x_detail;

namespace {

double now_us() {
  using namespace std::chrono;
  return duration<double, std::micro>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
  unsigned long long tag;
  int member;
  double dt;
  double t_submit;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  char* d_arena = nullptr;  // scratch of the batch's sets
  size_t arena_bytes = 0;
  char* d_desc = nullptr;   // set descriptors, block offsets and the sets' error words
  char* h_stage = nullptr;  // pinned staging of d_desc
  size_t desc_bytes = 0;
  unsigned* h_err = nullptr;  // pinned, max_batch words: the error word of each set
  bool busy = false;
  std::vector<Request> reqs;
};

}  // namespace

struct p3x_server {
  p3x_context* ctx = nullptr;
  p3x_runtime rt{};
  p3x_server_options opt{};
  p3d::Shape shape{};
  int occ1 = 1, occ2 = 1;  // resident blocks per CU of the part 1 and part 2 kernels
  int occ_sed[p3d::kMaxLevelsPerLane + 1] = {};
  struct Member {
    p3d::KParams k;
    size_t scratch;   // scratch bytes of one step
    size_t resident;  // bytes its device arrays span
    bool busy;
    // a process member (p3x_server_add_process): its process, fields and
    // neighbour stages (has_chain: chain holds them)
    p3x_process* proc = nullptr;
    p3x_atm_fields fields{};
    p3x_step_chain chain{};
    bool has_chain = false;
  };
  std::vector<Member> members;
  double resident_bytes = 0;  // of all members
  double scratch_total = 0;   // the slots' arenas
  std::deque<Request> pending;
  std::vector<Slot> slots;
  std::deque<p3x_completion> done;
  std::vector<double> latencies;
  long long submitted = 0, completed = 0, in_flight = 0;
  double colev = 0, t_first = -1, t_last = 0;
  int fatal = P3X_OK;
  bool started = false, stop = false;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::thread thr;
};

namespace {

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of descriptors a batch of n sets needs in a slot: [sets][sed sets],
// [first], [first2], [sed_first], [error words], each 256-aligned.
size_t desc_bytes_for(int n) {
  return align256(2 * (size_t)n * sizeof(p3d::KParams)) + 2 * align256((size_t)(n + 1) * sizeof(int)) +
         align256((size_t)(n + p3d::kMaxLevelsPerLane) * sizeof(int)) + align256((size_t)n * sizeof(unsigned));
}

// Bytes the device arrays of a member span (every field of its column set).
size_t member_bytes(const p3d::KParams& k) {
  const size_t lev = level_span(k);
  const size_t iface = (size_t)(k.ncol - 1) * k.fcs + (size_t)k.nlev * k.fks + 1;
  // 10 state + 13 input + 7 per-level output + 3 exchange fields, 2 fluxes, 2 surface fields
  size_t b = (33 * lev + 2 * iface + 2 * (size_t)k.ncol) * sizeof(double);
  if (k.dbg_ice_idx) b += lev * sizeof(int);
  if (k.dbg_status) b += (size_t)k.ncol * sizeof(int);
  if (k.dbg_substeps) b += 3 * (size_t)k.dbg_stride * sizeof(int);
  return b;
}

// Enqueue the batch of slot s (srv->mu held).  Returns a P3X status.
int launch_slot(p3x_server* srv, Slot& sl) {
  const int n = (int)sl.reqs.size();
  std::vector<p3d::KParams> sets(n);
  // staging layout: [sets][sed sets][first][first2][sed_first][error words]
  const size_t o_sets = 0, o_sed = (size_t)n * sizeof(p3d::KParams);
  const size_t o_first = align256(2 * (size_t)n * sizeof(p3d::KParams));
  const size_t o_first2 = o_first + align256((size_t)(n + 1) * sizeof(int));
  const size_t o_sfirst = o_first2 + align256((size_t)(n + 1) * sizeof(int));
  const size_t o_err = o_sfirst + align256((size_t)(n + p3d::kMaxLevelsPerLane) * sizeof(int));
  unsigned* d_err = (unsigned*)(sl.d_desc + o_err);
  size_t off = 0;
  for (int i = 0; i < n; ++i) {
    const Request& r = sl.reqs[i];
    p3d::KParams k = srv->members[r.member].k;
    k.dt = r.dt;
    k.scratch = (double*)(sl.d_arena + off);
    k.colstat = (int*)(k.scratch + (size_t)p3d::SC_COUNT * k.sfs);
    k.err = d_err + i;  // one error word per set: a failing step is charged to its own request
    off += srv->members[r.member].scratch;
    sets[i] = k;
  }
  Layout lay;
  int st = layout_batch(sets.data(), n, srv->shape, &lay);
  if (st != P3X_OK) return st;
  char* h = sl.h_stage;
  std::memcpy(h + o_sets, sets.data(), n * sizeof(p3d::KParams));
  for (int i = 0; i < n; ++i)
    std::memcpy(h + o_sed + i * sizeof(p3d::KParams), &sets[lay.sed_order[i]], sizeof(p3d::KParams));
  std::memcpy(h + o_first, lay.first.data(), lay.first.size() * sizeof(int));
  std::memcpy(h + o_first2, lay.first2.data(), lay.first2.size() * sizeof(int));
  std::memcpy(h + o_sfirst, lay.sed_first.data(), lay.sed_first.size() * sizeof(int));
  std::memset(h + o_err, 0, (size_t)n * sizeof(unsigned));  // the error words start at zero
  const size_t bytes = o_err + (size_t)n * sizeof(unsigned);

  p3d::BatchPlan plan{};
  plan.shape = srv->shape;
  plan.sets = (const p3d::KParams*)(sl.d_desc + o_sets);
  plan.first = (const int*)(sl.d_desc + o_first);
  plan.first2 = (const int*)(sl.d_desc + o_first2);
  plan.nsets = n;
  // one block per virtual block: the pointwise kernels then run without their
  // persistent loop, which costs them registers
  plan.grid = grid_for(srv->ctx->ncu, -1, lay.first[n], srv->occ1);
  plan.grid2 = grid_for(srv->ctx->ncu, -1, lay.first2[n], srv->occ2);
  plan.total2 = lay.first2[n];
  plan.nclass = (int)lay.cls.size();
  for (int c = 0; c < plan.nclass; ++c) {
    const Layout::Cls& lc = lay.cls[c];
    p3d::SedClass& sc = plan.cls[c];
    sc.L = lc.L;
    sc.sets = (const p3d::KParams*)(sl.d_desc + o_sed) + lc.set_off;
    sc.first = (const int*)(sl.d_desc + o_sfirst) + lc.first_off;
    sc.nsets = lc.nsets;
    sc.grid = grid_for(srv->ctx->ncu, -1, lc.total, srv->occ_sed[lc.L]);
  }
  // the neighbour stages of process members before their steps ...
  for (int i = 0; i < n; ++i) {
    const p3x_server::Member& m = srv->members[sl.reqs[i].member];
    if (!m.has_chain) continue;
    st = chain_before(m.proc, &m.fields, &m.chain, sl.stream);
    if (st != P3X_OK) return st;
  }
  hipError_t e = hipMemcpyAsync(sl.d_desc, h, bytes, hipMemcpyHostToDevice, sl.stream);
  if (e == hipSuccess) e = launch_plan(srv->opt.precision, plan, sl.stream);
  // ... and after them
  for (int i = 0; i < n && e == hipSuccess; ++i) {
    const p3x_server::Member& m = srv->members[sl.reqs[i].member];
    if (!m.has_chain) continue;
    st = chain_after(m.proc, &m.fields, &m.chain, sl.reqs[i].dt, sl.stream);
    if (st != P3X_OK) return st;
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(sl.h_err, d_err, (size_t)n * sizeof(unsigned), hipMemcpyDeviceToHost, sl.stream);
  if (e == hipSuccess) e = hipEventRecord(sl.done, sl.stream);
  if (e == hipErrorInvalidValue) return P3X_ERR_ARG;
  return e == hipSuccess ? P3X_OK : P3X_ERR_DEVICE;
}

// Retire finished slots (srv->mu held).  Returns true if any finished.
bool retire(p3x_server* srv) {
  bool any = false;
  for (Slot& sl : srv->slots) {
    if (!sl.busy) continue;
    const hipError_t q = hipEventQuery(sl.done);
    if (q == hipErrorNotReady) continue;
    const double t = now_us();
    for (size_t i = 0; i < sl.reqs.size(); ++i) {
      const Request& r = sl.reqs[i];
      int st = P3X_OK;
      if (q != hipSuccess) st = P3X_ERR_DEVICE;
      else if (sl.h_err[i]) st = status_of_err(sl.h_err[i]);
      p3x_completion c{r.tag, r.member, st, t - r.t_submit};
      srv->done.push_back(c);
      srv->latencies.push_back(c.latency_us);
      srv->members[r.member].busy = false;
      const p3d::KParams& k = srv->members[r.member].k;
      srv->colev += (double)k.ncol * k.nlev;
      srv->completed++;
      srv->in_flight--;
    }
    srv->t_last = t;
    sl.reqs.clear();
    sl.busy = false;
    any = true;
  }
  return any;
}

// Fill free slots from the queue (srv->mu held).  Returns true if a batch
// failed to launch (its requests are then complete, with the error).
bool admit(p3x_server* srv) {
  bool failed = false;
  for (Slot& sl : srv->slots) {
    if (sl.busy || srv->pending.empty()) continue;
    size_t used = 0;
    std::vector<char> taken(srv->members.size(), 0);
    for (auto it = srv->pending.begin(); it != srv->pending.end() && (int)sl.reqs.size() < srv->opt.max_batch;) {
      // a member's requests run in submission order: once its earliest queued
      // request is passed over, its later ones wait too
      auto& m = srv->members[it->member];
      const bool take = !m.busy && !taken[it->member] && used + m.scratch <= sl.arena_bytes;
      taken[it->member] = 1;
      if (take) {
        used += m.scratch;
        sl.reqs.push_back(*it);
        it = srv->pending.erase(it);
      } else {
        ++it;
      }
    }
    if (sl.reqs.empty()) continue;
    const int st = launch_slot(srv, sl);
    if (st != P3X_OK) {  // the batch never ran: fail its requests
      const double t = now_us();
      for (const Request& r : sl.reqs) {
        srv->done.push_back(p3x_completion{r.tag, r.member, st, t - r.t_submit});
        srv->completed++;
      }
      sl.reqs.clear();
      if (st == P3X_ERR_DEVICE) srv->fatal = st;
      failed = true;
      continue;
    }
    for (const Request& r : sl.reqs) srv->members[r.member].busy = true;
    srv->in_flight += (long long)sl.reqs.size();
    sl.busy = true;
  }
  return failed;
}

void scheduler(p3x_server* srv) {
  (void)hipSetDevice(srv->ctx->device);
  std::unique_lock<std::mutex> lk(srv->mu);
  while (true) {
    const bool retired = retire(srv);
    const bool failed = admit(srv);
    if (retired || failed) srv->cv_done.notify_all();
    bool busy = false;
    for (const Slot& sl : srv->slots) busy |= sl.busy;
    if (srv->stop && !busy) break;
    if (busy) {
      // poll the in-flight events; new requests wake the thread early
      srv->cv_work.wait_for(lk, std::chrono::microseconds(20));
    } else {
      srv->cv_work.wait(lk, [&] { return srv->stop || !srv->pending.empty(); });
    }
  }
  srv->cv_done.notify_all();
}

void free_server(p3x_server* s) {
  for (Slot& sl : s->slots) {
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.done) (void)hipEventDestroy(sl.done);
    (void)hipFree(sl.d_arena);
    (void)hipFree(sl.d_desc);
    (void)hipHostFree(sl.h_stage);
    (void)hipHostFree(sl.h_err);
  }
  delete s;
}

}  // namespace

extern "C" void p3x_server_defaults(p3x_server_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->precision = P3X_MIXED;
  o->slots = 4;
  o->max_batch = 32;
  o->scratch_budget_bytes = 8.0 * (1 << 30);
  o->hbm_budget_bytes = 200e9;
}

extern "C" int p3x_server_create(p3x_context* c, const p3x_runtime* rt, const p3x_server_options* opts,
                                 p3x_server** out) {
  if (!c || !rt || !out) return P3X_ERR_ARG;
  *out = nullptr;
  p3x_server_options o;
  p3x_server_defaults(&o);
  if (opts) {
    o.precision = opts->precision;
    if (opts->slots > 0) o.slots = opts->slots;
    if (opts->max_batch > 0) o.max_batch = opts->max_batch;
    if (opts->scratch_budget_bytes > 0) o.scratch_budget_bytes = opts->scratch_budget_bytes;
    if (opts->hbm_budget_bytes > 0) o.hbm_budget_bytes = opts->hbm_budget_bytes;
    o.columns_per_block = opts->columns_per_block;
    o.threads_per_block = opts->threads_per_block;
    o.sed_lanes_per_column = opts->sed_lanes_per_column;
    o.part12_waves_per_simd = opts->part12_waves_per_simd;
  }
  if (o.precision != P3X_FP64 && o.precision != P3X_MIXED) return P3X_ERR_ARG;
  if (o.slots > 64 || o.max_batch > 4096) return P3X_ERR_ARG;
  if (o.scratch_budget_bytes > o.hbm_budget_bytes) return P3X_ERR_BUDGET;
  p3x_server* s = new p3x_server();
  s->ctx = c;
  s->rt = *rt;
  s->opt = o;
  // one shape for every batch; members are checked against its lanes per column
  const int st = resolve_shape(o.precision, 1, o.columns_per_block, o.threads_per_block,
                               o.sed_lanes_per_column != 0 ? o.sed_lanes_per_column : 32,
                               o.part12_waves_per_simd, &s->shape);
  if (st != P3X_OK) {
    delete s;
    return st;
  }
  hipError_t e = hipSuccess;
  for (int L = 1; L <= p3d::kMaxLevelsPerLane && e == hipSuccess; ++L) {
    if (L == 7) continue;
    int r1 = 1, r2 = 1, rsed = 1;
    e = occupancy(o.precision, s->shape, L, &r1, &r2, &rsed);
    s->occ1 = r1;
    s->occ2 = r2;
    s->occ_sed[L] = rsed;
  }
  s->occ_sed[7] = s->occ_sed[8];
  s->slots.resize(o.slots);
  const size_t arena = (size_t)(o.scratch_budget_bytes / o.slots) & ~(size_t)255;
  s->scratch_total = (double)arena * o.slots;
  for (Slot& sl : s->slots) {
    if (e != hipSuccess) break;
    sl.arena_bytes = arena;
    sl.desc_bytes = desc_bytes_for(o.max_batch);
    e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc((void**)&sl.d_arena, arena);
    if (e == hipSuccess) e = hipMalloc((void**)&sl.d_desc, sl.desc_bytes);
    if (e == hipSuccess) e = hipHostMalloc((void**)&sl.h_stage, sl.desc_bytes, hipHostMallocDefault);
    if (e == hipSuccess)
      e = hipHostMalloc((void**)&sl.h_err, (size_t)o.max_batch * sizeof(unsigned), hipHostMallocDefault);
  }
  if (e != hipSuccess) {
    free_server(s);
    return dev_fail(e);
  }
  *out = s;
  return P3X_OK;
}

extern "C" int p3x_server_add_member(p3x_server* s, const p3x_column_set* set, int* member) {
  if (!s || !set || !member) return P3X_ERR_ARG;
  if (set->nj <= 0 || set->nk <= 0) return P3X_ERR_ARG;
  if (p3d::sed_levels_per_lane(set->nk, s->shape.lpc) == 0) return P3X_ERR_ARG;
  if (!fields_present(&set->prognostic_state, &set->diagnostic_inputs, &set->diagnostic_outputs,
                      &set->history_only))
    return P3X_ERR_ARG;
  const size_t need = scratch_bytes_for(set->nj, set->nk);
  if (need > s->slots[0].arena_bytes) return P3X_ERR_BUDGET;  // its scratch can never fit a slot
  p3x_server::Member m;
  p3x_infrastructure inf = set->infrastructure;
  inf.dt = 1.0;  // replaced per request
  const int st = fill_params(s->ctx, &s->rt, &set->prognostic_state, &set->diagnostic_inputs,
                             &set->diagnostic_outputs, &inf, &set->history_only, set->nj, set->nk, set->layout,
                             set->ld, set->ld_iface, set->col_status, set->sed_substeps, set->ice_table_idx,
                             nullptr, &m.k);
  if (st != P3X_OK) return st;
  m.k.dbg_check = set->check_values;
  m.scratch = need;
  m.resident = member_bytes(m.k);
  m.busy = false;
  // registration may happen while the scheduler serves others: the members are
  // only touched under the lock
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->resident_bytes + (double)m.resident + s->scratch_total > s->opt.hbm_budget_bytes) return P3X_ERR_BUDGET;
  s->resident_bytes += (double)m.resident;
  s->members.push_back(m);
  *member = (int)s->members.size() - 1;
  return P3X_OK;
}

extern "C" int p3x_server_add_process(p3x_server* s, p3x_process* proc, const p3x_atm_fields* fields,
                                      const p3x_step_chain* chain, int* member) {
  if (!s || !proc || !fields || !member) return P3X_ERR_ARG;
  p3x_server::Member m;
  int st = process_kparams(proc, fields, 1.0, &m.k);  // dt is replaced per request
  if (st != P3X_OK) return st;
  if (process_context(proc) != s->ctx) return P3X_ERR_ARG;  // the batch shares one context's tables
  if (p3d::sed_levels_per_lane(m.k.nlev, s->shape.lpc) == 0) return P3X_ERR_ARG;
  st = chain_check(proc, fields, chain);
  if (st != P3X_OK) return st;
  const size_t need = scratch_bytes_for(m.k.ncol, m.k.nlev);
  if (need > s->slots[0].arena_bytes) return P3X_ERR_BUDGET;
  m.proc = proc;
  m.fields = *fields;
  if (chain) {
    m.chain = *chain;
    m.has_chain = true;
  }
  m.scratch = need;
  // the field-manager fields p3_main sees, the process's own arrays and the chain's
  const size_t lev = level_span(m.k);
  m.resident = member_bytes(m.k) + process_local_bytes(proc) +
               (chain ? (5 * lev + 6 * (size_t)m.k.ncol) * sizeof(double) : 0);
  m.busy = false;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->resident_bytes + (double)m.resident + s->scratch_total > s->opt.hbm_budget_bytes) return P3X_ERR_BUDGET;
  s->resident_bytes += (double)m.resident;
  s->members.push_back(m);
  *member = (int)s->members.size() - 1;
  return P3X_OK;
}

extern "C" int p3x_server_submit(p3x_server* s, int member, double dt, unsigned long long tag) {
  if (!s || !(dt > 0)) return P3X_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  if (member < 0 || member >= (int)s->members.size() || s->stop) return P3X_ERR_ARG;
  if (s->fatal != P3X_OK) return s->fatal;
  const double t = now_us();
  if (s->t_first < 0) s->t_first = t;
  s->pending.push_back(Request{tag, member, dt, t});
  s->submitted++;
  if (!s->started) {
    s->started = true;
    s->thr = std::thread(scheduler, s);
  }
  s->cv_work.notify_one();
  return P3X_OK;
}

extern "C" int p3x_server_poll(p3x_server* s, p3x_completion* out, int max, int wait_ms) {
  if (!s || (max > 0 && !out) || max < 0) return -P3X_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->done.empty() && wait_ms > 0 && s->completed < s->submitted)
    s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !s->done.empty(); });
  int n = 0;
  while (n < max && !s->done.empty()) {
    out[n++] = s->done.front();
    s->done.pop_front();
  }
  return n;
}

extern "C" int p3x_server_drain(p3x_server* s) {
  if (!s) return P3X_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->completed >= s->submitted; });
  return s->fatal;
}

extern "C" int p3x_server_stats_get(p3x_server* s, p3x_server_stats* st) {
  if (!s || !st) return P3X_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(st, 0, sizeof(*st));
  st->submitted = s->submitted;
  st->completed = s->completed;
  st->column_levels = s->colev;
  st->busy_s = s->t_first < 0 ? 0 : (s->t_last - s->t_first) * 1e-6;
  st->resident_bytes = s->resident_bytes;
  st->scratch_bytes = s->scratch_total;
  st->members = (int)s->members.size();
  if (!s->latencies.empty()) {
    std::vector<double> v = s->latencies;
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) {
      const size_t i = (size_t)std::min<double>((double)v.size() - 1, q * ((double)v.size() - 1) + 0.5);
      return v[i];
    };
    st->p50_us = pct(0.50);
    st->p99_us = pct(0.99);
    st->max_us = v.back();
  }
  return P3X_OK;
}

extern "C" void p3x_server_destroy(p3x_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->thr.joinable()) s->thr.join();
  (void)hipDeviceSynchronize();
  free_server(s);
}
