/*
 * server.hip -- served form: independent cases admitted while others run.
 *
 * No reference counterpart: the reference runs one case per process
 * (Main.C:96-141).  The request loop here:
 *
 *   submit -> queue -> setup thread: generate the mesh from the options
 *   (ma_mesh_generate), plan the device layout and compute the exact device
 *   bytes (ma::plan_case) -> admission in submission order, when a slot is
 *   free and the bytes fit in the server's region -> build inside the region
 *   on the case's own stream and set its initial state -> scheduler thread:
 *   the cases of one scheme form a group whose single set of launches
 *   advances all of them by one RK4 step (continuous batching; each
 *   workgroup finds its case in a device table); a case joins its group at
 *   the next step and leaves after its last one, its state is then copied to
 *   the host, its region block returned and the next request admitted.
 *
 * The region is one hipMalloc made at creation, so admitting and retiring
 * cases never allocates or frees device memory (hipFree can wait for the
 * whole device, which would stall the other cases).  Each case owns its
 * buffers and goes through the same per-cell and per-face arithmetic in the
 * same order as a ma_solver_create run of it alone, so its result is bit for
 * bit the same.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/miniaero.h"
#include "common.hpp"
#include "internal.hpp"

namespace {

using Clock = std::chrono::steady_clock;

// First-fit allocator of [0, size) in 2 MiB granules, coalescing on release.
class Region {
 public:
  static size_t granule(size_t n) {
    const size_t g = size_t(2) << 20;
    return (n + g - 1) / g * g;
  }
  void reset(size_t size) {
    free_.clear();
    if (size) free_[0] = size;
    size_ = size;
    used_ = 0;
  }
  bool reserve(size_t n, size_t* off) {
    n = granule(n);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      *off = it->first;
      const size_t rest = it->second - n, at = it->first + n;
      free_.erase(it);
      if (rest) free_[at] = rest;
      used_ += n;
      return true;
    }
    return false;
  }
  void release(size_t off, size_t n) {
    n = granule(n);
    used_ -= n;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t used() const { return used_; }
  size_t size() const { return size_; }

 private:
  std::map<size_t, size_t> free_;
  size_t size_ = 0, used_ = 0;
};

struct Request {
  long long id = 0;
  ma_options opt{};
  ma_status status = MA_OK;
  std::string error;
  double t_submit = 0.0, t_admit = 0.0, t_finish = 0.0;
  long long cells = 0;
  size_t bytes = 0, offset = 0;
  bool has_region = false, finished = false;
  ma_solver* solver = nullptr;
  int total = 0, done = 0;  // steps asked for, steps run
  int remaining = 0;        // steps not yet queued
  int parity = 0;           // state buffer of sol_n = parity ^ the group's step parity
  std::vector<double> U;
};

// Cases that can share launches: same order, viscous flag, limiter,
// arithmetic and Prandtl number.
bool same_scheme(const ma_options& a, const ma_options& b) {
  const double pa = a.prandtl > 0.0 ? a.prandtl : 0.71, pb = b.prandtl > 0.0 ? b.prandtl : 0.71;
  return a.second_order_space == b.second_order_space && a.viscous == b.viscous && a.limiter == b.limiter &&
         a.precision == b.precision && pa == pb;
}

// The cases of one scheme and the launches that advance them together.
struct Group {
  ma_options scheme{};
  hipStream_t stream = nullptr;
  ma::Multi* multi = nullptr;
  std::vector<Request*> act;  // in the launches of the next step
  bool dirty = true;          // act changed since the last upload
  int parity = 0;             // parity of the next step
  struct Step {
    hipEvent_t ev;
    std::vector<Request*> finishing;  // cases whose last step this is
  };
  std::deque<Step> inflight;
  std::vector<hipEvent_t> spare;
};

// Layouts of the meshes the server has planned, by what the generated mesh
// and its layout depend on: the box (problem type, lengths, ramp angle,
// cells per direction, y-stretch), the boundary kinds per side and the tile
// shape.  Requests of a shape seen before skip mesh generation and layout
// planning (about 2 us per cell on one core); least recently used layouts are
// dropped beyond the cache's host-byte budget.
struct PlanKey {
  int problem_type, nx, ny, nz, bc_kind[6], tile_shape[3];
  double lx, ly, lz, angle, y_stretch;
  bool operator<(const PlanKey& o) const { return std::memcmp(this, &o, sizeof(PlanKey)) < 0; }
};

PlanKey plan_key(const ma_options& o) {
  PlanKey k;
  std::memset(&k, 0, sizeof(k));  // padding compares equal
  k.problem_type = o.problem_type;
  k.nx = o.nx;
  k.ny = o.ny;
  k.nz = o.nz;
  for (int i = 0; i < 6; ++i) k.bc_kind[i] = o.bc_kind[i];
  for (int i = 0; i < 3; ++i) k.tile_shape[i] = o.tile_shape[i];
  k.lx = o.lx;
  k.ly = o.ly;
  k.lz = o.lz;
  k.angle = o.angle;
  k.y_stretch = o.y_stretch;
  return k;
}

class PlanCache {
 public:
  struct Entry {
    std::shared_ptr<const ma::LayoutPlan> lp;
    long long cells = 0;
    size_t bytes = 0;
  };
  void set_budget(size_t b) { budget_ = b; }
  bool get(const PlanKey& k, Entry* e) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = map_.find(k);
    if (it == map_.end()) {
      ++misses_;
      return false;
    }
    lru_.splice(lru_.begin(), lru_, it->second.second);  // most recent first
    *e = it->second.first;
    ++hits_;
    return true;
  }
  void put(const PlanKey& k, const Entry& e) {
    if (e.bytes > budget_) return;
    std::lock_guard<std::mutex> lk(mu_);
    if (map_.count(k)) return;  // planned meanwhile by another thread
    lru_.push_front(k);
    map_[k] = {e, lru_.begin()};
    used_ += e.bytes;
    while (used_ > budget_ && !lru_.empty()) {
      auto it = map_.find(lru_.back());
      used_ -= it->second.first.bytes;
      map_.erase(it);
      lru_.pop_back();
    }
  }
  void counts(long long* hits, long long* misses) {
    std::lock_guard<std::mutex> lk(mu_);
    *hits = hits_;
    *misses = misses_;
  }

 private:
  std::mutex mu_;
  std::list<PlanKey> lru_;
  std::map<PlanKey, std::pair<Entry, std::list<PlanKey>::iterator>> map_;
  size_t budget_ = 0, used_ = 0;
  long long hits_ = 0, misses_ = 0;
};

}  // namespace

struct ma_server {
  ma_server_config cfg{};
  PlanCache plans;
  int device = 0;
  char* base = nullptr;
  Region region;
  Clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_work;   // setup threads: a request was queued, or stop
  std::condition_variable cv_admit;  // admission turn, slot or region changed
  std::condition_variable cv_sched;  // scheduler: a case became active, or stop
  std::condition_variable cv_done;   // pollers: a case finished
  std::deque<Request*> queue;        // submitted, not yet taken by a setup thread
  std::deque<Request*> incoming;     // built, not yet taken by the scheduler
  std::deque<Request*> done_unpolled;
  std::deque<Request*> completing;   // last step done, state not yet copied back (finisher thread)
  std::condition_variable cv_complete;
  bool sched_done = false;           // the scheduler has returned: the finisher drains and stops
  std::map<long long, std::unique_ptr<Request>> all;
  long long next_id = 1;
  long long admit_next = 1;  // admission follows submission order
  long long waiting = 0;     // submitted, neither admitted nor finished
  int active = 0;            // admitted, not finished
  long long finished = 0;
  bool stop = false;
  std::vector<std::thread> setup;
  std::thread sched, finish;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

// Under the lock: the request is done (successfully or not).  A case that
// held a slot finished when it gave the slot back.
void mark_finished(ma_server* S, Request* r) {
  r->finished = true;
  if (r->t_finish == 0.0) r->t_finish = S->now();
  S->done_unpolled.push_back(r);
  ++S->finished;
  S->cv_done.notify_all();
}

void fail_request(ma_server* S, Request* r, ma_status st, const std::string& err) {
  std::lock_guard<std::mutex> lk(S->mu);
  r->status = st;
  r->error = err;
  mark_finished(S, r);
}

// Frees a built case's device resources and returns its region block.
void retire(ma_server* S, Request* r) {
  if (r->solver) ma_solver_destroy(r->solver);
  r->solver = nullptr;
  std::lock_guard<std::mutex> lk(S->mu);
  if (r->has_region) {
    S->region.release(r->offset, r->bytes);
    r->has_region = false;
    r->t_finish = S->now();
    --S->active;
    S->cv_admit.notify_all();
  }
}

void setup_worker(ma_server* S) {
  (void)hipSetDevice(S->device);
  for (;;) {
    Request* r = nullptr;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_work.wait(lk, [&] { return S->stop || !S->queue.empty(); });
      if (S->stop) return;  // queued requests are cancelled by ma_server_destroy
      r = S->queue.front();
      S->queue.pop_front();
    }
    // plan on this thread, outside the lock: the mesh's layout from the
    // cache, or generated and planned here, then the case's own plan on it
    ma::PlannedCase* pc = nullptr;
    size_t bytes = 0;
    std::string err;
    ma_status st = MA_OK;
    const PlanKey key = plan_key(r->opt);
    PlanCache::Entry ent;
    if (!S->plans.get(key, &ent)) {
      ma_mesh* mesh = nullptr;
      st = ma_mesh_generate(&r->opt, &mesh);
      if (!st) {
        ma_mesh_data md;
        st = ma_mesh_view(mesh, &md);
        if (!st) st = ma::plan_layout_shared(&md, &r->opt, &ent.lp, &ent.cells, &ent.bytes);
      }
      if (mesh) ma_mesh_destroy(mesh);  // the layout holds its own copies
      if (!st) S->plans.put(key, ent);
    }
    if (!st) {
      r->cells = ent.cells;
      st = ma::plan_case_with(ent.lp, &r->opt, &pc, &bytes);
    }
    if (st) err = ma::last_error();

    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_admit.wait(lk, [&] { return S->stop || S->admit_next == r->id; });
      if (!st && !S->stop && Region::granule(bytes) > S->region.size()) {
        st = MA_ERR_NOMEM;
        err = "case needs " + std::to_string(bytes) + " device bytes, more than the server's region";
      }
      if (!st)
        S->cv_admit.wait(lk, [&] {
          return S->stop || (S->active < S->cfg.max_active && S->region.reserve(bytes, &r->offset));
        });
      if (!st && S->stop) {
        st = MA_ERR_INVALID;
        err = "server stopped before the case was admitted";
      }
      if (!st) {
        r->has_region = true;
        r->bytes = bytes;
        r->t_admit = S->now();
        ++S->active;
      }
      --S->waiting;
      ++S->admit_next;
      S->cv_admit.notify_all();
    }
    if (st) {
      ma::destroy_plan(pc);
      fail_request(S, r, st, err);
      continue;
    }
    // build inside the region, outside the lock
    ma_solver* s = nullptr;
    st = ma::create_from_plan(pc, S->base + r->offset, r->bytes, &s);
    ma::destroy_plan(pc);
    r->solver = s;
    if (!st) st = ma_solver_initialize(s);
    if (st) {
      err = ma::last_error();
      retire(S, r);
      fail_request(S, r, st, err);
      continue;
    }
    r->total = r->opt.ntimesteps;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->incoming.push_back(r);
    }
    S->cv_sched.notify_one();
  }
}

// Copies the final state to the host, frees the case's device resources.
void complete(ma_server* S, Request* r) {
  if (r->status == MA_OK) {
    r->U.resize((size_t)r->cells * 5);
    ma_status st = ma_solver_get_state(r->solver, r->U.data());
    if (st) {
      r->status = st;
      r->error = ma::last_error();
      r->U.clear();
    } else {
      long long bad = 0;
      for (double x : r->U) bad += !std::isfinite(x);
      if (bad) {  // the state is kept for inspection
        r->status = MA_ERR_NONFINITE;
        r->error = std::to_string(bad) + " non-finite values in the final state";
      }
    }
  }
  retire(S, r);
  std::lock_guard<std::mutex> lk(S->mu);
  mark_finished(S, r);
}

// The scheduler's side of a finished case: the finisher thread copies the
// state back and frees the case, so that the scheduler goes on queueing the
// other cases' steps meanwhile.  The scheduler does not touch r afterwards.
void hand_off(ma_server* S, Request* r) {
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->completing.push_back(r);
  }
  S->cv_complete.notify_one();
}

void finisher(ma_server* S) {
  (void)hipSetDevice(S->device);
  for (;;) {
    Request* r = nullptr;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_complete.wait(lk, [&] { return S->sched_done || !S->completing.empty(); });
      if (S->completing.empty()) return;  // the scheduler is gone and nothing is left
      r = S->completing.front();
      S->completing.pop_front();
    }
    complete(S, r);
  }
}

void fail_group(ma_server* S, Group& g, const std::string& why) {
  (void)hipStreamSynchronize(g.stream);
  std::vector<Request*> all(g.act);
  for (auto& st : g.inflight) {
    all.insert(all.end(), st.finishing.begin(), st.finishing.end());
    g.spare.push_back(st.ev);
  }
  g.inflight.clear();
  g.act.clear();
  g.dirty = true;
  for (Request* r : all) {
    r->status = MA_ERR_DEVICE;
    r->error = why;
    hand_off(S, r);
  }
}

// The request loop's device side: every group queues one RK step of all its
// cases at a time (at most steps_in_flight ahead), cases join a group at its
// next step and leave after their last one.
void scheduler(ma_server* S) {
  (void)hipSetDevice(S->device);
  std::vector<std::unique_ptr<Group>> groups;
  auto idle = [&] {
    for (auto& g : groups)
      if (!g->act.empty() || !g->inflight.empty()) return false;
    return true;
  };
  for (;;) {
    std::vector<Request*> joining;
    bool stopping = false;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      if (idle()) S->cv_sched.wait(lk, [&] { return S->stop || !S->incoming.empty(); });
      joining.assign(S->incoming.begin(), S->incoming.end());
      S->incoming.clear();
      stopping = S->stop;
    }
    if (stopping) {
      for (auto& g : groups) {
        (void)hipStreamSynchronize(g->stream);
        std::vector<Request*> all(g->act);
        for (auto& st : g->inflight) all.insert(all.end(), st.finishing.begin(), st.finishing.end());
        all.insert(all.end(), joining.begin(), joining.end());
        joining.clear();
        for (Request* r : all) {
          r->status = MA_ERR_INVALID;
          r->error = "server stopped before the case finished";
          retire(S, r);
          std::lock_guard<std::mutex> lk(S->mu);
          mark_finished(S, r);
        }
      }
      for (Request* r : joining) {
        r->status = MA_ERR_INVALID;
        r->error = "server stopped before the case finished";
        retire(S, r);
        std::lock_guard<std::mutex> lk(S->mu);
        mark_finished(S, r);
      }
      for (auto& g : groups) {
        for (auto& st : g->inflight) (void)hipEventDestroy(st.ev);
        for (hipEvent_t e : g->spare) (void)hipEventDestroy(e);
        ma::multi_destroy(g->multi);
        (void)hipStreamDestroy(g->stream);
      }
      return;
    }
    bool progress = false;
    for (Request* r : joining) {
      progress = true;
      if (r->total == 0) {
        hand_off(S, r);  // the initial state is the result
        continue;
      }
      Group* g = nullptr;
      for (auto& x : groups)
        if (same_scheme(x->scheme, r->opt)) g = x.get();
      if (!g) {
        std::unique_ptr<Group> ng(new Group());
        ng->scheme = r->opt;
        ma_status st = MA_OK;
        if (hipStreamCreateWithFlags(&ng->stream, hipStreamNonBlocking) != hipSuccess)
          st = ma::fail(MA_ERR_DEVICE, "hipStreamCreate failed");
        if (!st) st = ma::multi_create(r->opt, S->cfg.max_active, ng->stream, &ng->multi);
        if (st) {
          if (ng->stream) (void)hipStreamDestroy(ng->stream);
          r->status = st;
          r->error = ma::last_error();
          hand_off(S, r);
          continue;
        }
        groups.push_back(std::move(ng));
        g = groups.back().get();
      }
      r->parity = ma::solver_cur(r->solver) ^ g->parity;
      r->remaining = r->total;
      g->act.push_back(r);
      g->dirty = true;
    }
    for (auto& gp : groups) {
      Group& g = *gp;
      while (!g.inflight.empty()) {  // finished steps
        const hipError_t q = hipEventQuery(g.inflight.front().ev);
        if (q == hipErrorNotReady) break;
        if (q != hipSuccess) {
          fail_group(S, g, std::string("group stream: ") + hipGetErrorString(q));
          break;
        }
        Group::Step st = g.inflight.front();
        g.inflight.pop_front();
        g.spare.push_back(st.ev);
        for (Request* r : st.finishing) {
          ma::solver_set_progress(r->solver, r->total);
          r->done = r->total;
          hand_off(S, r);  // last use of r here
        }
        progress = true;
      }
      while (!g.act.empty() && (int)g.inflight.size() < S->cfg.steps_in_flight) {  // queue steps
        ma_status st = MA_OK;
        if (g.dirty) {
          std::vector<ma_solver*> sv;
          std::vector<int> par;
          for (Request* r : g.act) {
            sv.push_back(r->solver);
            par.push_back(r->parity);
          }
          st = ma::multi_set(g.multi, sv.data(), par.data(), (int)sv.size());
          g.dirty = false;
        }
        if (!st) st = ma::multi_step(g.multi, g.parity);
        hipEvent_t ev = nullptr;
        if (!st) {
          if (!g.spare.empty()) {
            ev = g.spare.back();
            g.spare.pop_back();
          } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            st = ma::fail(MA_ERR_DEVICE, "hipEventCreate failed");
          }
        }
        if (!st && hipEventRecord(ev, g.stream) != hipSuccess) st = ma::fail(MA_ERR_DEVICE, "hipEventRecord failed");
        if (st) {
          if (ev) g.spare.push_back(ev);
          fail_group(S, g, ma::last_error());
          break;
        }
        Group::Step step{ev, {}};
        std::vector<Request*> keep;
        for (Request* r : g.act) {
          if (--r->remaining == 0) step.finishing.push_back(r);
          else keep.push_back(r);
        }
        if (keep.size() != g.act.size()) {
          g.act.swap(keep);
          g.dirty = true;
        }
        g.inflight.push_back(std::move(step));
        g.parity ^= 1;
        progress = true;
      }
    }
    if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(50));
  }
}

}  // namespace

// The most cases the server keeps on the device at once.  A stream with no
// slot limit (about 1,400 cases admitted by free HBM alone) stopped completing
// cases without a device error, for a cause not found; the largest limit run
// since is 64.
constexpr int kMaxActive = 64;

extern "C" ma_status ma_server_create(const ma_server_config* cfg, ma_server** out) {
  if (!out) return ma::fail(MA_ERR_INVALID, "ma_server_create: null argument");
  *out = nullptr;
  ma_server_config c = cfg ? *cfg : ma_server_config{};
  if (c.hbm_budget_bytes < 0 || c.max_active < 0 || c.setup_threads < 0 || c.steps_in_flight < 0)
    return ma::fail(MA_ERR_INVALID, "ma_server_create: negative setting");
  if (c.hbm_budget_bytes == 0) c.hbm_budget_bytes = 200000000000ll;
  if (c.max_active == 0) c.max_active = 16;
  if (c.max_active > kMaxActive)
    return ma::fail(MA_ERR_INVALID, "ma_server_create: max_active above " + std::to_string(kMaxActive) +
                                        " is not supported");
  if (c.setup_threads == 0) c.setup_threads = 4;
  if (c.steps_in_flight == 0) c.steps_in_flight = 2;
  if (c.plan_cache_bytes == 0) c.plan_cache_bytes = 16ll << 30;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return ma::fail(MA_ERR_DEVICE, "no HIP device available");
  size_t freeb = 0, totalb = 0;
  if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) return ma::fail(MA_ERR_DEVICE, "hipMemGetInfo failed");
  // leave room for the runtime, streams and graphs of the active cases
  const size_t margin = std::max<size_t>(size_t(2) << 30, freeb / 50);
  size_t size = std::min<size_t>((size_t)c.hbm_budget_bytes, freeb > margin ? freeb - margin : 0);
  size = size / (size_t(2) << 20) * (size_t(2) << 20);
  if (size < (size_t(64) << 20)) return ma::fail(MA_ERR_NOMEM, "ma_server_create: less than 64 MiB of HBM free");
  ma_server* S = new (std::nothrow) ma_server();
  if (!S) return ma::fail(MA_ERR_NOMEM, "ma_server_create: out of host memory");
  S->cfg = c;
  S->plans.set_budget(c.plan_cache_bytes > 0 ? (size_t)c.plan_cache_bytes : 0);
  (void)hipGetDevice(&S->device);
  if (hipMalloc((void**)&S->base, size) != hipSuccess) {
    delete S;
    return ma::fail(MA_ERR_NOMEM, "ma_server_create: hipMalloc of the case region failed");
  }
  S->region.reset(size);
  S->t0 = Clock::now();
  try {
    for (int i = 0; i < c.setup_threads; ++i) S->setup.emplace_back(setup_worker, S);
    S->sched = std::thread(scheduler, S);
    S->finish = std::thread(finisher, S);
  } catch (const std::exception&) {
    ma_server_destroy(S);
    return ma::fail(MA_ERR_DEVICE, "ma_server_create: could not start threads");
  }
  *out = S;
  return MA_OK;
}

extern "C" ma_status ma_server_submit(ma_server* S, const ma_options* opts, long long* id) {
  if (!S || !opts || !id) return ma::fail(MA_ERR_INVALID, "ma_server_submit: null argument");
  if (opts->ntimesteps < 0) return ma::fail(MA_ERR_INVALID, "ma_server_submit: negative ntimesteps");
  std::unique_ptr<Request> r(new (std::nothrow) Request());
  if (!r) return ma::fail(MA_ERR_NOMEM, "ma_server_submit: out of host memory");
  r->opt = *opts;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) return ma::fail(MA_ERR_INVALID, "ma_server_submit: server is stopping");
    r->id = S->next_id++;
    r->t_submit = S->now();
    *id = r->id;
    S->queue.push_back(r.get());
    ++S->waiting;
    S->all[r->id] = std::move(r);
  }
  S->cv_work.notify_one();
  return MA_OK;
}

extern "C" int ma_server_poll(ma_server* S, ma_case_result* out, int max, double timeout_s) {
  if (!S || (!out && max > 0) || max < 0) {
    ma::fail(MA_ERR_INVALID, "ma_server_poll: bad argument");
    return -1;
  }
  std::unique_lock<std::mutex> lk(S->mu);
  if (S->done_unpolled.empty() && timeout_s > 0.0)
    S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s),
                        [&] { return !S->done_unpolled.empty() || S->stop; });
  int n = 0;
  while (n < max && !S->done_unpolled.empty()) {
    Request* r = S->done_unpolled.front();
    S->done_unpolled.pop_front();
    ma_case_result& o = out[n++];
    o.id = r->id;
    o.status = r->status;
    o.steps = r->done;
    o.cells = r->cells;
    o.device_bytes = (long long)r->bytes;
    o.t_submit = r->t_submit;
    o.t_admit = r->t_admit;
    o.t_finish = r->t_finish;
  }
  return n;
}

extern "C" ma_status ma_server_get_state(ma_server* S, long long id, double* U) {
  if (!S || !U) return ma::fail(MA_ERR_INVALID, "ma_server_get_state: null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->all.find(id);
  if (it == S->all.end()) return ma::fail(MA_ERR_INVALID, "ma_server_get_state: unknown or released id");
  Request* r = it->second.get();
  if (!r->finished) return ma::fail(MA_ERR_INVALID, "ma_server_get_state: case not finished");
  if (r->U.empty() || (r->status != MA_OK && r->status != MA_ERR_NONFINITE))
    return ma::fail(r->status ? (ma_status)r->status : MA_ERR_INVALID, "case failed: " + r->error);
  std::memcpy(U, r->U.data(), r->U.size() * sizeof(double));
  if (r->status == MA_ERR_NONFINITE) return ma::fail(MA_ERR_NONFINITE, r->error);  // U is filled
  return MA_OK;
}

extern "C" ma_status ma_server_release(ma_server* S, long long id) {
  if (!S) return ma::fail(MA_ERR_INVALID, "ma_server_release: null server");
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->all.find(id);
  if (it == S->all.end()) return ma::fail(MA_ERR_INVALID, "ma_server_release: unknown or released id");
  if (!it->second->finished) return ma::fail(MA_ERR_INVALID, "ma_server_release: case not finished");
  S->done_unpolled.erase(std::remove(S->done_unpolled.begin(), S->done_unpolled.end(), it->second.get()),
                         S->done_unpolled.end());
  S->all.erase(it);
  return MA_OK;
}

extern "C" ma_status ma_server_stats(ma_server* S, long long* stats) {
  if (!S || !stats) return ma::fail(MA_ERR_INVALID, "ma_server_stats: null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  stats[0] = S->next_id - 1;
  stats[1] = S->waiting;
  stats[2] = S->active;
  stats[3] = S->finished;
  stats[4] = (long long)S->region.used();
  stats[5] = (long long)S->region.size();
  return MA_OK;
}

extern "C" ma_status ma_server_plan_cache(ma_server* S, long long* hits, long long* misses) {
  if (!S || !hits || !misses) return ma::fail(MA_ERR_INVALID, "ma_server_plan_cache: null argument");
  S->plans.counts(hits, misses);
  return MA_OK;
}

extern "C" double ma_server_now(ma_server* S) { return S ? S->now() : 0.0; }

extern "C" void ma_server_destroy(ma_server* S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  S->cv_admit.notify_all();
  S->cv_sched.notify_all();
  S->cv_done.notify_all();
  for (auto& t : S->setup)
    if (t.joinable()) t.join();
  if (S->sched.joinable()) S->sched.join();
  {  // the finisher completes what the scheduler handed it, then stops
    std::lock_guard<std::mutex> lk(S->mu);
    S->sched_done = true;
  }
  S->cv_complete.notify_all();
  if (S->finish.joinable()) S->finish.join();
  // built after the scheduler stopped
  for (Request* r : S->incoming) retire(S, r);
  S->incoming.clear();
  if (S->base) (void)hipFree(S->base);
  delete S;
}
