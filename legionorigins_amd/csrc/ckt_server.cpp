// ckt_server.cpp -- continuous batching of circuit simulations.
//
// A request is a circuit (generator parameters) and a loop count.  A
// scheduler thread runs the serving loop; requests reach it from
// ckt_server_submit, from any host thread at any time, or from a trace that
// ckt_server_run_trace replays in real time on top of the same loop:
//   arrival   -> FIFO queue
//   admission -> when the HBM budget, the resident-circuit cap and the
//                per-iteration wire cap allow, the request's device footprint
//                is reserved and a host thread walks its generator chain
//                (each wire's first drand48 state, 8 bytes per wire)
//   generate  -> the device generator is enqueued behind the running
//                iteration (the GPU draws the circuit in HBM; nothing is
//                uploaded but the start states); the circuit joins the next
//                iteration
//   iteration -> one loop (calc_new_currents, distribute_charge,
//                update_voltages) of every resident circuit with loops left,
//                each task one launch over all of them
//   retire    -> a circuit whose loops are done is checksummed (its final
//                voltages are copied behind its last iteration), optionally
//                kept, and its slab released.
// Each pass launches the next iteration first and then retires, takes
// arrivals and cancellations and admits while it runs, so the GPU waits only
// for the launch.  A cancelled request leaves the queue at once, or its
// circuit leaves at the next iteration boundary (it is not in any iteration
// launched after the cancel is seen).  Circuits are independent, so a
// circuit's result does not depend on what else shares its iterations
// (bit-identical with EXACT + CSR).
#include <time.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ckt_engine.h"

using namespace ckt;

namespace {

enum RecState { kQueued = 0, kPlanning = 1, kResident = 2, kDone = 3 };

// One submitted request.  gen, loops, timeout_s, need and wires are fixed at
// submission; the rest belongs to the scheduler thread and is read by other
// threads under the server's mutex once state is kDone.
struct Rec {
  ckt_gen_params_t gen{};
  int32_t loops = 0;
  double timeout_s = 0;
  uint64_t need = 0;     // slab bytes
  uint32_t wires = 0;
  int state = kQueued;
  bool cancel = false;   // set by the scheduler when it applies a cancel
  uint32_t circuit = 0;  // engine circuit id while resident
  ckt_request_result_t res{};
};

struct KeptState {
  std::vector<ckt_node_t> nodes;
  std::vector<ckt_wire_t> wires;
};

}  // namespace

struct ckt_server {
  Engine eng;
  ckt_server_config_t cfg{};
  bool keep_state = false;
  uint32_t keep_max_wires = UINT32_MAX;  // keep only requests with at most this many wires
  std::unordered_map<uint32_t, KeptState> kept;  // by request id (trace index after a trace)

  // Shared with the scheduler thread, under mu.
  std::mutex mu;
  std::condition_variable wake;      // scheduler: submissions, cancels, stop, finished plans
  std::condition_variable finished;  // ckt_server_wait
  uint64_t wake_seq = 0;
  std::deque<Rec> recs;              // by request id; references stay valid as it grows
  std::vector<uint32_t> incoming;    // submitted, not yet taken by the scheduler
  std::vector<uint32_t> cancels;     // cancel requests not yet applied
  bool running = false, stop = false, drain = true;
  int sched_status = CKT_OK;
  std::string sched_error;
  std::thread thread;
  double t0 = 0;
  // Counters of the current run (scheduler thread; read after it stops).
  double wire_updates = 0;
  int32_t iterations = 0, max_resident = 0;
  double last_finish = 0;

  void poke() {
    std::lock_guard<std::mutex> g(mu);
    wake_seq++;
    wake.notify_all();
  }
};

namespace {

double now_s() {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

struct Ready {
  uint32_t id;
  std::unique_ptr<DevgenPlan> plan;
  int status;
  std::string error;
};

// Host threads that walk admitted requests' generator chains; each finished
// plan wakes the scheduler.
class GenPool {
 public:
  GenPool(int threads, ckt_server* s) : s_(s) {
    for (int i = 0; i < threads; i++) th_.emplace_back([this] { work(); });
  }
  ~GenPool() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  void submit(uint32_t id, const ckt_gen_params_t& gen) {
    {
      std::lock_guard<std::mutex> g(mu_);
      jobs_.emplace_back(id, gen);
    }
    cv_.notify_one();
  }
  void take(std::vector<Ready>& out) {
    std::lock_guard<std::mutex> g(mu_);
    for (auto& r : done_) out.push_back(std::move(r));
    done_.clear();
  }

 private:
  void work() {
    for (;;) {
      std::pair<uint32_t, ckt_gen_params_t> job;
      {
        std::unique_lock<std::mutex> g(mu_);
        cv_.wait(g, [this] { return stop_ || !jobs_.empty(); });
        if (stop_) return;
        job = jobs_.front();
        jobs_.pop_front();
      }
      Ready r;
      r.id = job.first;
      r.plan.reset(new DevgenPlan());
      r.status = devgen_plan(&job.second, *r.plan);
      if (r.status) {
        r.error = ckt_last_error();
        r.plan.reset();
      }
      {
        std::lock_guard<std::mutex> g(mu_);
        done_.push_back(std::move(r));
      }
      s_->poke();
    }
  }
  ckt_server* s_;
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<std::pair<uint32_t, ckt_gen_params_t>> jobs_;
  std::vector<Ready> done_;
  bool stop_ = false;
};

// Grow-only pinned host buffer (device-to-host copies into it are
// asynchronous); grown only when no copy into it is pending.
struct PinnedBuffer {
  void* p = nullptr;
  size_t cap = 0;
  ~PinnedBuffer() {
    if (p) (void)hipHostFree(p);
  }
  int reserve(size_t bytes) {
    if (bytes <= cap) return CKT_OK;
    if (p) CKT_HIP(hipHostFree(p));
    p = nullptr;
    cap = 0;
    CKT_HIP(hipHostMalloc(&p, bytes, hipHostMallocDefault));
    cap = bytes;
    return CKT_OK;
  }
  template <class T>
  T* at(size_t elem_offset) {
    return reinterpret_cast<T*>(p) + elem_offset;
  }
};

double percentile(std::vector<double> v, double p) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = p * (double)(v.size() - 1);
  const size_t lo = (size_t)pos;
  const size_t hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (pos - (double)lo) * (v[hi] - v[lo]);
}

// Sizes of a request, or why it can never run under this budget.
int validate(const ckt_gen_params_t& gen, int32_t loops, uint64_t budget, uint64_t* need,
             uint32_t* wires) {
  uint32_t N = 0, W = 0, P = 0;
  if (int st = gen_validate(&gen, &N, &W, &P)) return st;
  if (loops < 0) return fail(CKT_ERR_INVALID_ARG, "server: negative loop count");
  *need = SlabLayout::make(N, W, P).total;
  *wires = W;
  if (*need > budget)
    return fail(CKT_ERR_BUDGET, "server: request needs " + std::to_string(*need) +
                                    " bytes, more than the HBM budget");
  return CKT_OK;
}

// Records a submitted request (under s->mu) and wakes the scheduler.
uint32_t enqueue(ckt_server* s, const ckt_gen_params_t& gen, int32_t loops, double timeout_s,
                 double arrival_s, uint64_t need, uint32_t wires) {
  const uint32_t id = (uint32_t)s->recs.size();
  s->recs.emplace_back();
  Rec& r = s->recs.back();
  r.gen = gen;
  r.loops = loops;
  r.timeout_s = timeout_s;
  r.need = need;
  r.wires = wires;
  r.res.status = CKT_ERR_STATE;
  r.res.num_wires = wires;
  r.res.num_loops = loops;
  r.res.arrival_s = arrival_s;
  s->incoming.push_back(id);
  s->wake_seq++;
  s->wake.notify_all();
  return id;
}

// The serving loop (scheduler thread).
void schedule(ckt_server* s) {
  Engine& E = s->eng;
  const ckt_server_config_t cfg = s->cfg;
  int st = CKT_OK;
  if (hipSetDevice(E.device()) != hipSuccess) st = fail(CKT_ERR_HIP, "server: hipSetDevice");
  GenPool pool(cfg.gen_threads, s);
  std::deque<uint32_t> queue;       // arrived, not admitted
  uint64_t reserved = 0;            // bytes of admitted, not retired requests
  uint64_t admitted_wires = 0;      // wires of admitted, not retired requests
  int32_t admitted = 0;             // admitted, not retired requests
  std::vector<uint32_t> request_of_circuit;
  std::vector<Ready> ready;
  std::vector<uint32_t> in_flight;  // requests advanced by the running iteration
  // Circuits finishing with an iteration, as (circuit id, offset in that
  // iteration's download buffer).  Two buffers alternate: the voltages of the
  // iteration that just ended are summed while the next one, launched first,
  // copies its own finishers' voltages into the other buffer.
  std::vector<std::pair<uint32_t, size_t>> finishing[2];
  PinnedBuffer downloads[2];
  int cur = 0;
  auto clock = [&] { return now_s() - s->t0; };
  // The scheduler's own view of the records: pointers taken under the lock
  // when a request is first seen (elements of the deque never move; the
  // deque's index itself may change under a concurrent submit).
  std::vector<Rec*> recp;
  auto rec = [&](uint32_t id) -> Rec& { return *recp[id]; };

  auto retire = [&](uint32_t id, int status, double checksum) {
    std::lock_guard<std::mutex> g(s->mu);
    Rec& r = rec(id);
    r.res.status = status;
    r.res.finish_s = clock();
    r.res.final_voltage_checksum = checksum;
    r.state = kDone;
    s->last_finish = std::max(s->last_finish, r.res.finish_s);
    s->finished.notify_all();
  };
  auto release = [&](uint32_t id) {
    reserved -= rec(id).need;
    admitted_wires -= rec(id).wires;
    admitted--;
  };

  while (!st) {
    // Inputs: submissions, cancels, stop.
    std::vector<uint32_t> arrivals, cancels;
    bool stop, drain;
    uint64_t seen;
    {
      std::lock_guard<std::mutex> g(s->mu);
      arrivals.swap(s->incoming);
      cancels.swap(s->cancels);
      stop = s->stop;
      drain = s->drain;
      seen = s->wake_seq;
      for (uint32_t id : arrivals) {
        if (recp.size() <= id) recp.resize(id + 1, nullptr);
        recp[id] = &s->recs[id];
      }
    }
    for (uint32_t id : arrivals) queue.push_back(id);
    if (stop && !drain) {
      for (uint32_t id : queue) cancels.push_back(id);
      for (Circuit* c : E.slots()) cancels.push_back(request_of_circuit[c->id]);
      for (uint32_t id = 0; id < recp.size(); id++)
        if (recp[id] && rec(id).state == kPlanning) cancels.push_back(id);
    }
    for (uint32_t id : cancels) {
      Rec& r = rec(id);
      if (r.state == kQueued) {
        auto q = std::find(queue.begin(), queue.end(), id);
        if (q != queue.end()) {
          queue.erase(q);
          retire(id, CKT_ERR_CANCELLED, 0);
        }
      } else if (r.state == kPlanning) {
        r.cancel = true;  // retired when its plan comes back
      } else if (r.state == kResident && !r.cancel) {
        // Leaves before the next iteration, unless its last one has run.
        Circuit* c = E.find(r.circuit);
        if (c && c->loops_left > 0) {
          r.cancel = true;
          c->loops_left = 0;
        }
      }
    }

    // Launch one iteration of every resident circuit with loops left.  It is
    // the first GPU work after the previous iteration ended, so that the
    // retirement, arrivals and admission below run while it does.
    std::vector<PieceRef> refs;
    in_flight.clear();
    finishing[cur].clear();
    for (uint32_t slot = 0; slot < E.slots().size(); slot++) {
      Circuit* c = E.slots()[slot];
      if (c->loops_left <= 0) continue;
      in_flight.push_back(request_of_circuit[c->id]);
      for (uint32_t k = 0; k < c->num_pieces; k++) refs.push_back(PieceRef{slot, k});
    }
    if (!refs.empty()) {
      if ((st = E.loop_once(refs, cfg.policy, cfg.reduce_mode))) break;
      s->iterations++;
      // Circuits that finish with this iteration: their final voltages are
      // copied to pinned memory behind it, for the checksum after it ends.
      size_t total = 0;
      for (uint32_t ri : in_flight) {
        const Circuit* c = E.find(rec(ri).circuit);
        if (c->loops_left == 1) {
          finishing[cur].emplace_back(c->id, total);
          total += c->num_nodes;
        }
      }
      if ((st = downloads[cur].reserve(total * sizeof(float)))) break;
      for (auto& f : finishing[cur])
        if ((st = E.copy_voltages_async(f.first, downloads[cur].at<float>(f.second)))) break;
      if (st) break;
    }
    // Retire circuits with no loops left: the previous iteration's finishers,
    // whose voltages are in the other buffer, zero-loop requests and
    // cancelled ones.  None of them is in the iteration just launched.
    std::vector<uint32_t> done;
    for (Circuit* c : E.slots())
      if (c->loops_left <= 0) done.push_back(c->id);
    const auto& prev = finishing[cur ^ 1];
    std::vector<std::pair<uint32_t, double>> sums;
    for (uint32_t id : done) {
      const uint32_t ri = request_of_circuit[id];
      if (rec(ri).cancel) continue;
      double sum = 0;
      auto f = std::find_if(prev.begin(), prev.end(),
                            [id](const std::pair<uint32_t, size_t>& x) { return x.first == id; });
      if (f != prev.end()) {
        const float* v = downloads[cur ^ 1].at<float>(f->second);
        for (uint32_t k = 0, nn = E.find(id)->num_nodes; k < nn; k++) sum += v[k];
      } else {  // never ran (zero loops)
        std::vector<float> v;
        if ((st = E.download_voltages(id, v))) break;
        for (float x : v) sum += x;
      }
      sums.emplace_back(ri, sum);
      if (s->keep_state && rec(ri).wires <= s->keep_max_wires) {
        Circuit* c = E.find(id);
        KeptState k;
        k.nodes.resize(c->num_nodes);
        k.wires.resize(c->num_wires);
        if ((st = E.store(id, k.nodes.data(), k.wires.data()))) break;
        std::lock_guard<std::mutex> g(s->mu);
        s->kept[ri] = std::move(k);
      }
    }
    if (st) break;
    // One removal for the whole set.  Slabs from the stream's memory pool are
    // released in stream order, behind the running iteration, which does not
    // use them; the slot table is rewritten by a copy queued behind it too.
    if ((st = E.remove_many(done.data(), done.size()))) break;
    for (uint32_t id : done) {
      const uint32_t ri = request_of_circuit[id];
      release(ri);
      if (rec(ri).cancel) retire(ri, CKT_ERR_CANCELLED, 0);
    }
    for (auto& x : sums) retire(x.first, CKT_OK, x.second);
    const double now = clock();
    // Load shedding: queued requests past their admission deadline leave.
    double next_deadline = 1e300;
    if (!queue.empty()) {
      std::deque<uint32_t> keep;
      for (uint32_t i : queue) {
        const Rec& r = rec(i);
        if (r.timeout_s > 0 && now - r.res.arrival_s > r.timeout_s) {
          retire(i, CKT_ERR_TIMEOUT, 0);
        } else {
          keep.push_back(i);
          if (r.timeout_s > 0) next_deadline = std::min(next_deadline, r.res.arrival_s + r.timeout_s);
        }
      }
      queue.swap(keep);
    }
    // Admission, first come first served.
    while (!queue.empty()) {
      const uint32_t i = queue.front();
      const Rec& r = rec(i);
      const bool fits_mem = reserved + r.need <= cfg.hbm_budget_bytes;
      const bool fits_slots = cfg.max_active <= 0 || admitted < cfg.max_active;
      const bool fits_wires = cfg.max_active_wires == 0 || admitted == 0 ||
                              admitted_wires + r.wires <= cfg.max_active_wires;
      if (!(fits_mem && fits_slots && fits_wires)) break;
      queue.pop_front();
      reserved += r.need;
      admitted_wires += r.wires;
      admitted++;
      {
        std::lock_guard<std::mutex> g(s->mu);
        rec(i).state = kPlanning;
      }
      pool.submit(i, r.gen);
    }
    // With nothing on the GPU, sleep until something changes: a submission,
    // a cancel, a stop, a finished plan or the next admission deadline.
    if (refs.empty()) {
      std::unique_lock<std::mutex> g(s->mu);
      const bool idle_done = s->stop && queue.empty() && admitted == 0 && E.count() == 0 &&
                             s->incoming.empty();
      if (idle_done) break;
      const double wait = std::max(0.0, std::min(0.05, next_deadline - clock()));
      s->wake.wait_for(g, std::chrono::duration<double>(wait),
                       [&] { return s->wake_seq != seen; });
    }
    // Overlap: enqueue the generation of planned circuits behind the
    // running iteration.
    ready.clear();
    pool.take(ready);
    for (Ready& r : ready) {
      Rec& q = rec(r.id);
      if (q.cancel || r.status) {
        release(r.id);
        retire(r.id, q.cancel ? CKT_ERR_CANCELLED : r.status, 0);
        continue;
      }
      uint32_t id = 0;
      int as = E.add_planned(*r.plan, q.loops, &id);
      r.plan.reset();
      if (as) {
        release(r.id);
        retire(r.id, as, 0);
        continue;
      }
      if (request_of_circuit.size() <= id) request_of_circuit.resize(id + 1, 0);
      request_of_circuit[id] = r.id;
      q.circuit = id;
      std::lock_guard<std::mutex> g(s->mu);
      q.state = kResident;
      q.res.admit_s = clock();
    }
    s->max_resident = std::max<int32_t>(s->max_resident, (int32_t)E.count());
    if (!refs.empty()) {
      if ((st = E.sync())) break;
      for (uint32_t ri : in_flight) {
        Circuit* c = E.find(rec(ri).circuit);
        c->loops_left--;
        s->wire_updates += c->num_wires;
      }
    }
    cur ^= 1;
  }
  // Leave nothing resident, whatever happened; on an error every unfinished
  // request ends with that error.
  std::vector<uint32_t> left;
  for (Circuit* c : E.slots()) left.push_back(c->id);
  (void)E.remove_many(left.data(), left.size());
  (void)E.sync();
  std::lock_guard<std::mutex> g(s->mu);
  if (st) {
    s->sched_status = st;
    s->sched_error = ckt_last_error();
    for (Rec& r : s->recs)
      if (r.state != kDone) {
        r.state = kDone;
        r.res.status = st;
        r.res.finish_s = clock();
      }
  }
  s->finished.notify_all();
}

int start(ckt_server* s) {
  std::lock_guard<std::mutex> g(s->mu);
  if (s->running) return fail(CKT_ERR_STATE, "server: already started");
  if (s->eng.count()) return fail(CKT_ERR_STATE, "server: circuits left resident");
  s->recs.clear();
  s->incoming.clear();
  s->cancels.clear();
  s->kept.clear();
  s->stop = false;
  s->drain = true;
  s->sched_status = CKT_OK;
  s->sched_error.clear();
  s->wire_updates = 0;
  s->iterations = 0;
  s->max_resident = 0;
  s->last_finish = 0;
  s->t0 = now_s();
  s->running = true;
  s->thread = std::thread(schedule, s);
  return CKT_OK;
}

int stop(ckt_server* s, bool drain) {
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (!s->running) return fail(CKT_ERR_STATE, "server: not started");
    s->stop = true;
    s->drain = drain;
    s->wake_seq++;
    s->wake.notify_all();
  }
  s->thread.join();
  std::lock_guard<std::mutex> g(s->mu);
  s->running = false;
  if (s->sched_status) return fail(s->sched_status, s->sched_error);
  return CKT_OK;
}

}  // namespace

extern "C" {

int ckt_server_create(int device, const ckt_server_config_t* config, ckt_server_t** out) {
  if (!out) return fail(CKT_ERR_INVALID_ARG, "ckt_server_create: null out");
  *out = nullptr;
  ckt_server_config_t cfg{};
  if (config) cfg = *config;
  if (cfg.policy < CKT_POLICY_EXACT || cfg.policy > CKT_POLICY_SCALED)
    return fail(CKT_ERR_INVALID_ARG, "server: unknown policy");
  if (cfg.reduce_mode < CKT_REDUCE_ATOMIC || cfg.reduce_mode > CKT_REDUCE_CSR_KAHAN64)
    return fail(CKT_ERR_INVALID_ARG, "server: unknown reduction mode");
  if (cfg.gen_threads <= 0) cfg.gen_threads = 8;
  ckt_server_t* s = new ckt_server_t();
  int st = s->eng.init(device, cfg.hbm_budget_bytes);
  if (st) {
    delete s;
    return st;
  }
  cfg.hbm_budget_bytes = s->eng.budget();
  s->cfg = cfg;
  *out = s;
  return CKT_OK;
}

int ckt_server_destroy(ckt_server_t* s) {
  if (!s) return CKT_OK;
  bool running;
  {
    std::lock_guard<std::mutex> g(s->mu);
    running = s->running;
  }
  if (running) (void)stop(s, false);
  delete s;
  return CKT_OK;
}

int ckt_server_set_keep_state(ckt_server_t* s, int keep_state) {
  if (!s) return fail(CKT_ERR_INVALID_ARG, "null server");
  std::lock_guard<std::mutex> g(s->mu);
  if (s->running) return fail(CKT_ERR_STATE, "server: set keep_state before starting");
  s->keep_state = keep_state != 0;
  return CKT_OK;
}

int ckt_server_set_keep_limit(ckt_server_t* s, uint32_t max_wires) {
  if (!s) return fail(CKT_ERR_INVALID_ARG, "null server");
  std::lock_guard<std::mutex> g(s->mu);
  if (s->running) return fail(CKT_ERR_STATE, "server: set the keep limit before starting");
  s->keep_max_wires = max_wires;
  return CKT_OK;
}

int ckt_server_fetch(ckt_server_t* s, uint32_t index, ckt_node_t* nodes, ckt_wire_t* wires) {
  if (!s) return fail(CKT_ERR_INVALID_ARG, "null server");
  std::lock_guard<std::mutex> g(s->mu);
  auto k = s->kept.find(index);
  if (k == s->kept.end())
    return fail(CKT_ERR_NOT_FOUND, "no kept state for request " + std::to_string(index));
  if (nodes) std::copy(k->second.nodes.begin(), k->second.nodes.end(), nodes);
  if (wires) std::copy(k->second.wires.begin(), k->second.wires.end(), wires);
  return CKT_OK;
}

int ckt_server_start(ckt_server_t* s) {
  if (!s) return fail(CKT_ERR_INVALID_ARG, "null server");
  return start(s);
}

int ckt_server_stop(ckt_server_t* s, int drain) {
  if (!s) return fail(CKT_ERR_INVALID_ARG, "null server");
  return stop(s, drain != 0);
}

int ckt_server_submit(ckt_server_t* s, const ckt_gen_params_t* gen, int32_t num_loops,
                      double timeout_s, uint32_t* id) {
  if (!s || !gen || !id) return fail(CKT_ERR_INVALID_ARG, "ckt_server_submit: null argument");
  uint64_t need = 0;
  uint32_t wires = 0;
  if (int st = validate(*gen, num_loops, s->cfg.hbm_budget_bytes, &need, &wires)) return st;
  std::lock_guard<std::mutex> g(s->mu);
  if (!s->running || s->stop) return fail(CKT_ERR_STATE, "server: not accepting requests");
  *id = enqueue(s, *gen, num_loops, timeout_s, now_s() - s->t0, need, wires);
  return CKT_OK;
}

int ckt_server_poll(ckt_server_t* s, uint32_t id, int* finished, ckt_request_result_t* result) {
  if (!s || !finished) return fail(CKT_ERR_INVALID_ARG, "ckt_server_poll: null argument");
  std::lock_guard<std::mutex> g(s->mu);
  if (id >= s->recs.size()) return fail(CKT_ERR_NOT_FOUND, "server: unknown request id");
  const Rec& r = s->recs[id];
  *finished = r.state == kDone;
  if (result) *result = r.res;  // admit_s > 0 once resident
  return CKT_OK;
}

int ckt_server_wait(ckt_server_t* s, uint32_t id, double timeout_s, ckt_request_result_t* result) {
  if (!s) return fail(CKT_ERR_INVALID_ARG, "null server");
  std::unique_lock<std::mutex> g(s->mu);
  if (id >= s->recs.size()) return fail(CKT_ERR_NOT_FOUND, "server: unknown request id");
  auto done = [&] { return s->recs[id].state == kDone; };
  if (timeout_s < 0) s->finished.wait(g, done);
  else if (!s->finished.wait_for(g, std::chrono::duration<double>(timeout_s), done))
    return fail(CKT_ERR_TIMEOUT, "server: request " + std::to_string(id) + " still running");
  if (result) *result = s->recs[id].res;
  return CKT_OK;
}

int ckt_server_cancel(ckt_server_t* s, uint32_t id) {
  if (!s) return fail(CKT_ERR_INVALID_ARG, "null server");
  std::lock_guard<std::mutex> g(s->mu);
  if (id >= s->recs.size()) return fail(CKT_ERR_NOT_FOUND, "server: unknown request id");
  if (s->recs[id].state == kDone)
    return fail(CKT_ERR_STATE, "server: request " + std::to_string(id) + " already finished");
  s->cancels.push_back(id);
  s->wake_seq++;
  s->wake.notify_all();
  return CKT_OK;
}

int ckt_server_run_trace(ckt_server_t* s, const ckt_request_t* reqs, uint32_t n,
                         ckt_request_result_t* results, ckt_server_report_t* report) {
  if (!s) return fail(CKT_ERR_INVALID_ARG, "null server");
  if (n && (!reqs || !results)) return fail(CKT_ERR_INVALID_ARG, "server: null request/result array");
  // Requests that can never run are rejected up front.
  std::vector<uint64_t> need(n, 0);
  std::vector<uint32_t> wires(n, 0);
  std::vector<int> reject(n, CKT_OK);
  int32_t rejected = 0;
  double offered = 0, last_arrival = 0;
  for (uint32_t i = 0; i < n; i++) {
    reject[i] = validate(reqs[i].gen, reqs[i].num_loops, s->cfg.hbm_budget_bytes, &need[i],
                         &wires[i]);
    if (reject[i]) {
      rejected++;
      continue;
    }
    offered += (double)wires[i] * reqs[i].num_loops;
    last_arrival = std::max(last_arrival, reqs[i].arrival_s);
  }
  if (int st = start(s)) return st;
  std::vector<uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t a, uint32_t b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
  // Each request is submitted at its arrival offset from the scheduler's
  // start, with that offset as its arrival time.
  std::vector<int64_t> id_of(n, -1);
  for (uint32_t i : order) {
    if (reject[i]) continue;
    const double wake_at = s->t0 + reqs[i].arrival_s;
    for (double t = now_s(); t < wake_at; t = now_s())
      std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wake_at - t, 0.01)));
    std::lock_guard<std::mutex> g(s->mu);
    id_of[i] = enqueue(s, reqs[i].gen, reqs[i].num_loops, reqs[i].timeout_s, reqs[i].arrival_s,
                       need[i], wires[i]);
  }
  const int st = stop(s, true);
  std::vector<double> latencies;
  int32_t timed_out = 0;
  {
    std::lock_guard<std::mutex> g(s->mu);
    std::unordered_map<uint32_t, KeptState> by_index;
    for (uint32_t i = 0; i < n; i++) {
      ckt_request_result_t& r = results[i];
      if (id_of[i] < 0) {
        r = ckt_request_result_t{};
        r.status = reject[i];
        r.num_loops = reqs[i].num_loops;
        r.arrival_s = r.finish_s = reqs[i].arrival_s;
        continue;
      }
      r = s->recs[(size_t)id_of[i]].res;
      if (r.status == CKT_OK) latencies.push_back(r.finish_s - r.arrival_s);
      if (r.status == CKT_ERR_TIMEOUT) timed_out++;
      auto k = s->kept.find((uint32_t)id_of[i]);
      if (k != s->kept.end()) by_index[i] = std::move(k->second);
    }
    s->kept.swap(by_index);
  }
  if (st) return st;
  if (report) {
    const double elapsed = s->last_finish;
    report->elapsed_s = elapsed;
    report->wire_updates = s->wire_updates;
    report->wire_updates_per_s = elapsed > 0 ? s->wire_updates / elapsed : 0;
    report->offered_wire_updates_per_s = last_arrival > 0 ? offered / last_arrival : 0;
    report->latency_p50_s = percentile(latencies, 0.50);
    report->latency_p99_s = percentile(latencies, 0.99);
    report->latency_mean_s =
        latencies.empty() ? 0 : std::accumulate(latencies.begin(), latencies.end(), 0.0) / latencies.size();
    report->latency_max_s = latencies.empty() ? 0 : *std::max_element(latencies.begin(), latencies.end());
    report->completed = (int32_t)latencies.size();
    report->rejected = rejected;
    report->timed_out = timed_out;
    report->iterations = s->iterations;
    report->max_resident = s->max_resident;
  }
  return CKT_OK;
}

}  // extern "C"
