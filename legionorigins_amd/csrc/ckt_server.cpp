// ckt_server.cpp -- continuous batching of circuit simulations.
//
// A request is a circuit (generator parameters) and a loop count.  The
// scheduler replays a trace in real time:
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
// arrivals and admits while it runs, so the GPU waits only for the launch.
// Circuits are independent, so a circuit's result does not depend on what
// else shares its iterations (bit-identical with EXACT + CSR).
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
#include <vector>

#include "ckt_engine.h"

using namespace ckt;

struct ckt_server {
  Engine eng;
  ckt_server_config_t cfg{};
  bool keep_state = false;
  std::vector<std::vector<ckt_node_t>> kept_nodes;
  std::vector<std::vector<ckt_wire_t>> kept_wires;
};

namespace {

double now_s() {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

struct Ready {
  uint32_t index;
  std::unique_ptr<DevgenPlan> plan;
  int status;
  std::string error;
};

// Host threads that walk admitted requests' generator chains.
class GenPool {
 public:
  GenPool(int threads, const ckt_request_t* reqs) : reqs_(reqs) {
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
  void submit(uint32_t index) {
    {
      std::lock_guard<std::mutex> g(mu_);
      jobs_.push_back(index);
    }
    cv_.notify_one();
  }
  // Moves finished images out; waits up to `timeout_s` if none is ready.
  void take(std::vector<Ready>& out, double timeout_s) {
    std::unique_lock<std::mutex> g(mu_);
    if (done_.empty() && timeout_s > 0)
      done_cv_.wait_for(g, std::chrono::duration<double>(timeout_s), [this] { return !done_.empty(); });
    for (auto& r : done_) out.push_back(std::move(r));
    done_.clear();
  }

 private:
  void work() {
    for (;;) {
      uint32_t idx;
      {
        std::unique_lock<std::mutex> g(mu_);
        cv_.wait(g, [this] { return stop_ || !jobs_.empty(); });
        if (stop_) return;
        idx = jobs_.front();
        jobs_.pop_front();
      }
      Ready r;
      r.index = idx;
      r.plan.reset(new DevgenPlan());
      r.status = devgen_plan(&reqs_[idx].gen, *r.plan);
      if (r.status) {
        r.error = ckt_last_error();
        r.plan.reset();
      }
      {
        std::lock_guard<std::mutex> g(mu_);
        done_.push_back(std::move(r));
      }
      done_cv_.notify_all();
    }
  }
  const ckt_request_t* reqs_;
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::deque<uint32_t> jobs_;
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
  delete s;
  return CKT_OK;
}

int ckt_server_set_keep_state(ckt_server_t* s, int keep_state) {
  if (!s) return fail(CKT_ERR_INVALID_ARG, "null server");
  s->keep_state = keep_state != 0;
  return CKT_OK;
}

int ckt_server_fetch(ckt_server_t* s, uint32_t index, ckt_node_t* nodes, ckt_wire_t* wires) {
  if (!s) return fail(CKT_ERR_INVALID_ARG, "null server");
  if (index >= s->kept_nodes.size() || s->kept_nodes[index].empty())
    return fail(CKT_ERR_NOT_FOUND, "no kept state for request " + std::to_string(index));
  if (nodes) std::copy(s->kept_nodes[index].begin(), s->kept_nodes[index].end(), nodes);
  if (wires) std::copy(s->kept_wires[index].begin(), s->kept_wires[index].end(), wires);
  return CKT_OK;
}

int ckt_server_run_trace(ckt_server_t* s, const ckt_request_t* reqs, uint32_t n,
                         ckt_request_result_t* results, ckt_server_report_t* report) {
  if (!s) return fail(CKT_ERR_INVALID_ARG, "null server");
  if (n && (!reqs || !results)) return fail(CKT_ERR_INVALID_ARG, "server: null request/result array");
  Engine& E = s->eng;
  CKT_HIP(hipSetDevice(E.device()));
  if (E.count()) return fail(CKT_ERR_STATE, "server: circuits left from a previous trace");
  const ckt_server_config_t& cfg = s->cfg;
  if (s->keep_state) {
    s->kept_nodes.assign(n, {});
    s->kept_wires.assign(n, {});
  }

  // Per-request bookkeeping.
  std::vector<uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t a, uint32_t b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
  std::vector<uint64_t> need(n, 0);
  std::vector<uint32_t> wires_of(n, 0);
  double offered = 0, last_arrival = 0;
  for (uint32_t i = 0; i < n; i++) {
    ckt_request_result_t& r = results[i];
    r = ckt_request_result_t{};
    r.status = CKT_ERR_CANCELLED;
    r.num_loops = reqs[i].num_loops;
    r.arrival_s = reqs[i].arrival_s;
    uint32_t N = 0, W = 0, P = 0;
    int st = gen_validate(&reqs[i].gen, &N, &W, &P);
    if (st || reqs[i].num_loops < 0) {
      r.status = st ? st : CKT_ERR_INVALID_ARG;
      need[i] = UINT64_MAX;
      continue;
    }
    need[i] = SlabLayout::make(N, W, P).total;
    wires_of[i] = W;
    r.num_wires = W;
    offered += (double)W * reqs[i].num_loops;
    last_arrival = std::max(last_arrival, reqs[i].arrival_s);
  }

  GenPool pool(cfg.gen_threads, reqs);
  std::deque<uint32_t> queue;            // arrived, not admitted
  size_t next_arrival = 0;               // index into `order`
  uint64_t reserved = 0;                 // bytes of admitted, not retired requests
  uint64_t admitted_wires = 0;           // wires of admitted, not retired requests
  int32_t admitted = 0;                  // admitted, not retired requests
  std::vector<int64_t> circuit_of(n, -1);  // request -> circuit id
  std::vector<uint32_t> request_of_circuit;
  std::vector<int32_t> loops_left(n, 0);
  uint32_t finished = 0, rejected = 0, timed_out = 0;
  int32_t iterations = 0, max_resident = 0;
  double wire_updates = 0;
  std::vector<double> latencies;
  int st = CKT_OK;
  const double t0 = now_s();
  auto clock = [&] { return now_s() - t0; };

  auto retire = [&](uint32_t ri, int status) {
    ckt_request_result_t& r = results[ri];
    r.status = status;
    r.finish_s = clock();
    if (status == CKT_OK) latencies.push_back(r.finish_s - r.arrival_s);
    finished++;
  };

  // Requests that can never fit are rejected up front.
  for (uint32_t i = 0; i < n; i++)
    if (need[i] == UINT64_MAX || need[i] > cfg.hbm_budget_bytes) {
      if (need[i] != UINT64_MAX) results[i].status = CKT_ERR_BUDGET;
      results[i].finish_s = results[i].arrival_s;
      finished++;
      rejected++;
    }

  std::vector<Ready> ready;
  std::vector<uint32_t> in_flight;  // requests advanced by the running iteration
  // Circuits finishing with an iteration, as (circuit id, offset in that
  // iteration's download buffer).  Two buffers alternate: the voltages of the
  // iteration that just ended are summed while the next one, launched first,
  // copies its own finishers' voltages into the other buffer.
  std::vector<std::pair<uint32_t, size_t>> finishing[2];
  PinnedBuffer downloads[2];
  int cur = 0;
  while (finished < n && !st) {
    // Launch one iteration of every resident circuit with loops left.  It is
    // the first thing after the previous iteration ended, so that the
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
      iterations++;
      // Circuits that finish with this iteration: their final voltages are
      // copied to pinned memory behind it, for the checksum after it ends.
      size_t total = 0;
      for (uint32_t ri : in_flight) {
        const Circuit* c = E.find((uint32_t)circuit_of[ri]);
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
    // whose voltages are in the other buffer, and zero-loop requests.  None
    // of them is in the iteration just launched.
    std::vector<uint32_t> done;
    for (Circuit* c : E.slots())
      if (c->loops_left <= 0) done.push_back(c->id);
    const auto& prev = finishing[cur ^ 1];
    for (uint32_t id : done) {
      const uint32_t ri = request_of_circuit[id];
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
      results[ri].final_voltage_checksum = sum;
      if (s->keep_state) {
        Circuit* c = E.find(id);
        s->kept_nodes[ri].resize(c->num_nodes);
        s->kept_wires[ri].resize(c->num_wires);
        if ((st = E.store(id, s->kept_nodes[ri].data(), s->kept_wires[ri].data()))) break;
      }
    }
    if (st) break;
    // One removal for the whole set.  Slabs from the stream's memory pool are
    // released in stream order, behind the running iteration, which does not
    // use them; the slot table is rewritten by a copy queued behind it too.
    if ((st = E.remove_many(done.data(), done.size()))) break;
    for (uint32_t id : done) {
      const uint32_t ri = request_of_circuit[id];
      reserved -= need[ri];
      admitted_wires -= wires_of[ri];
      admitted--;
      retire(ri, CKT_OK);
    }
    const double now = clock();
    // Arrivals.
    while (next_arrival < order.size() && reqs[order[next_arrival]].arrival_s <= now) {
      const uint32_t i = order[next_arrival++];
      if (need[i] != UINT64_MAX && need[i] <= cfg.hbm_budget_bytes) queue.push_back(i);
    }
    // Load shedding: queued requests past their admission deadline leave.
    if (!queue.empty()) {
      std::deque<uint32_t> keep;
      for (uint32_t i : queue) {
        if (reqs[i].timeout_s > 0 && now - reqs[i].arrival_s > reqs[i].timeout_s) {
          retire(i, CKT_ERR_TIMEOUT);
          timed_out++;
        } else {
          keep.push_back(i);
        }
      }
      queue.swap(keep);
    }
    // Admission, first come first served.
    while (!queue.empty()) {
      const uint32_t i = queue.front();
      const bool fits_mem = reserved + need[i] <= cfg.hbm_budget_bytes;
      const bool fits_slots = cfg.max_active <= 0 || admitted < cfg.max_active;
      const bool fits_wires = cfg.max_active_wires == 0 || admitted == 0 ||
                              admitted_wires + wires_of[i] <= cfg.max_active_wires;
      if (!(fits_mem && fits_slots && fits_wires)) break;
      queue.pop_front();
      reserved += need[i];
      admitted_wires += wires_of[i];
      admitted++;
      pool.submit(i);
    }
    // Overlap: enqueue the generation of planned circuits behind the
    // running iteration.
    ready.clear();
    double wait = 0;
    if (refs.empty()) {
      // Idle GPU: sleep until the next arrival or a finished plan.
      const double next = next_arrival < order.size() ? reqs[order[next_arrival]].arrival_s : now + 0.05;
      wait = std::max(0.0, std::min(0.05, next - clock()));
      if (admitted == 0 && next_arrival >= order.size() && queue.empty()) wait = 0;
    }
    pool.take(ready, wait);
    for (Ready& r : ready) {
      if (r.status) {
        reserved -= need[r.index];
        admitted_wires -= wires_of[r.index];
        admitted--;
        retire(r.index, r.status);
        continue;
      }
      uint32_t id = 0;
      int as = E.add_planned(*r.plan, reqs[r.index].num_loops, &id);
      r.plan.reset();
      if (as) {
        reserved -= need[r.index];
        admitted_wires -= wires_of[r.index];
        admitted--;
        retire(r.index, as);
        continue;
      }
      if (request_of_circuit.size() <= id) request_of_circuit.resize(id + 1, 0);
      request_of_circuit[id] = r.index;
      circuit_of[r.index] = id;
      results[r.index].admit_s = clock();
      loops_left[r.index] = reqs[r.index].num_loops;
    }
    max_resident = std::max<int32_t>(max_resident, (int32_t)E.count());
    if (!refs.empty()) {
      if ((st = E.sync())) break;
      for (uint32_t ri : in_flight) {
        Circuit* c = E.find((uint32_t)circuit_of[ri]);
        c->loops_left--;
        wire_updates += c->num_wires;
      }
    }
    cur ^= 1;
  }
  const double elapsed = clock();
  // Leave nothing resident, whatever happened.
  std::vector<uint32_t> left;
  for (Circuit* c : E.slots()) left.push_back(c->id);
  (void)E.remove_many(left.data(), left.size());
  if (st) return st;
  if (report) {
    report->elapsed_s = elapsed;
    report->wire_updates = wire_updates;
    report->wire_updates_per_s = elapsed > 0 ? wire_updates / elapsed : 0;
    report->offered_wire_updates_per_s = last_arrival > 0 ? offered / last_arrival : 0;
    report->latency_p50_s = percentile(latencies, 0.50);
    report->latency_p99_s = percentile(latencies, 0.99);
    report->latency_mean_s =
        latencies.empty() ? 0 : std::accumulate(latencies.begin(), latencies.end(), 0.0) / latencies.size();
    report->latency_max_s = latencies.empty() ? 0 : *std::max_element(latencies.begin(), latencies.end());
    report->completed = (int32_t)latencies.size();
    report->rejected = (int32_t)rejected;
    report->timed_out = (int32_t)timed_out;
    report->iterations = iterations;
    report->max_resident = max_resident;
  }
  return CKT_OK;
}

}  // extern "C"
