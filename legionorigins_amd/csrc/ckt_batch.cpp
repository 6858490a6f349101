// ckt_batch.cpp -- many independent circuits advanced by shared launches.
//
// Every batch iteration runs one loop (calc_new_currents, distribute_charge,
// update_voltages) of every circuit that still has loops left, each task as
// ONE launch whose tiles span all those circuits.  Iterations with an
// unchanged circuit set replay one captured hipGraph.
#include <time.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ckt_engine.h"

using namespace ckt;

// Calls on one batch from several host threads are serialised by its mutex.
struct ckt_batch {
  Engine eng;
  ckt_batch_config_t applied{};
  bool have_applied = false;
  std::recursive_mutex mu;
};

namespace {

// Holds the batch's mutex for the rest of an entry point (nothing for a null
// batch, which the entry point then rejects).
class BatchLock {
 public:
  explicit BatchLock(ckt_batch_t* b) {
    if (b) lk_ = std::unique_lock<std::recursive_mutex>(b->mu);
  }

 private:
  std::unique_lock<std::recursive_mutex> lk_;
};

double now_s() {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

int check(ckt_batch_t* b) {
  if (!b) return fail(CKT_ERR_INVALID_ARG, "null batch");
  hipError_t e = hipSetDevice(b->eng.device());
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  return CKT_OK;
}

int apply_config(Engine& E, const ckt_batch_config_t& cfg) {
  int st;
  if (cfg.block_threads && (st = E.set_option(OPT_CNC_BLOCK, cfg.block_threads))) return st;
  if ((st = E.set_option(OPT_PERSISTENT, cfg.persistent))) return st;
  if (cfg.waves_per_simd && (st = E.set_option(OPT_WAVES_PER_SIMD, cfg.waves_per_simd))) return st;
  if (cfg.wires_per_lane) {
    if ((st = E.set_option(OPT_WIRES_PER_LANE, cfg.wires_per_lane))) return st;
    if ((st = E.set_option(OPT_TILE_SPAN, cfg.tile_span))) return st;
    if ((st = E.set_option(OPT_TILE_ORDER, cfg.tile_order))) return st;
  }
  return CKT_OK;
}

int apply_config(ckt_batch_t* b, const ckt_batch_config_t& cfg) { return apply_config(b->eng, cfg); }

// The probes of ckt_batch_autotune / ckt_batch_probe: calc_new_currents over
// every circuit at a shortened step count, timed with events on the engine
// stream.  The constructor saves each circuit's currents and internal
// voltages (one contiguous region of its slab) and the engine options; the
// destructor puts both back.
class Prober {
 public:
  explicit Prober(Engine& E) : E_(E) {}
  int begin(int32_t probe_steps) {
    for (int o = 0; o < kOpts; o++) opts_[o] = E_.option(kOptIds[o]);
    began_ = true;
    hipStream_t s = E_.stream();
    for (Circuit* c : E_.slots()) {
      const size_t bytes = c->lay.ind - c->lay.cur;
      void* p = nullptr;
      CKT_HIP(hipMalloc(&p, bytes));
      saved_.push_back(p);
      CKT_HIP(hipMemcpyAsync(p, (uint8_t*)c->slab + c->lay.cur, bytes, hipMemcpyDeviceToDevice, s));
    }
    E_.all_refs(refs_, false);
    CKT_HIP(hipEventCreate(&a_));
    CKT_HIP(hipEventCreate(&z_));
    return E_.set_option(OPT_STEPS, probe_steps);
  }
  // The faster of two launches under `cfg`'s launch shape and schedule.
  int time(int32_t policy, const ckt_batch_config_t& cfg, double* ms) {
    int st;
    if ((st = apply_config(E_, cfg))) return st;
    double t = 1e30;
    for (int rep = 0; rep < 2; rep++) {
      if ((st = E_.sync())) return st;
      CKT_HIP(hipEventRecord(a_, E_.stream()));
      if ((st = E_.cnc(refs_, policy))) return st;
      CKT_HIP(hipEventRecord(z_, E_.stream()));
      if ((st = E_.sync())) return st;
      float f = 0;
      CKT_HIP(hipEventElapsedTime(&f, a_, z_));
      t = std::min(t, (double)f);
    }
    *ms = t;
    return CKT_OK;
  }
  ~Prober() {
    if (a_) (void)hipEventDestroy(a_);
    if (z_) (void)hipEventDestroy(z_);
    if (began_)
      for (int o = 0; o < kOpts; o++) (void)E_.set_option(kOptIds[o], opts_[o]);
    hipStream_t s = E_.stream();
    for (size_t i = 0; i < saved_.size(); i++) {
      Circuit* c = E_.slots()[i];
      (void)hipMemcpyAsync((uint8_t*)c->slab + c->lay.cur, saved_[i], c->lay.ind - c->lay.cur,
                           hipMemcpyDeviceToDevice, s);
    }
    (void)hipStreamSynchronize(s);
    for (void* p : saved_) (void)hipFree(p);
  }

 private:
  static constexpr int kOpts = 7;
  static constexpr int kOptIds[kOpts] = {OPT_STEPS,          OPT_CNC_BLOCK, OPT_PERSISTENT,
                                         OPT_WAVES_PER_SIMD, OPT_WIRES_PER_LANE,
                                         OPT_TILE_SPAN,      OPT_TILE_ORDER};
  Engine& E_;
  std::vector<void*> saved_;
  std::vector<PieceRef> refs_;
  int64_t opts_[kOpts] = {};
  bool began_ = false;
  hipEvent_t a_ = nullptr, z_ = nullptr;
};

}  // namespace

extern "C" {

int ckt_batch_create(int device, uint64_t hbm_budget_bytes, ckt_batch_t** out) {
  if (!out) return fail(CKT_ERR_INVALID_ARG, "ckt_batch_create: null out");
  *out = nullptr;
  ckt_batch_t* b = new ckt_batch_t();
  int st = b->eng.init(device, hbm_budget_bytes);
  if (st) {
    delete b;
    return st;
  }
  *out = b;
  return CKT_OK;
}

int ckt_batch_destroy(ckt_batch_t* b) {
  delete b;
  return CKT_OK;
}

int ckt_batch_set_option(ckt_batch_t* b, int option, int64_t value) {
  BatchLock lk(b);
  int st = check(b);
  if (st) return st;
  return b->eng.set_option(option, value);
}

int ckt_batch_set_timing(ckt_batch_t* b, int enable) {
  BatchLock lk(b);
  int st = check(b);
  if (st) return st;
  b->eng.set_timing(enable != 0);
  return CKT_OK;
}

int ckt_batch_kernel_time(ckt_batch_t* b, int task_id, double* total_ms, uint64_t* launches) {
  BatchLock lk(b);
  int st = check(b);
  if (st) return st;
  if (task_id < 0 || task_id >= CKT_NUM_TASK_IDS) return fail(CKT_ERR_INVALID_ARG, "bad task id");
  return b->eng.kernel_time(task_id, total_ms, launches);
}

int ckt_footprint_bytes(uint32_t num_nodes, uint32_t num_wires, uint32_t num_pieces, uint64_t* bytes) {
  if (!bytes) return fail(CKT_ERR_INVALID_ARG, "null out");
  *bytes = SlabLayout::make(num_nodes, num_wires, num_pieces).total;
  return CKT_OK;
}

int ckt_batch_add_generated(ckt_batch_t* b, const ckt_gen_params_t* params, int32_t num_loops,
                            uint32_t* id) {
  BatchLock lk(b);
  int st = check(b);
  if (st) return st;
  if (num_loops < 0) return fail(CKT_ERR_INVALID_ARG, "negative loop count");
  uint32_t N, W, P;
  if ((st = gen_validate(params, &N, &W, &P))) return st;
  const uint64_t need = SlabLayout::make(N, W, P).total;
  if (b->eng.budget() && b->eng.bytes_in_use() + need > b->eng.budget())
    return fail(CKT_ERR_BUDGET, "circuit does not fit in the remaining HBM budget");
  SlabImage img;
  if ((st = image_from_gen(img, params))) return st;
  return b->eng.add(img, num_loops, id);
}

int ckt_batch_add_generated_device(ckt_batch_t* b, const ckt_gen_params_t* params,
                                   int32_t num_loops, uint32_t* id) {
  BatchLock lk(b);
  int st = check(b);
  if (st) return st;
  if (num_loops < 0) return fail(CKT_ERR_INVALID_ARG, "negative loop count");
  return b->eng.add_generated_on_device(params, num_loops, id);
}

int ckt_batch_add_generated_many(ckt_batch_t* b, const ckt_gen_params_t* params,
                                 const int32_t* num_loops, uint32_t n, int32_t threads,
                                 uint32_t* ids) {
  BatchLock lk(b);
  int st = check(b);
  if (st) return st;
  if (n && (!params || !num_loops || !ids)) return fail(CKT_ERR_INVALID_ARG, "null array");
  uint64_t need = 0;
  for (uint32_t i = 0; i < n; i++) {
    uint32_t N, W, P;
    if ((st = gen_validate(&params[i], &N, &W, &P))) return st;
    if (num_loops[i] < 0) return fail(CKT_ERR_INVALID_ARG, "negative loop count");
    need += SlabLayout::make(N, W, P).total;
  }
  if (b->eng.budget() && b->eng.bytes_in_use() + need > b->eng.budget())
    return fail(CKT_ERR_BUDGET, "circuits do not fit in the remaining HBM budget");
  if (threads < 1) threads = 1;
  // Images are built in windows of `threads` circuits so that host memory
  // holds at most one window at a time.
  for (uint32_t base = 0; base < n; base += (uint32_t)threads) {
    const uint32_t cnt = std::min<uint32_t>((uint32_t)threads, n - base);
    std::vector<SlabImage> imgs(cnt);
    std::vector<int> rc(cnt, CKT_OK);
    std::vector<std::string> err(cnt);
    std::vector<std::thread> th;
    for (uint32_t k = 0; k < cnt; k++)
      th.emplace_back([&, k] {
        rc[k] = image_from_gen(imgs[k], &params[base + k]);
        if (rc[k]) err[k] = ckt_last_error();
      });
    for (auto& t : th) t.join();
    for (uint32_t k = 0; k < cnt; k++) {
      if (rc[k]) return fail(rc[k], err[k]);
      if ((st = b->eng.add(imgs[k], num_loops[base + k], &ids[base + k]))) return st;
      imgs[k] = SlabImage();
    }
  }
  return CKT_OK;
}

int ckt_batch_add(ckt_batch_t* b, const ckt_node_t* nodes, uint32_t num_nodes,
                  const ckt_wire_t* wires, uint32_t num_wires, const int32_t* locator,
                  const ckt_piece_t* pieces, uint32_t num_pieces, int32_t num_loops, uint32_t* id) {
  BatchLock lk(b);
  int st = check(b);
  if (st) return st;
  if (num_loops < 0) return fail(CKT_ERR_INVALID_ARG, "negative loop count");
  SlabImage img;
  if ((st = image_from_aos(img, nodes, num_nodes, wires, num_wires, locator, pieces, num_pieces)))
    return st;
  return b->eng.add(img, num_loops, id);
}

int ckt_batch_remove(ckt_batch_t* b, uint32_t id) {
  BatchLock lk(b);
  int st = check(b);
  if (st) return st;
  return b->eng.remove(id);
}

int ckt_batch_count(ckt_batch_t* b, uint32_t* num_circuits, uint64_t* bytes_in_use) {
  BatchLock lk(b);
  int st = check(b);
  if (st) return st;
  if (num_circuits) *num_circuits = (uint32_t)b->eng.count();
  if (bytes_in_use) *bytes_in_use = b->eng.bytes_in_use();
  return CKT_OK;
}

int ckt_batch_loops_left(ckt_batch_t* b, uint32_t id, int32_t* loops_left) {
  BatchLock lk(b);
  int st = check(b);
  if (st) return st;
  Circuit* c = b->eng.find(id);
  if (!c) return fail(CKT_ERR_NOT_FOUND, "no circuit with id " + std::to_string(id));
  if (loops_left) *loops_left = c->loops_left;
  return CKT_OK;
}

int ckt_batch_step(ckt_batch_t* b, const ckt_batch_config_t* config, int32_t* active) {
  BatchLock lk(b);
  int st = check(b);
  if (st) return st;
  const ckt_batch_config_t def = {CKT_POLICY_FAST, CKT_REDUCE_ATOMIC, 1, 0, 0, 0};
  const ckt_batch_config_t& cfg = config ? *config : def;
  if ((st = apply_config(b, cfg))) return st;
  Engine& E = b->eng;
  std::vector<PieceRef> refs;
  E.all_refs(refs, true);
  if (!refs.empty()) {
    if (cfg.use_graph) {
      if ((st = E.run_graph(refs, cfg.policy, cfg.reduce_mode, 1))) return st;
    } else {
      if ((st = E.loop_once(refs, cfg.policy, cfg.reduce_mode))) return st;
    }
  }
  int32_t n = 0;
  for (Circuit* c : E.slots())
    if (c->loops_left > 0) {
      c->loops_left--;
      if (c->loops_left > 0) n++;
    }
  if (active) *active = n;
  return CKT_OK;
}

int ckt_batch_run(ckt_batch_t* b, const ckt_batch_config_t* config, ckt_batch_report_t* report) {
  BatchLock lk(b);
  int st = check(b);
  if (st) return st;
  const ckt_batch_config_t def = {CKT_POLICY_FAST, CKT_REDUCE_ATOMIC, 1, 0, 0, 0};
  const ckt_batch_config_t& cfg = config ? *config : def;
  if ((st = apply_config(b, cfg))) return st;
  Engine& E = b->eng;
  double cnc0 = 0;
  E.kernel_time(CKT_CALC_NEW_CURRENTS, &cnc0, nullptr);
  if ((st = E.sync())) return st;
  double wu = 0;
  int32_t iters = 0;
  const double t0 = now_s();
  std::vector<PieceRef> refs;
  for (;;) {
    E.all_refs(refs, true);
    if (refs.empty()) break;
    // Replay the same graph while the set of active circuits is unchanged.
    int32_t k = INT32_MAX;
    for (Circuit* c : E.slots())
      if (c->loops_left > 0) k = std::min(k, c->loops_left);
    if (cfg.use_graph) {
      if ((st = E.run_graph(refs, cfg.policy, cfg.reduce_mode, k))) return st;
    } else {
      for (int32_t i = 0; i < k; i++)
        if ((st = E.loop_once(refs, cfg.policy, cfg.reduce_mode))) return st;
    }
    for (Circuit* c : E.slots())
      if (c->loops_left > 0) {
        c->loops_left -= k;
        wu += (double)c->num_wires * k;
      }
    iters += k;
  }
  if ((st = E.sync())) return st;
  const double elapsed = now_s() - t0;
  if (report) {
    double cnc1 = 0;
    E.kernel_time(CKT_CALC_NEW_CURRENTS, &cnc1, nullptr);
    report->elapsed_s = elapsed;
    report->wire_updates = wu;
    report->wire_updates_per_s = elapsed > 0 ? wu / elapsed : 0;
    report->cnc_ms = cnc1 - cnc0;
    report->loops_run = iters;
  }
  return CKT_OK;
}

int ckt_batch_store(ckt_batch_t* b, uint32_t id, ckt_node_t* nodes, ckt_wire_t* wires) {
  BatchLock lk(b);
  int st = check(b);
  if (st) return st;
  return b->eng.store(id, nodes, wires);
}

// Two stages, each keeping the fastest candidate: the launch shape
// (workgroup size x plain grid or persistent queue x resident waves) with the
// default tile schedule, then the tile schedule (wires per lane x span x
// order) with that shape.
int ckt_batch_autotune(ckt_batch_t* b, int32_t policy, int32_t probe_steps,
                       ckt_batch_config_t* best, double* best_ms) {
  BatchLock lk(b);
  int st = check(b);
  if (st) return st;
  if (!best) return fail(CKT_ERR_INVALID_ARG, "null best");
  if (probe_steps < 1) probe_steps = 500;
  Engine& E = b->eng;
  if (E.count() == 0) return fail(CKT_ERR_STATE, "empty batch");
  Prober P(E);
  if ((st = P.begin(probe_steps))) return st;
  double best_t = 1e30;
  ckt_batch_config_t bc{policy, CKT_REDUCE_ATOMIC, 1, 256, 0, 8, 1, 0, 0};
  for (int pers = 0; pers <= 1; pers++)
    for (int block : {64, 128, 256, 512})
      for (int waves : {4, 8}) {
        if (!pers && waves != 8) continue;
        ckt_batch_config_t c = bc;
        c.block_threads = block;
        c.persistent = pers;
        c.waves_per_simd = waves;
        double t;
        if ((st = P.time(policy, c, &t))) return st;
        if (t < best_t) best_t = t, bc = c;
      }
  const ckt_batch_config_t shape = bc;
  for (int lane : {1, 2, 4})
    for (int span : {0, 1})
      for (int order : {0, 1, 2}) {
        if (lane == 1 && span == 0 && order == 0) continue;  // timed in the first stage
        ckt_batch_config_t c = shape;
        c.wires_per_lane = lane;
        c.tile_span = span;
        c.tile_order = order;
        double t;
        if ((st = P.time(policy, c, &t))) return st;
        if (t < best_t) best_t = t, bc = c;
      }
  *best = bc;
  if (best_ms) *best_ms = best_t;
  return CKT_OK;
}

int ckt_batch_probe(ckt_batch_t* b, const ckt_batch_config_t* config, int32_t probe_steps,
                    double* ms) {
  BatchLock lk(b);
  int st = check(b);
  if (st) return st;
  if (!config || !ms) return fail(CKT_ERR_INVALID_ARG, "null argument");
  if (probe_steps < 1) return fail(CKT_ERR_INVALID_ARG, "probe_steps must be positive");
  Engine& E = b->eng;
  if (E.count() == 0) return fail(CKT_ERR_STATE, "empty batch");
  Prober P(E);
  if ((st = P.begin(probe_steps))) return st;
  return P.time(config->policy, *config, ms);
}

}  // extern "C"
