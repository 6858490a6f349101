// ckt_ctx.cpp -- the single-circuit context: task entry points, the
// task-variant table and the region_main loop (apps/circuit/circuit.cc:123-346).
#include <time.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "ckt_engine.h"

using namespace ckt;

// Calls on one context from several host threads are serialised by its
// mutex (recursive: the task-variant table calls the task entry points).
struct ckt_ctx {
  Engine eng;
  uint32_t id = 0;
  bool loaded = false;
  std::recursive_mutex mu;
};

namespace {

// Holds the context's mutex for the rest of an entry point (nothing for a
// null context, which the entry point then rejects).
class CtxLock {
 public:
  explicit CtxLock(ckt_ctx_t* ctx) {
    if (ctx) lk_ = std::unique_lock<std::recursive_mutex>(ctx->mu);
  }

 private:
  std::unique_lock<std::recursive_mutex> lk_;
};

int check_ctx(ckt_ctx_t* ctx) {
  if (!ctx) return fail(CKT_ERR_INVALID_ARG, "null context");
  if (!ctx->loaded) return fail(CKT_ERR_STATE, "no circuit loaded in this context");
  hipError_t e = hipSetDevice(ctx->eng.device());
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  return CKT_OK;
}

// The loaded pieces named by `pieces`, each at most once (the loaded pieces
// partition the circuit, so a piece named twice is the only way one launch
// could process a wire or node twice).
int refs_from(ckt_ctx_t* ctx, const ckt_piece_t* pieces, uint32_t n, std::vector<PieceRef>& refs) {
  if (n && !pieces) return fail(CKT_ERR_INVALID_ARG, "null piece array");
  refs.resize(n);
  for (uint32_t i = 0; i < n; i++) {
    int st = ctx->eng.ref_for(0, pieces[i], &refs[i]);
    if (st) return st;
  }
  std::vector<uint32_t> seen(n);
  for (uint32_t i = 0; i < n; i++) seen[i] = refs[i].piece;
  std::sort(seen.begin(), seen.end());
  if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
    return fail(CKT_ERR_INVALID_ARG, "a piece is named twice in one call");
  return CKT_OK;
}

int install(ckt_ctx_t* ctx, SlabImage& img) {
  if (ctx->loaded) {
    int st = ctx->eng.remove(ctx->id);
    if (st) return st;
    ctx->loaded = false;
  }
  int st = ctx->eng.add(img, 0, &ctx->id);
  if (st) return st;
  ctx->loaded = true;
  return CKT_OK;
}

double now_s() {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

}  // namespace

extern "C" {

int ckt_ctx_create(int device, ckt_ctx_t** out) {
  if (!out) return fail(CKT_ERR_INVALID_ARG, "ckt_ctx_create: null out");
  *out = nullptr;
  ckt_ctx_t* ctx = new ckt_ctx_t();
  int st = ctx->eng.init(device, 0);
  if (st) {
    delete ctx;
    return st;
  }
  *out = ctx;
  return CKT_OK;
}

int ckt_ctx_destroy(ckt_ctx_t* ctx) {
  delete ctx;
  return CKT_OK;
}

void* ckt_ctx_stream(ckt_ctx_t* ctx) { return ctx ? (void*)ctx->eng.stream() : nullptr; }

int ckt_ctx_synchronize(ckt_ctx_t* ctx) {
  CtxLock lk(ctx);
  if (!ctx) return fail(CKT_ERR_INVALID_ARG, "null context");
  return ctx->eng.sync();
}

int ckt_ctx_load(ckt_ctx_t* ctx, const ckt_node_t* nodes, uint32_t num_nodes,
                 const ckt_wire_t* wires, uint32_t num_wires, const int32_t* locator,
                 const ckt_piece_t* pieces, uint32_t num_pieces) {
  CtxLock lk(ctx);
  if (!ctx) return fail(CKT_ERR_INVALID_ARG, "null context");
  SlabImage img;
  int st = image_from_aos(img, nodes, num_nodes, wires, num_wires, locator, pieces, num_pieces);
  if (st) return st;
  return install(ctx, img);
}

int ckt_ctx_generate(ckt_ctx_t* ctx, const ckt_gen_params_t* params) {
  CtxLock lk(ctx);
  if (!ctx) return fail(CKT_ERR_INVALID_ARG, "null context");
  SlabImage img;
  int st = image_from_gen(img, params);
  if (st) return st;
  return install(ctx, img);
}

int ckt_ctx_generate_device(ckt_ctx_t* ctx, const ckt_gen_params_t* params) {
  CtxLock lk(ctx);
  if (!ctx) return fail(CKT_ERR_INVALID_ARG, "null context");
  int st = gen_validate(params, nullptr, nullptr, nullptr);
  if (st) return st;
  if (ctx->loaded) {
    if ((st = ctx->eng.remove(ctx->id))) return st;
    ctx->loaded = false;
  }
  if ((st = ctx->eng.add_generated_on_device(params, 0, &ctx->id))) return st;
  ctx->loaded = true;
  return CKT_OK;
}

int ckt_ctx_store(ckt_ctx_t* ctx, ckt_node_t* nodes, ckt_wire_t* wires) {
  CtxLock lk(ctx);
  int st = check_ctx(ctx);
  if (st) return st;
  return ctx->eng.store(ctx->id, nodes, wires);
}

int ckt_ctx_store_topology(ckt_ctx_t* ctx, int32_t* locator, uint32_t* csr_row,
                           uint32_t* csr_ent) {
  CtxLock lk(ctx);
  int st = check_ctx(ctx);
  if (st) return st;
  return ctx->eng.store_topology(ctx->id, locator, csr_row, csr_ent);
}

int ckt_ctx_sizes(ckt_ctx_t* ctx, uint32_t* num_nodes, uint32_t* num_wires, uint32_t* num_pieces) {
  CtxLock lk(ctx);
  int st = check_ctx(ctx);
  if (st) return st;
  const Circuit* c = ctx->eng.find(ctx->id);
  if (num_nodes) *num_nodes = c->num_nodes;
  if (num_wires) *num_wires = c->num_wires;
  if (num_pieces) *num_pieces = c->num_pieces;
  return CKT_OK;
}

int ckt_ctx_pieces(ckt_ctx_t* ctx, ckt_piece_t* pieces) {
  CtxLock lk(ctx);
  int st = check_ctx(ctx);
  if (st) return st;
  if (!pieces) return fail(CKT_ERR_INVALID_ARG, "null piece array");
  const Circuit* c = ctx->eng.find(ctx->id);
  for (uint32_t i = 0; i < c->num_pieces; i++) pieces[i] = c->pieces[i];
  return CKT_OK;
}

int ckt_calc_new_currents(ckt_ctx_t* ctx, const ckt_piece_t* pieces, uint32_t n, int policy) {
  CtxLock lk(ctx);
  int st = check_ctx(ctx);
  if (st) return st;
  std::vector<PieceRef> refs;
  if ((st = refs_from(ctx, pieces, n, refs))) return st;
  return ctx->eng.cnc(refs, policy);
}

int ckt_distribute_charge(ckt_ctx_t* ctx, const ckt_piece_t* pieces, uint32_t n, int mode) {
  CtxLock lk(ctx);
  int st = check_ctx(ctx);
  if (st) return st;
  std::vector<PieceRef> refs;
  if ((st = refs_from(ctx, pieces, n, refs))) return st;
  return ctx->eng.dsc(refs, mode);
}

int ckt_update_voltages(ckt_ctx_t* ctx, const ckt_piece_t* pieces, uint32_t n) {
  CtxLock lk(ctx);
  int st = check_ctx(ctx);
  if (st) return st;
  std::vector<PieceRef> refs;
  if ((st = refs_from(ctx, pieces, n, refs))) return st;
  return ctx->eng.upv(refs);
}

int ckt_sanity_check_wires(ckt_ctx_t* ctx, const ckt_piece_t* pieces, uint32_t n,
                           uint64_t* num_bad) {
  CtxLock lk(ctx);
  int st = check_ctx(ctx);
  if (st) return st;
  std::vector<PieceRef> refs;
  if ((st = refs_from(ctx, pieces, n, refs))) return st;
  return ctx->eng.sanity(refs, num_bad);
}

int ckt_execute_index_space(ckt_ctx_t* ctx, int task_id, const ckt_piece_t* pieces, uint32_t n,
                            const ckt_task_config_t* config) {
  CtxLock lk(ctx);
  const ckt_task_config_t def = {CKT_POLICY_EXACT, CKT_REDUCE_ATOMIC};
  const ckt_task_config_t* cfg = config ? config : &def;
  switch (task_id) {
    case CKT_CALC_NEW_CURRENTS: return ckt_calc_new_currents(ctx, pieces, n, cfg->policy);
    case CKT_DISTRIBUTE_CHARGE: return ckt_distribute_charge(ctx, pieces, n, cfg->reduce_mode);
    case CKT_UPDATE_VOLTAGES: return ckt_update_voltages(ctx, pieces, n);
    case CKT_CALC_NEW_CURRENTS_NOP:
    case CKT_DISTRIBUTE_CHARGE_NOP:
    case CKT_UPDATE_VOLTAGES_NOP: {
      int st = check_ctx(ctx);
      if (st) return st;
      return ctx->eng.nop();
    }
    case CKT_SANITY_CHECK_WIRES: {
      uint64_t bad = 0;
      int st = ckt_sanity_check_wires(ctx, pieces, n, &bad);
      if (st) return st;
      if (bad)
        return fail(CKT_ERR_INVALID_ARG, std::to_string(bad) + " wires fail the sanity check");
      return CKT_OK;
    }
    case CKT_REGION_MAIN:
      return fail(CKT_ERR_UNSUPPORTED, "REGION_MAIN is the host driver (ckt_run_loops)");
    default: return fail(CKT_ERR_INVALID_ARG, "unknown task id " + std::to_string(task_id));
  }
}

int ckt_run_loops(ckt_ctx_t* ctx, int num_loops, const ckt_loop_config_t* config,
                  ckt_loop_report_t* report) {
  CtxLock lk(ctx);
  int st = check_ctx(ctx);
  if (st) return st;
  if (num_loops < 0) return fail(CKT_ERR_INVALID_ARG, "negative loop count");
  const ckt_loop_config_t def = {CKT_POLICY_EXACT, CKT_REDUCE_ATOMIC, 0, 0, 0};
  const ckt_loop_config_t& cfg = config ? *config : def;
  if (cfg.sync < 0 || cfg.sync > 2) return fail(CKT_ERR_INVALID_ARG, "sync must be 0, 1 or 2");
  Engine& E = ctx->eng;
  const Circuit* c = E.find(ctx->id);
  std::vector<PieceRef> all;
  E.all_refs(all, false);
  double ms0[3] = {0, 0, 0};
  for (int k = 0; k < 3; k++) E.kernel_time(CKT_CALC_NEW_CURRENTS + k, &ms0[k], nullptr);

  // Each step of the loop as the reference orders it (circuit.cc:232-307).
  auto phase = [&](int nop_task, int task) -> int {
    int rc;
    if (cfg.sync) {
      if ((rc = E.nop())) return rc;
      if (cfg.sync == 1 && (rc = E.sync())) return rc;
    }
    // The CSR reductions sum per node over the whole index space, so in
    // per-piece mode that phase is still one launch.
    const bool whole = task == CKT_DISTRIBUTE_CHARGE && cfg.reduce_mode != CKT_REDUCE_ATOMIC;
    if (cfg.per_piece && !whole) {
      for (const PieceRef& r : all) {
        std::vector<PieceRef> one(1, r);
        if (task == CKT_CALC_NEW_CURRENTS) rc = E.cnc(one, cfg.policy);
        else if (task == CKT_DISTRIBUTE_CHARGE) rc = E.dsc(one, cfg.reduce_mode);
        else rc = E.upv(one);
        if (rc) return rc;
        if ((rc = E.sync())) return rc;
      }
    } else {
      if (task == CKT_CALC_NEW_CURRENTS) rc = E.cnc(all, cfg.policy);
      else if (task == CKT_DISTRIBUTE_CHARGE) rc = E.dsc(all, cfg.reduce_mode);
      else rc = E.upv(all);
      if (rc) return rc;
    }
    (void)nop_task;
    return CKT_OK;
  };

  if ((st = E.sync())) return st;
  const double t0 = now_s();
  if (cfg.use_graph && !cfg.sync && !cfg.per_piece) {
    if ((st = E.run_graph(all, cfg.policy, cfg.reduce_mode, num_loops))) return st;
  } else if (!cfg.sync && !cfg.per_piece) {
    for (int i = 0; i < num_loops; i++)
      if ((st = E.loop_once(all, cfg.policy, cfg.reduce_mode))) return st;
  } else {
    for (int i = 0; i < num_loops; i++) {
      if ((st = phase(CKT_CALC_NEW_CURRENTS_NOP, CKT_CALC_NEW_CURRENTS))) return st;
      if (cfg.sync == 1 && (st = E.sync())) return st;
      if ((st = phase(CKT_DISTRIBUTE_CHARGE_NOP, CKT_DISTRIBUTE_CHARGE))) return st;
      if (cfg.sync == 1 && (st = E.sync())) return st;
      if ((st = phase(CKT_UPDATE_VOLTAGES_NOP, CKT_UPDATE_VOLTAGES))) return st;
      if (i != num_loops - 1 && cfg.sync == 1 && (st = E.sync())) return st;
    }
  }
  if ((st = E.sync())) return st;
  const double elapsed = now_s() - t0;
  if (report) {
    report->elapsed_s = elapsed;
    // circuit.cc:320-334, with the configured inner step count.
    const double steps = (double)E.option(OPT_STEPS);
    double ops = (double)c->num_wires * (kSeg * 6 + (kSeg - 1) * 4) * steps;
    ops += (double)c->num_wires * 4;
    ops += (double)c->num_nodes * 4;
    ops *= num_loops;
    report->gflops = elapsed > 0 ? 1e-9 * ops / elapsed : 0.0;
    report->wire_updates_per_s = elapsed > 0 ? (double)c->num_wires * num_loops / elapsed : 0.0;
    for (int k = 0; k < 3; k++) {
      double ms = 0;
      E.kernel_time(CKT_CALC_NEW_CURRENTS + k, &ms, nullptr);
      report->kernel_ms[k] = ms - ms0[k];
    }
  }
  return CKT_OK;
}

int ckt_ctx_set_option(ckt_ctx_t* ctx, int option, int64_t value) {
  CtxLock lk(ctx);
  if (!ctx) return fail(CKT_ERR_INVALID_ARG, "null context");
  return ctx->eng.set_option(option, value);
}

int ckt_ctx_set_timing(ckt_ctx_t* ctx, int enable) {
  CtxLock lk(ctx);
  if (!ctx) return fail(CKT_ERR_INVALID_ARG, "null context");
  ctx->eng.set_timing(enable != 0);
  return CKT_OK;
}

int ckt_ctx_kernel_time(ckt_ctx_t* ctx, int task_id, double* total_ms, uint64_t* launches) {
  CtxLock lk(ctx);
  if (!ctx) return fail(CKT_ERR_INVALID_ARG, "null context");
  if (task_id < 0 || task_id >= CKT_NUM_TASK_IDS) return fail(CKT_ERR_INVALID_ARG, "bad task id");
  return ctx->eng.kernel_time(task_id, total_ms, launches);
}

int ckt_ctx_reset_timing(ckt_ctx_t* ctx) {
  CtxLock lk(ctx);
  if (!ctx) return fail(CKT_ERR_INVALID_ARG, "null context");
  ctx->eng.reset_timing();
  return CKT_OK;
}

}  // extern "C"
