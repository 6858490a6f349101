// ckt_main.cpp -- host driver: the region_main of apps/circuit/circuit.cc
// (:123-346) over the C ABI.  Same command line (parse_input_args,
// circuit.cc:991-1040) and the same report lines (:144-146, :165, :314-334).
//
//   ckt_sim [-l loops] [-p pieces] [-npp nodes/piece] [-wpp wires/piece]
//           [-pct pct_in_piece] [-s seed] [-sync 0|1|2]
//           [-policy exact|recip|fast|scaled] [-mode atomic|csr|kahan64]
//           [-profile reference|stable] [-graph] [-per_piece] [-device d]
//           [-devgen]   (generate the circuit in HBM instead of on the host)
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/circuit_abi.h"

static int check(int st, const char* what) {
  if (st != CKT_OK) {
    std::fprintf(stderr, "%s failed: %s: %s\n", what, ckt_status_string(st), ckt_last_error());
    std::exit(1);
  }
  return st;
}

int main(int argc, char** argv) {
  // Defaults of region_main (circuit.cc:128-134).
  int num_loops = 2, num_pieces = 4, nodes_per_piece = 2, wires_per_piece = 4;
  int pct_wire_in_piece = 95, random_seed = 12345, sync = 0;
  int policy = CKT_POLICY_EXACT, mode = CKT_REDUCE_ATOMIC, profile = CKT_PROFILE_REFERENCE;
  int graph = 0, per_piece = 0, device = 0, devgen = 0;
  for (int i = 1; i < argc; i++) {
    auto next_int = [&](int& v) {
      if (i + 1 < argc) v = std::atoi(argv[++i]);
    };
    auto next_str = [&]() -> const char* { return i + 1 < argc ? argv[++i] : ""; };
    if (!std::strcmp(argv[i], "-l")) next_int(num_loops);
    else if (!std::strcmp(argv[i], "-p")) next_int(num_pieces);
    else if (!std::strcmp(argv[i], "-npp")) next_int(nodes_per_piece);
    else if (!std::strcmp(argv[i], "-wpp")) next_int(wires_per_piece);
    else if (!std::strcmp(argv[i], "-pct")) next_int(pct_wire_in_piece);
    else if (!std::strcmp(argv[i], "-s")) next_int(random_seed);
    else if (!std::strcmp(argv[i], "-sync")) next_int(sync);
    else if (!std::strcmp(argv[i], "-device")) next_int(device);
    else if (!std::strcmp(argv[i], "-graph")) graph = 1;
    else if (!std::strcmp(argv[i], "-per_piece")) per_piece = 1;
    else if (!std::strcmp(argv[i], "-devgen")) devgen = 1;
    else if (!std::strcmp(argv[i], "-policy")) {
      const char* s = next_str();
      policy = !std::strcmp(s, "scaled") ? CKT_POLICY_SCALED
               : !std::strcmp(s, "fast")   ? CKT_POLICY_FAST
               : !std::strcmp(s, "recip")  ? CKT_POLICY_RECIP
                                           : CKT_POLICY_EXACT;
    } else if (!std::strcmp(argv[i], "-mode")) {
      const char* s = next_str();
      mode = !std::strcmp(s, "csr") ? CKT_REDUCE_CSR
             : !std::strcmp(s, "kahan64") ? CKT_REDUCE_CSR_KAHAN64 : CKT_REDUCE_ATOMIC;
    } else if (!std::strcmp(argv[i], "-profile")) {
      profile = !std::strcmp(next_str(), "stable") ? CKT_PROFILE_STABLE : CKT_PROFILE_REFERENCE;
    }
  }
  std::printf("circuit settings: loops=%d pieces=%d nodes/piece=%d wires/piece=%d "
              "pct_in_piece=%d seed=%d\n",
              num_loops, num_pieces, nodes_per_piece, wires_per_piece, pct_wire_in_piece,
              random_seed);
  ckt_ctx_t* ctx = nullptr;
  check(ckt_ctx_create(device, &ctx), "ckt_ctx_create");
  ckt_gen_params_t gp = {num_pieces, nodes_per_piece, wires_per_piece, pct_wire_in_piece,
                         random_seed, profile};
  check(devgen ? ckt_ctx_generate_device(ctx, &gp) : ckt_ctx_generate(ctx, &gp), "load_circuit");
  check(ckt_ctx_set_timing(ctx, 1), "timing");
  std::printf("Starting main simulation loop\n");
  ckt_loop_config_t cfg = {policy, mode, sync, graph, per_piece};
  ckt_loop_report_t rep;
  check(ckt_run_loops(ctx, num_loops, &cfg, &rep), "simulation loop");
  std::printf("SUCCESS!\n");
  std::printf("ELAPSED TIME = %7.3f s\n", rep.elapsed_s);
  std::printf("GFLOPS = %7.3f GFLOPS\n", rep.gflops);
  // Per-task kernel time, in the spirit of DetailedTimer::report_timers.
  const char* names[3] = {"calc_new_currents", "distribute_charge", "update_voltages"};
  for (int k = 0; k < 3; k++) {
    double ms = 0;
    uint64_t n = 0;
    ckt_ctx_kernel_time(ctx, CKT_CALC_NEW_CURRENTS + k, &ms, &n);
    std::printf("KERNEL %-18s %10.3f ms total, %6llu launches, %10.3f ms avg\n", names[k], ms,
                (unsigned long long)n, n ? ms / (double)n : 0.0);
  }
  std::printf("WIRE-UPDATES/S = %.4g\n", rep.wire_updates_per_s);
  ckt_ctx_destroy(ctx);
  return 0;
}
