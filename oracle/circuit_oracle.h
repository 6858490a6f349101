/*
 * circuit_oracle.h -- CPU restatement of the Legion circuit hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle the GPU results are
 * checked against: tests, the smoke entry point and the CPU-baseline leg of
 * the benchmark load it.  The product
 * (legionorigins_amd) never links, loads or calls it.
 *
 * Each function cites the reference lines it restates (paths relative to the
 * reference tree).  FP32 functions reproduce the reference's float arithmetic
 * operation by operation; FP64 twins run the same operations in double from
 * the same float inputs, with dt = (double)(float)1e-6 so that only the
 * working precision differs.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   generator         POSIX drand48 closed form (independent integer LCG),
 *                     draw-count and classification invariants, hand-checked
 *                     tiny circuit.
 *   calc_new_currents Ohm's law after one step, the backward-Euler RLC ladder
 *                     solution as the fixed point of the relaxation (numpy
 *                     linear solve), the FP32 divergence signature of the
 *                     reference parameters.
 *   distribute_charge charge conservation, numpy scatter-add in float64.
 *   update_voltages   special cases q=0 and leak=0, charge cleared.
 */
#ifndef CIRCUIT_ORACLE_H
#define CIRCUIT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_WIRE_SEGMENTS 10 /* circuit.h:42 */
#define ORC_STEPS 10000      /* circuit.h:43 */
#define ORC_DELTAT 1e-6      /* circuit.h:44 */

enum { ORC_PRIVATE_PTR = 0, ORC_SHARED_PTR = 1, ORC_GHOST_PTR = 2 }; /* circuit.h:13-17 */

/* circuit.h:35-40 */
typedef struct {
  float charge, voltage, capacitance, leakage;
} orc_node_t;

/* circuit.h:46-55 (128 bytes) */
typedef struct {
  uint32_t in_ptr, out_ptr;
  int32_t in_loc, out_loc;
  float inductance, resistance;
  float current[ORC_WIRE_SEGMENTS];
  float capacitance;
  float voltage[ORC_WIRE_SEGMENTS - 1];
  float pad[6];
} orc_wire_t;

typedef struct {
  double charge, voltage, capacitance, leakage;
} orc_node64_t;

typedef struct {
  uint32_t in_ptr, out_ptr;
  int32_t in_loc, out_loc;
  double inductance, resistance;
  double current[ORC_WIRE_SEGMENTS];
  double capacitance;
  double voltage[ORC_WIRE_SEGMENTS - 1];
} orc_wire64_t;

typedef struct {
  uint32_t first_wire, num_wires, first_node, num_nodes;
} orc_piece_t;

/* Seeding of the relaxation's voltage vector.
 * ORC_SEED_GPU        : new_v[0] = in_node.voltage, new_v[1..9] = wire.voltage,
 *                       new_v[10] = out_node.voltage (circuit_gpu.cu:107-112).
 * ORC_SEED_CPU_LITERAL: new_v[0..8] = wire.voltage[0..8], new_v[10] = out
 *                       voltage (circuit.cc:380-384); new_v[9], uninitialised
 *                       in the reference for the first step, is taken as 0. */
enum { ORC_SEED_GPU = 0, ORC_SEED_CPU_LITERAL = 1 };

enum { ORC_PROFILE_REFERENCE = 0, ORC_PROFILE_STABLE = 1 };

int orc_sizeof_node(void);
int orc_sizeof_wire(void);

/* load_circuit, circuit.cc:1071-1288, random_element :1042-1049,
 * find_location :1051-1069.  glibc srand48/drand48 are called directly.
 * first_wire/first_node receive num_pieces entries; *draws the number of
 * drand48 calls.  Returns 0, or -1 on invalid sizes. */
int orc_generate(int num_pieces, int nodes_per_piece, int wires_per_piece, int pct_wire_in_piece,
                 int random_seed, int profile, orc_node_t* nodes, orc_wire_t* wires,
                 int32_t* locator, orc_piece_t* pieces, uint64_t* draws);

/* glibc drand48 after srand48(seed): `count` values (known-answer tests). */
void orc_drand48_stream(int seed, double* out, int count);

/* calculate_currents_task, circuit.cc:352-413 (seeding per ORC_SEED_*; the
 * GPU variant circuit_gpu.cu:61-149 when seeding == ORC_SEED_GPU).  Node
 * reads go through get_node (circuit.cc:80-96): pvt/shr/ghost instances
 * indexed by absolute pointer.  Returns the number of wires with a bad
 * location (0 normally). */
int orc_calc_new_currents_f32(int seeding, int steps, orc_wire_t* wires, const orc_node_t* pvt,
                              const orc_node_t* shr, const orc_node_t* ghost,
                              uint32_t first_wire, uint32_t num_wires);
int orc_calc_new_currents_f64(int seeding, int steps, orc_wire64_t* wires,
                              const orc_node64_t* nodes, uint32_t first_wire, uint32_t num_wires);

/* distribute_charge_task, circuit.cc:415-443 with reduce_node :98-116 and
 * AccumulateCharge::apply<true> (:32-36): sequential in wire order, in-end
 * before out-end.  shr/ghost are node instances reduced in place. */
int orc_distribute_charge_f32(const orc_wire_t* wires, orc_node_t* pvt, orc_node_t* shr,
                              orc_node_t* ghost, uint32_t first_wire, uint32_t num_wires);
/* Same, with shared/ghost contributions folded into float reduction buffers
 * (identity 0.0f, circuit.cc:30; fold<true> :51-55), as the GPU variant's
 * AccessorGPUReductionFold instances receive them (circuit_gpu.cu:219-243). */
int orc_distribute_charge_fold_f32(const orc_wire_t* wires, orc_node_t* pvt, float* shr_fold,
                                   float* ghost_fold, uint32_t first_wire, uint32_t num_wires);
/* Apply a fold buffer to the owner instance (apply<true>, circuit.cc:32-36;
 * lowlevel_dma.cc:324-358). */
void orc_fold_apply_f32(orc_node_t* owner, const float* fold, uint32_t first, uint32_t count);
int orc_distribute_charge_f64(const orc_wire64_t* wires, orc_node64_t* nodes, uint32_t first_wire,
                              uint32_t num_wires);
/* All pieces in global wire order; shared nodes (locator == 0) accumulate in
 * double with Kahan compensation and are rounded once, private nodes as
 * orc_distribute_charge_f32.  This is the build's CSR_KAHAN64 mode. */
void orc_distribute_charge_kahan64(const orc_wire_t* wires, uint32_t num_wires, orc_node_t* nodes,
                                   const int32_t* locator, uint32_t num_nodes);

/* update_voltages_task + update_region_voltages, circuit.cc:445-460,1290-1307;
 * locator dispatch as circuit_gpu.cu:325-345. */
void orc_update_voltages_f32(orc_node_t* pvt, orc_node_t* shr, const int32_t* locator,
                             uint32_t first_node, uint32_t num_nodes);
void orc_update_voltages_f64(orc_node64_t* nodes, uint32_t first_node, uint32_t num_nodes);

/* The region_main loop (circuit.cc:232-307): per loop, calc_new_currents over
 * every piece, then distribute_charge over every piece in piece order, then
 * update_voltages over every piece.  calc_new_currents runs on `nthreads`
 * threads; grain 0 hands out whole pieces (one piece per task, as the
 * reference's SharedMapper, circuit.cc:623-635), grain > 0 hands out chunks of
 * `grain` wires.  reduce: 0 sequential FP32, 1 Kahan FP64 for shared nodes. */
int orc_run_f32(int seeding, int steps, int loops, int reduce, int nthreads, int grain,
                orc_node_t* nodes, uint32_t num_nodes, orc_wire_t* wires, uint32_t num_wires,
                const int32_t* locator, const orc_piece_t* pieces, uint32_t num_pieces);
int orc_run_f64(int seeding, int steps, int loops, int nthreads, orc_node64_t* nodes,
                uint32_t num_nodes, orc_wire64_t* wires, uint32_t num_wires,
                const orc_piece_t* pieces, uint32_t num_pieces);

/* FP64 limit of the relaxation.  The 10000-step iteration of
 * calculate_currents_task is a fixed-point iteration whose fixed point is the
 * backward-Euler step of the RLC ladder:
 *   KVL per segment i:      L (I_i - I0_i)/dt + R I_i = v_{i+1} - v_i
 *   KCL per internal node k: C (v_k - V0_k)/dt = I_{k-1} - I_k
 * (v_0, v_10 the in/out node voltages).  In the stable profile it contracts
 * by < 0.12 per step, so 10000 steps reach the fixed point to the last bit of
 * FP64.  This solves the 19x19 system per wire (Gaussian elimination with
 * partial pivoting) instead of iterating; only meaningful when the
 * relaxation converges.  The loop driver is orc_run_f64's. */
int orc_calc_new_currents_f64_limit(orc_wire64_t* wires, const orc_node64_t* nodes,
                                    uint32_t first_wire, uint32_t num_wires);
int orc_run_f64_limit(int loops, int nthreads, orc_node64_t* nodes, uint32_t num_nodes,
                      orc_wire64_t* wires, uint32_t num_wires, const orc_piece_t* pieces,
                      uint32_t num_pieces);

/* calc_new_currents over an explicit list of wire indices on `nthreads`
 * threads (CPU-baseline sample).  Returns wall seconds. */
double orc_cnc_sample_f32(int seeding, int steps, int nthreads, orc_wire_t* wires,
                          const orc_node_t* nodes, const uint32_t* wire_idx, uint32_t count);

void orc_promote(const orc_node_t* n32, uint32_t num_nodes, const orc_wire_t* w32,
                 uint32_t num_wires, orc_node64_t* n64, orc_wire64_t* w64);

#ifdef __cplusplus
}
#endif

#endif
