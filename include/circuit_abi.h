/*
 * circuit_abi.h -- C ABI of the MI355X (gfx950) circuit hot path.
 *
 * The three per-piece tasks of the Legion circuit application
 *   calc_new_currents   (apps/circuit/circuit.cc:352-413, circuit_gpu.cu:61-149,186-205)
 *   distribute_charge   (apps/circuit/circuit.cc:415-443, circuit_gpu.cu:219-304)
 *   update_voltages     (apps/circuit/circuit.cc:445-460,1290-1307, circuit_gpu.cu:306-368)
 * behind the reference's task IDs, region field layouts and piece descriptor.
 *
 * Conventions
 *   - Plain C types only; every function returns an int status (CKT_OK == 0)
 *     and never aborts.  ckt_last_error() gives a thread-local message.
 *   - Node and wire indices are absolute element indices, as ptr_t::value in
 *     the reference (runtime/common.h:12-164): piece p owns wires
 *     [first_wire, first_wire+num_wires) and nodes [first_node, first_node+num_nodes).
 *   - "ctx" calls are stream-ordered on the context's own HIP stream and return
 *     before the GPU finishes; ckt_ctx_synchronize() waits.  The accessor-form
 *     calls (suffix _gpu) wait for completion before returning, as the
 *     reference's host launchers do (circuit_gpu.cu:204,303,367).
 *   - A stream argument is a hipStream_t passed as void* (NULL = null stream).
 *   - Threading: calls on one context or batch handle may come from several
 *     host threads; the handle's mutex runs them one at a time (the reference
 *     runs a GPU's tasks one at a time on one worker thread,
 *     lowlevel_gpu.cc:117-160).  A server's submit, poll, wait and cancel may
 *     be called from any thread while its scheduler runs.  Distinct handles are
 *     independent and may be used concurrently.  Destroying a handle must not
 *     overlap other calls on it.  The accessor-form calls and the generator are
 *     re-entrant.
 */
#ifndef CIRCUIT_ABI_H
#define CIRCUIT_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CKT_ABI_VERSION 1

/* Physics constants, apps/circuit/circuit.h:42-44. */
#define CKT_WIRE_SEGMENTS 10
#define CKT_STEPS 10000
#define CKT_DELTAT 1e-6

/* PointerLocation, apps/circuit/circuit.h:13-17. */
enum { CKT_PRIVATE_PTR = 0, CKT_SHARED_PTR = 1, CKT_GHOST_PTR = 2 };

/* Task IDs, apps/circuit/circuit.h:19-28, and the reduction op ID, :30-32. */
enum {
  CKT_REGION_MAIN = 0,
  CKT_CALC_NEW_CURRENTS = 1,
  CKT_DISTRIBUTE_CHARGE = 2,
  CKT_UPDATE_VOLTAGES = 3,
  CKT_CALC_NEW_CURRENTS_NOP = 4,
  CKT_DISTRIBUTE_CHARGE_NOP = 5,
  CKT_UPDATE_VOLTAGES_NOP = 6,
  CKT_SANITY_CHECK_WIRES = 7,
  CKT_NUM_TASK_IDS = 8
};
enum { CKT_REDUCE_ID = 1 };

/* CircuitNode, apps/circuit/circuit.h:35-40 (16 bytes). */
typedef struct ckt_node {
  float charge;
  float voltage;
  float capacitance;
  float leakage;
} ckt_node_t;

/* CircuitWire, apps/circuit/circuit.h:46-55 (128 bytes, including the six
 * padding floats crap..lastcrap).  in_ptr/out_ptr are ptr_t<CircuitNode>
 * values; in_loc/out_loc hold CKT_*_PTR. */
typedef struct ckt_wire {
  uint32_t in_ptr, out_ptr;
  int32_t in_loc, out_loc;
  float inductance;
  float resistance;
  float current[CKT_WIRE_SEGMENTS];
  float capacitance;
  float voltage[CKT_WIRE_SEGMENTS - 1];
  float pad[6];
} ckt_wire_t;

/* The fields of CircuitPiece (apps/circuit/circuit.h:63-70) that the GPU
 * variants read (circuit_gpu.cu:195-203,360-365). */
typedef struct ckt_piece {
  uint32_t first_wire;
  uint32_t num_wires;
  uint32_t first_node;
  uint32_t num_nodes;
} ckt_piece_t;

/* Status codes. */
enum {
  CKT_OK = 0,
  CKT_ERR_INVALID_ARG = 1,   /* null pointer, bad enum, inconsistent sizes */
  CKT_ERR_OUT_OF_RANGE = 2,  /* piece or index outside the loaded circuit */
  CKT_ERR_NO_MEMORY = 3,     /* device or host allocation failed */
  CKT_ERR_BUDGET = 4,        /* request exceeds the configured HBM budget */
  CKT_ERR_HIP = 5,           /* a HIP runtime call failed */
  CKT_ERR_NO_DEVICE = 6,     /* no usable gfx950 device */
  CKT_ERR_STATE = 7,         /* call not valid in the object's current state */
  CKT_ERR_UNSUPPORTED = 8,
  CKT_ERR_NOT_FOUND = 9,     /* unknown circuit / request id */
  CKT_ERR_CANCELLED = 10,
  CKT_ERR_TIMEOUT = 11
};

/* Arithmetic policy of calc_new_currents.
 *   EXACT : the reference kernel's operation order with correctly rounded
 *           division (circuit_gpu.cu:114-129); bit-identical to a strict FP32
 *           CPU run.
 *   RECIP : per-wire reciprocals (1/R, L/dt, dt/C) hoisted out of the step
 *           loop and fused multiply-adds; voltages materialised every step.
 *   FAST  : RECIP with the internal-voltage update substituted into the next
 *           step's current update (a tridiagonal update of the 10 currents per
 *           step: 20 FMAs and 8 adds).  Every step's currents are computed; internal
 *           voltages are formed from the final currents.
 *   SCALED: FAST's update iterated as the deviation from its fixed point (one
 *           tridiagonal solve per wire and loop), scaled by the coupling
 *           coefficient and split into its mirror-symmetric and
 *           antisymmetric halves (segment i against segment 9 - i), so that
 *           each step is 6 adds and 10 FMAs, plus 10 multiplies per 4 steps.
 *           Used for wires shown to contract or to diverge; other wires take
 *           FAST's step.
 * RECIP, FAST and SCALED are not bit-exact; their error is checked against
 * 1e-5.  FAST and SCALED write a wire whose currents left the finite range as
 * all NaN, which is where the reference recurrence ends such a wire. */
enum { CKT_POLICY_EXACT = 0, CKT_POLICY_RECIP = 1, CKT_POLICY_FAST = 2, CKT_POLICY_SCALED = 3 };

/* Reduction mode of distribute_charge.
 *   ATOMIC      : float atomics (LDS-privatised per piece when the piece's
 *                 node range fits in LDS), order nondeterministic.
 *   CSR         : deterministic per-node sum over a node->(wire,side) index in
 *                 global (wire, in-before-out) order; bit-identical to the
 *                 sequential CPU order of the reference task bodies.
 *   CSR_KAHAN64 : CSR, with shared nodes accumulated in FP64 with Kahan
 *                 compensation and rounded once to FP32. */
enum { CKT_REDUCE_ATOMIC = 0, CKT_REDUCE_CSR = 1, CKT_REDUCE_CSR_KAHAN64 = 2 };

/* Parameter profiles of the generator.
 *   REFERENCE : load_circuit's distributions (circuit.cc:1122-1164).
 *   STABLE    : same draws and topology, inductance scaled by 1e-6 and wire
 *               capacitance U(0.001, 0.011), so the relaxation converges. */
enum { CKT_PROFILE_REFERENCE = 0, CKT_PROFILE_STABLE = 1 };

/* Arguments of load_circuit (circuit.cc:1071-1073; CLI circuit.cc:991-1040). */
typedef struct ckt_gen_params {
  int32_t num_pieces;
  int32_t nodes_per_piece;
  int32_t wires_per_piece;
  int32_t pct_wire_in_piece;
  int32_t random_seed;
  int32_t profile; /* CKT_PROFILE_* */
} ckt_gen_params_t;

/* ------------------------------------------------------------------------ */
/* Library / device utilities                                                */

int ckt_abi_version(void);
const char* ckt_status_string(int status);
const char* ckt_last_error(void);
/* Number of visible devices and whether device `device` is gfx950. */
int ckt_device_info(int device, int* is_gfx950, int* num_cus, size_t* lds_per_block,
                    size_t* hbm_bytes);
/* Plain device allocations for hosts without their own allocator (the
 * reference's runtime owns instances; lowlevel_gpu.cc:91-99). */
int ckt_device_malloc(int device, void** ptr, size_t bytes);
int ckt_device_free(void* ptr);
int ckt_memcpy_htod(void* dst, const void* src, size_t bytes);
int ckt_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int ckt_device_synchronize(int device);

/* ------------------------------------------------------------------------ */
/* Generator: load_circuit (circuit.cc:1071-1288) with O(1) node choice.      */

/* Sizes the caller must provide for ckt_generate. */
int ckt_generate_sizes(const ckt_gen_params_t* params, uint32_t* num_nodes,
                       uint32_t* num_wires, uint32_t* num_pieces);
/* Fills host arrays in the reference layouts: nodes[num_nodes], wires[num_wires],
 * locator[num_nodes] (1 = private, 0 = shared, circuit.cc:1215,1219) and
 * pieces[num_pieces]. `locator` and `pieces` may be NULL. */
int ckt_generate(const ckt_gen_params_t* params, ckt_node_t* nodes, ckt_wire_t* wires,
                 int32_t* locator, ckt_piece_t* pieces);

/* ------------------------------------------------------------------------ */
/* Accessor-form drop-ins: the entry points of circuit.h:83-104.             */
/* Arguments are device pointers to instances in the reference layouts,      */
/* indexed by absolute element index (GPU_Accessor::ref, lowlevel.h:749-833).*/
/* `flag` is the reference's unused trailing argument; here it selects the   */
/* policy (calc_new_currents) or must be 0.  Each call waits for completion. */

/* calc_new_currents_gpu, circuit.h:83-88 / circuit_gpu.cu:186-205. */
int ckt_calc_new_currents_gpu(const ckt_piece_t* piece, ckt_wire_t* wires,
                              const ckt_node_t* pvt, const ckt_node_t* owned,
                              const ckt_node_t* ghost, int flag, void* stream);
/* distribute_charge_gpu, circuit.h:90-95 / circuit_gpu.cu:288-304.  Private
 * nodes are reduced into pvt[].charge; shared and ghost contributions are
 * folded into the float reduction buffers owned_fold[] / ghost_fold[]
 * (AccessorGPUReductionFold, lowlevel.h:835-881), indexed absolutely. */
int ckt_distribute_charge_gpu(const ckt_piece_t* piece, const ckt_wire_t* wires,
                              ckt_node_t* pvt, float* owned_fold, float* ghost_fold,
                              int flag, void* stream);
/* update_voltages_gpu, circuit.h:97-101 / circuit_gpu.cu:352-368.  locator[i]
 * != 0 selects pvt[i], else owned[i]. */
int ckt_update_voltages_gpu(const ckt_piece_t* piece, ckt_node_t* pvt, ckt_node_t* owned,
                            const int32_t* locator, int flag, void* stream);
/* sanity_check_wires_gpu, circuit.h:103-104 / circuit_gpu.cu:151-217, without
 * the hard-coded 10000/2500 sizes: counts wires whose in/out location is not
 * a PointerLocation, whose in_ptr is outside the piece's node range, or whose
 * out_ptr is >= num_nodes_total.  *num_bad receives the count. */
int ckt_sanity_check_wires_gpu(const ckt_piece_t* piece, const ckt_wire_t* wires,
                               uint32_t num_nodes_total, uint64_t* num_bad, void* stream);
/* The three tasks above over a table of `num_pieces` pieces (host memory) in
 * one launch each: an index-space launch of the point tasks
 * (circuit.cc:250-251,276-277,298-299) on the same instances, with the
 * results of one call per piece in piece order (distribute_charge's float
 * atomics aside, whose order is unspecified in either form) and one wait at
 * the end.  The per-piece calls are this form with a table of one.  Each
 * call allocates a few bytes of device scratch per piece (the piece table and
 * a bad-record counter) from the stream-ordered allocator on `stream` and
 * releases it before returning. */
int ckt_calc_new_currents_index_gpu(const ckt_piece_t* pieces, uint32_t num_pieces,
                                    ckt_wire_t* wires, const ckt_node_t* pvt,
                                    const ckt_node_t* owned, const ckt_node_t* ghost, int flag,
                                    void* stream);
int ckt_distribute_charge_index_gpu(const ckt_piece_t* pieces, uint32_t num_pieces,
                                    const ckt_wire_t* wires, ckt_node_t* pvt, float* owned_fold,
                                    float* ghost_fold, int flag, void* stream);
int ckt_update_voltages_index_gpu(const ckt_piece_t* pieces, uint32_t num_pieces, ckt_node_t* pvt,
                                  ckt_node_t* owned, const int32_t* locator, int flag,
                                  void* stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): ckt_calc_new_currents_index_gpu without any allocation.
 * Its scratch (the piece table and the bad-record counter) is a static device
 * array of CKT_NOALLOC_SCRATCH_BYTES, one per device, held until the process
 * ends; a table of more than CKT_NOALLOC_MAX_PIECES pieces is rejected with
 * CKT_ERR_OUT_OF_RANGE.  Every call on a device shares that device's array
 * (each call rewrites it, counter included, before its launch): calls that
 * overlap on one device, from several host threads or streams, race, and the
 * caller must not overlap them. */
#define CKT_NOALLOC_MAX_PIECES 4096
#define CKT_NOALLOC_SCRATCH_BYTES 81952
int ckt_calc_new_currents_index_gpu_noalloc(const ckt_piece_t* pieces, uint32_t num_pieces,
                                            ckt_wire_t* wires, const ckt_node_t* pvt,
                                            const ckt_node_t* owned, const ckt_node_t* ghost,
                                            int flag, void* stream);
/* Reduction-instance semantics: fill a fold buffer with AccumulateCharge's
 * identity 0.0f (circuit.cc:30) and apply it to the owner instance with
 * AccumulateCharge::apply<true> (circuit.cc:32-36; lowlevel_dma.cc:324-358). */
int ckt_fold_init_gpu(float* fold, uint32_t first, uint32_t count, void* stream);
int ckt_fold_apply_gpu(ckt_node_t* owner, const float* fold, uint32_t first, uint32_t count,
                       void* stream);

/* ------------------------------------------------------------------------ */
/* Context form: one circuit resident in HBM as structure-of-arrays.         */

typedef struct ckt_ctx ckt_ctx_t;

int ckt_ctx_create(int device, ckt_ctx_t** out);
int ckt_ctx_destroy(ckt_ctx_t* ctx);
void* ckt_ctx_stream(ckt_ctx_t* ctx);
int ckt_ctx_synchronize(ckt_ctx_t* ctx);
/* Upload a circuit given in the reference layouts (host memory) and build the
 * device node->wire index used by the CSR reduction modes. */
int ckt_ctx_load(ckt_ctx_t* ctx, const ckt_node_t* nodes, uint32_t num_nodes,
                 const ckt_wire_t* wires, uint32_t num_wires, const int32_t* locator,
                 const ckt_piece_t* pieces, uint32_t num_pieces);
/* Generate (ckt_generate) and load in one call. */
int ckt_ctx_generate(ckt_ctx_t* ctx, const ckt_gen_params_t* params);
/* Same circuit, bit for bit, generated in HBM (load_circuit,
 * circuit.cc:1071-1288, without the host image or its upload): the host
 * walks the drand48 chain once to find each wire's first draw, the device
 * draws every node and wire, marks shared nodes, derives the locations and
 * sorts the node->wire index. */
int ckt_ctx_generate_device(ckt_ctx_t* ctx, const ckt_gen_params_t* params);
/* Download the current state back into the reference layouts (host memory;
 * either pointer may be NULL).  Waits for the stream. */
int ckt_ctx_store(ckt_ctx_t* ctx, ckt_node_t* nodes, ckt_wire_t* wires);
/* Download the loaded circuit's topology (any pointer may be NULL): the
 * locator (1 private, 0 shared; circuit.cc:1210-1220) as num_nodes ints, and
 * the node->wire index the CSR reduction modes use: csr_row[num_nodes+1]
 * offsets into csr_ent[2*num_wires], entry = wire << 1 | end (0 in, 1 out),
 * each node's entries in (wire, in-before-out) order.  Waits for the
 * stream. */
int ckt_ctx_store_topology(ckt_ctx_t* ctx, int32_t* locator, uint32_t* csr_row,
                           uint32_t* csr_ent);
int ckt_ctx_sizes(ckt_ctx_t* ctx, uint32_t* num_nodes, uint32_t* num_wires,
                  uint32_t* num_pieces);
int ckt_ctx_pieces(ckt_ctx_t* ctx, ckt_piece_t* pieces);

/* The three tasks over `n` pieces of the loaded circuit in one launch each
 * (an index-space launch, circuit.cc:250-251,276-277,298-299). */
int ckt_calc_new_currents(ckt_ctx_t* ctx, const ckt_piece_t* pieces, uint32_t n, int policy);
int ckt_distribute_charge(ckt_ctx_t* ctx, const ckt_piece_t* pieces, uint32_t n, int mode);
int ckt_update_voltages(ckt_ctx_t* ctx, const ckt_piece_t* pieces, uint32_t n);
int ckt_sanity_check_wires(ckt_ctx_t* ctx, const ckt_piece_t* pieces, uint32_t n,
                           uint64_t* num_bad);

/* Task-variant table keyed by the reference task IDs (circuit.cc:577-590).
 * The *_NOP IDs are accepted and do nothing (dummy_task, circuit.cc:462-469). */
typedef struct ckt_task_config {
  int32_t policy;      /* CKT_POLICY_* for CALC_NEW_CURRENTS */
  int32_t reduce_mode; /* CKT_REDUCE_* for DISTRIBUTE_CHARGE */
} ckt_task_config_t;
int ckt_execute_index_space(ckt_ctx_t* ctx, int task_id, const ckt_piece_t* pieces, uint32_t n,
                            const ckt_task_config_t* config);

/* The region_main loop (circuit.cc:232-335) over all pieces. */
typedef struct ckt_loop_config {
  int32_t policy;
  int32_t reduce_mode;
  int32_t sync;      /* -sync: 0 none, 1 wait after every launch, 2 NOP launches only */
  int32_t use_graph; /* capture one loop in a hipGraph and replay it */
  int32_t per_piece; /* 1: one launch per piece per task with a wait after each,
                        as the reference runtime runs GPU point tasks
                        (lowlevel_gpu.cc:117-160); 0: one launch per task */
} ckt_loop_config_t;
typedef struct ckt_loop_report {
  double elapsed_s;          /* ELAPSED TIME (circuit.cc:316-318) */
  double gflops;             /* GFLOPS with the reference's operation count (:320-334) */
  double wire_updates_per_s; /* wires * loops / elapsed */
  double kernel_ms[3];       /* summed event time of the three tasks (if timing is on) */
} ckt_loop_report_t;
int ckt_run_loops(ckt_ctx_t* ctx, int num_loops, const ckt_loop_config_t* config,
                  ckt_loop_report_t* report);

/* Tuning and test options (context and batch). */
enum {
  CKT_OPT_STEPS = 0,             /* inner steps per loop (default CKT_STEPS) */
  CKT_OPT_CNC_BLOCK = 1,         /* calc_new_currents workgroup size: 64/128/256/512;
                                    -1 (default): 64 for EXACT, 256 otherwise */
  CKT_OPT_PERSISTENT = 2,        /* 1: persistent tile-queue calc_new_currents, 0: plain
                                    grid, -1 (default): persistent, except for EXACT with
                                    its workgroup size set */
  CKT_OPT_WAVES_PER_SIMD = 3,    /* persistent grid: waves resident per SIMD (1..8);
                                    -1 (default): 7 for EXACT, 4 otherwise */
  CKT_OPT_DSC_ATOMIC_IMPL = 4,   /* ATOMIC reduction: 0 auto, 1 global atomics, 2 LDS */
  CKT_OPT_FUSE_DSC = 5,          /* 1 (default): in ckt_run_loops (sync 0, per_piece 0),
                                    batches and servers, the ATOMIC distribute_charge is
                                    done by calc_new_currents' epilogue from registers */
  /* calc_new_currents' tile schedule (results do not depend on it): */
  CKT_OPT_WIRES_PER_LANE = 6,    /* a tile holds workgroup size x this (1 default, 2, 4)
                                    wires; each lane relaxes that many one after another */
  CKT_OPT_TILE_SPAN = 7,         /* 0 (default): tiles end at piece ends; 1: adjacent pieces
                                    of a circuit form one run, only its last tile is partial */
  CKT_OPT_TILE_ORDER = 8         /* 0 (default): runs in the order given; 1: longest run
                                    first; 2: interleaved, tile k of every run before tile k+1 */
};
int ckt_ctx_set_option(ckt_ctx_t* ctx, int option, int64_t value);

/* Per-task kernel timing with HIP events on the context stream. */
int ckt_ctx_set_timing(ckt_ctx_t* ctx, int enable);
int ckt_ctx_kernel_time(ckt_ctx_t* ctx, int task_id, double* total_ms, uint64_t* launches);
int ckt_ctx_reset_timing(ckt_ctx_t* ctx);

/* ------------------------------------------------------------------------ */
/* Batching: many independent circuits advanced by shared launches.           */

typedef struct ckt_batch ckt_batch_t;

typedef struct ckt_batch_config {
  int32_t policy;
  int32_t reduce_mode;
  int32_t use_graph;        /* capture each loop's launches in a hipGraph */
  int32_t block_threads;    /* 64..1024, multiple of 64; 0 = autotuned/default */
  int32_t persistent;       /* 1: calc_new_currents as a persistent tile-queue kernel */
  int32_t waves_per_simd;   /* persistent grid size = CUs * 4 * waves_per_simd */
  int32_t wires_per_lane;   /* tile schedule (CKT_OPT_WIRES_PER_LANE, _TILE_SPAN,      */
  int32_t tile_span;        /* _TILE_ORDER); wires_per_lane 0 leaves the three as they */
  int32_t tile_order;       /* are                                                     */
} ckt_batch_config_t;

typedef struct ckt_batch_report {
  double elapsed_s;
  double wire_updates;       /* sum over circuits of wires * loops run */
  double wire_updates_per_s;
  double cnc_ms;             /* event time of calc_new_currents launches */
  int32_t loops_run;         /* number of batch iterations */
} ckt_batch_report_t;

int ckt_batch_create(int device, uint64_t hbm_budget_bytes, ckt_batch_t** out);
int ckt_batch_set_option(ckt_batch_t* b, int option, int64_t value);
int ckt_batch_set_timing(ckt_batch_t* b, int enable);
int ckt_batch_kernel_time(ckt_batch_t* b, int task_id, double* total_ms, uint64_t* launches);
int ckt_batch_destroy(ckt_batch_t* b);
/* Add a circuit that will run `num_loops` loops.  CKT_ERR_BUDGET if its device
 * footprint does not fit in the remaining budget. */
int ckt_batch_add_generated(ckt_batch_t* b, const ckt_gen_params_t* params, int32_t num_loops,
                            uint32_t* id);
/* ckt_batch_add_generated with the circuit generated in HBM (as
 * ckt_ctx_generate_device). */
int ckt_batch_add_generated_device(ckt_batch_t* b, const ckt_gen_params_t* params,
                                   int32_t num_loops, uint32_t* id);
/* Generate `n` circuits on `threads` host threads and add them in order;
 * ids[i] receives the id of params[i].  Stops at the first failure. */
int ckt_batch_add_generated_many(ckt_batch_t* b, const ckt_gen_params_t* params,
                                 const int32_t* num_loops, uint32_t n, int32_t threads,
                                 uint32_t* ids);
int ckt_batch_add(ckt_batch_t* b, const ckt_node_t* nodes, uint32_t num_nodes,
                  const ckt_wire_t* wires, uint32_t num_wires, const int32_t* locator,
                  const ckt_piece_t* pieces, uint32_t num_pieces, int32_t num_loops,
                  uint32_t* id);
int ckt_batch_remove(ckt_batch_t* b, uint32_t id);
int ckt_batch_count(ckt_batch_t* b, uint32_t* num_circuits, uint64_t* bytes_in_use);
/* Device bytes a circuit of this shape occupies. */
int ckt_footprint_bytes(uint32_t num_nodes, uint32_t num_wires, uint32_t num_pieces,
                        uint64_t* bytes);
/* Run every circuit through its remaining loops. */
int ckt_batch_run(ckt_batch_t* b, const ckt_batch_config_t* config, ckt_batch_report_t* report);
/* Run one loop of every circuit that has loops left (one batch iteration). */
int ckt_batch_step(ckt_batch_t* b, const ckt_batch_config_t* config, int32_t* active);
int ckt_batch_loops_left(ckt_batch_t* b, uint32_t id, int32_t* loops_left);
int ckt_batch_store(ckt_batch_t* b, uint32_t id, ckt_node_t* nodes, ckt_wire_t* wires);
/* Time candidate schedules for the current batch and write the fastest into
 * *best: first the launch shape (workgroup size x plain grid or persistent
 * queue x resident waves), then, with that shape, the tile schedule (wires
 * per lane x tile span x tile order).  Only calc_new_currents is timed, with
 * `probe_steps` inner steps per candidate; circuit state is restored
 * afterwards. */
int ckt_batch_autotune(ckt_batch_t* b, int32_t policy, int32_t probe_steps,
                       ckt_batch_config_t* best, double* best_ms);
/* The probe ckt_batch_autotune times, for one configuration: milliseconds of
 * calc_new_currents over every circuit at `probe_steps` inner steps (the
 * faster of two launches); circuit state is restored afterwards. */
int ckt_batch_probe(ckt_batch_t* b, const ckt_batch_config_t* config, int32_t probe_steps,
                    double* ms);

/* ------------------------------------------------------------------------ */
/* Serving: continuous batching of circuit simulations.                      */

typedef struct ckt_server ckt_server_t;

typedef struct ckt_server_config {
  uint64_t hbm_budget_bytes;   /* admission budget for device state */
  uint64_t max_active_wires;   /* cap on wires advanced per iteration (0 = none) */
  int32_t max_active;          /* cap on resident circuits (0 = none) */
  int32_t policy;
  int32_t reduce_mode;
  int32_t gen_threads;         /* host threads walking admitted circuits' generator chains (the
                                  circuits themselves are generated on the device) */
} ckt_server_config_t;

typedef struct ckt_request {
  ckt_gen_params_t gen;
  int32_t num_loops;
  double arrival_s;            /* offset from the start of the trace */
  double timeout_s;            /* > 0: not admitted within this many seconds of
                                  arrival -> dropped with CKT_ERR_TIMEOUT */
} ckt_request_t;

typedef struct ckt_request_result {
  int32_t status;              /* CKT_OK, or why the request failed (CKT_ERR_CANCELLED,
                                  CKT_ERR_TIMEOUT, ...) */
  uint32_t num_wires;
  int32_t num_loops;
  double arrival_s, admit_s, finish_s; /* seconds from trace start */
  double final_voltage_checksum;       /* sum of node voltages at completion */
} ckt_request_result_t;

typedef struct ckt_server_report {
  double elapsed_s;
  double wire_updates;
  double wire_updates_per_s;
  double offered_wire_updates_per_s;
  double latency_p50_s, latency_p99_s, latency_mean_s, latency_max_s;
  int32_t completed, rejected, iterations, max_resident;
  int32_t timed_out;           /* dropped at their admission deadline */
} ckt_server_report_t;

int ckt_server_create(int device, const ckt_server_config_t* config, ckt_server_t** out);
int ckt_server_destroy(ckt_server_t* s);
/* Online serving.  ckt_server_start launches the scheduler on a host thread
 * of its own; requests then arrive by ckt_server_submit, from any host thread
 * at any time, and are queued, admitted when the budget and caps allow,
 * generated on the device and advanced by continuous batching.  Times in
 * results are seconds since the start.  ckt_server_stop(s, drain) ends the
 * run: with drain != 0 after every submitted request has finished, otherwise
 * after cancelling the unfinished ones.  submit, poll, wait and cancel may be
 * called concurrently with each other and with the scheduler. */
int ckt_server_start(ckt_server_t* s);
int ckt_server_stop(ckt_server_t* s, int drain);
/* Queue a request; *id numbers requests from 0 in submission order since the
 * start.  A request that can never run (invalid sizes, a negative loop count,
 * a footprint above the budget) is refused with the status saying why.
 * timeout_s > 0: not admitted within that many seconds of its submission ->
 * it ends with CKT_ERR_TIMEOUT. */
int ckt_server_submit(ckt_server_t* s, const ckt_gen_params_t* gen, int32_t num_loops,
                      double timeout_s, uint32_t* id);
/* *finished = 1 and *result filled once request `id` has ended (result->status
 * says how); *finished = 0 while it is queued or running. */
int ckt_server_poll(ckt_server_t* s, uint32_t id, int* finished, ckt_request_result_t* result);
/* Wait for request `id` to end (timeout_s < 0: without limit); CKT_ERR_TIMEOUT
 * if it has not ended within timeout_s. */
int ckt_server_wait(ckt_server_t* s, uint32_t id, double timeout_s, ckt_request_result_t* result);
/* Cancel request `id`: a queued request leaves the queue, a running one leaves
 * at the next iteration boundary; either ends with CKT_ERR_CANCELLED.  A
 * request whose last loop has already run when the scheduler sees the cancel
 * ends normally.  CKT_ERR_STATE if it had already ended. */
int ckt_server_cancel(ckt_server_t* s, uint32_t id);
/* Replay a request trace in real time on the same scheduler (started and
 * stopped by this call): each request is submitted at its arrival offset,
 * with that offset as its arrival time.  results[i] corresponds to reqs[i]. */
int ckt_server_run_trace(ckt_server_t* s, const ckt_request_t* reqs, uint32_t n,
                         ckt_request_result_t* results, ckt_server_report_t* report);
/* Keep each finished request's final state for ckt_server_fetch (set before
 * a start or a trace); after a trace `index` is the trace index, otherwise
 * the request id. */
int ckt_server_set_keep_state(ckt_server_t* s, int keep_state);
/* Keep only the states of requests with at most max_wires wires (default:
 * all). */
int ckt_server_set_keep_limit(ckt_server_t* s, uint32_t max_wires);
int ckt_server_fetch(ckt_server_t* s, uint32_t index, ckt_node_t* nodes, ckt_wire_t* wires);

/* Debug builds (CKT_BOUNDS_CHECKS, the counterpart of the reference's
 * POINTER_CHECKS accessors, runtime/lowlevel.h:818-832): the number of
 * out-of-range indices the hot-path kernels found on `device` since the last
 * call (those accesses were skipped), after waiting for the device.  Builds
 * without the checks return CKT_ERR_UNSUPPORTED. */
int ckt_debug_bounds_violations(int device, uint64_t* count);

/* HBM calibration kernel: copies `bytes` from src to dst on the device with a
 * plain streaming kernel (used to scale profiler byte counters). */
int ckt_copy_kernel(void* dst, const void* src, size_t bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CIRCUIT_ABI_H */
