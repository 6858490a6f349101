/*
 * circuit_oracle.c -- CPU restatement of the Legion circuit hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle; see circuit_oracle.h).  Built with
 * gcc -O3 -ffp-contract=off and no fast-math, so every float operation is a
 * single IEEE-754 binary32 operation as in the reference build
 * (runtime/runtime.mk:38; x86-64 SSE arithmetic has no excess precision).
 */
#define _GNU_SOURCE
#include "circuit_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define SEG ORC_WIRE_SEGMENTS

int orc_sizeof_node(void) { return (int)sizeof(orc_node_t); }
int orc_sizeof_wire(void) { return (int)sizeof(orc_wire_t); }

/* ------------------------------------------------------------------------ */
/* Generator                                                                  */

static uint64_t g_draws;
static double draw(void) {
  g_draws++;
  return drand48();
}

void orc_drand48_stream(int seed, double* out, int count) {
  srand48(seed);
  for (int i = 0; i < count; i++) out[i] = drand48();
}

/* random_element (circuit.cc:1042-1049): an index drawn uniformly into the
 * ordered set, then the element at that position. */
static uint32_t random_element(const uint32_t* set, uint32_t size) {
  int index = (int)(draw() * (double)size);
  return set[index];
}

static int cmp_u32(const void* a, const void* b) {
  uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  return (x > y) - (x < y);
}

static int contains(const uint32_t* set, uint32_t size, uint32_t v) {
  return size && bsearch(&v, set, size, sizeof(uint32_t), cmp_u32) != NULL;
}

typedef struct {
  uint32_t* v;
  uint32_t n, cap;
} vec_u32;

static int vec_push(vec_u32* s, uint32_t x) {
  if (s->n == s->cap) {
    uint32_t nc = s->cap ? 2 * s->cap : 16;
    uint32_t* nv = (uint32_t*)realloc(s->v, (size_t)nc * sizeof(uint32_t));
    if (!nv) return -1;
    s->v = nv;
    s->cap = nc;
  }
  s->v[s->n++] = x;
  return 0;
}

static void vec_sort_unique(vec_u32* s) {
  if (s->n < 2) return;
  qsort(s->v, s->n, sizeof(uint32_t), cmp_u32);
  uint32_t k = 1;
  for (uint32_t i = 1; i < s->n; i++)
    if (s->v[i] != s->v[k - 1]) s->v[k++] = s->v[i];
  s->n = k;
}

/* find_location, circuit.cc:1051-1069: private, then own shared, then ghost. */
static int32_t find_location(uint32_t ptr, const vec_u32* priv, const vec_u32* shr,
                             const vec_u32* ghost) {
  if (contains(priv->v, priv->n, ptr)) return ORC_PRIVATE_PTR;
  if (contains(shr->v, shr->n, ptr)) return ORC_SHARED_PTR;
  if (contains(ghost->v, ghost->n, ptr)) return ORC_GHOST_PTR;
  return -1;
}

int orc_generate(int num_pieces, int nodes_per_piece, int wires_per_piece, int pct_wire_in_piece,
                 int random_seed, int profile, orc_node_t* nodes, orc_wire_t* wires,
                 int32_t* locator, orc_piece_t* pieces, uint64_t* draws) {
  if (num_pieces < 1 || nodes_per_piece < 1 || wires_per_piece < 0) return -1;
  const uint32_t P = (uint32_t)num_pieces, npp = (uint32_t)nodes_per_piece,
                 wpp = (uint32_t)wires_per_piece;
  const uint64_t N = (uint64_t)P * npp, W = (uint64_t)P * wpp;
  if (N >= (1ull << 31) || W >= (1ull << 31)) return -1;

  /* private_node_map[n] while the wires are drawn: every node of piece n, in
   * pointer order (circuit.cc:1132); privacy_map as a per-node flag
   * (0 = privacy_map[0], 1 = privacy_map[1], circuit.cc:1096,1133,1183-1184). */
  uint32_t* all_nodes_of = (uint32_t*)malloc((size_t)N * sizeof(uint32_t));
  uint8_t* privacy = (uint8_t*)calloc((size_t)N, 1);
  vec_u32* ghost = (vec_u32*)calloc(P, sizeof(vec_u32));
  vec_u32* priv = (vec_u32*)calloc(P, sizeof(vec_u32));
  vec_u32* shr = (vec_u32*)calloc(P, sizeof(vec_u32));
  int rc = 0;
  if (!all_nodes_of || !privacy || !ghost || !priv || !shr) {
    rc = -1;
    goto done;
  }

  g_draws = 0;
  srand48(random_seed);

  /* Nodes, circuit.cc:1104-1139. */
  for (uint32_t n = 0; n < P; n++) {
    if (pieces) {
      pieces[n].first_node = n * npp;
      pieces[n].num_nodes = npp;
      pieces[n].first_wire = n * wpp;
      pieces[n].num_wires = wpp;
    }
    for (uint32_t i = 0; i < npp; i++) {
      uint32_t ptr = n * npp + i;
      orc_node_t node;
      node.charge = 0.f;
      node.voltage = 2 * draw() - 1;
      node.capacitance = draw() + 1;
      node.leakage = 0.1f * draw();
      nodes[ptr] = node;
      all_nodes_of[ptr] = ptr;
    }
  }

  /* Wires, circuit.cc:1144-1196. */
  for (uint32_t n = 0; n < P; n++) {
    for (uint32_t i = 0; i < wpp; i++) {
      orc_wire_t wire;
      memset(&wire, 0, sizeof(wire));
      for (int j = 0; j < SEG; j++) wire.current[j] = 0.f;
      for (int j = 0; j < SEG - 1; j++) wire.voltage[j] = 0.f;
      if (profile == ORC_PROFILE_STABLE) {
        wire.resistance = draw() * 10 + 1;
        wire.inductance = (draw() * 0.01 + 0.1) * 1e-6;
        wire.capacitance = draw() * 0.01 + 0.001;
      } else {
        wire.resistance = draw() * 10 + 1;
        wire.inductance = draw() * 0.01 + 0.1;
        wire.capacitance = draw() * 0.1;
      }
      wire.in_ptr = random_element(&all_nodes_of[(size_t)n * npp], npp);
      if ((100 * draw()) < pct_wire_in_piece) {
        wire.out_ptr = random_element(&all_nodes_of[(size_t)n * npp], npp);
      } else {
        int nn = (int)(draw() * (num_pieces - 1));
        if (nn >= (int)n) nn++;
        /* With one piece the reference indexes piece 1, which does not exist;
         * the pick wraps to piece 0. */
        if (nn >= num_pieces) nn -= num_pieces;
        wire.out_ptr = random_element(&all_nodes_of[(size_t)nn * npp], npp);
        privacy[wire.out_ptr] = 1;
        if (vec_push(&ghost[n], wire.out_ptr)) {
          rc = -1;
          goto done;
        }
      }
      wires[(size_t)n * wpp + i] = wire;
    }
  }

  /* Second pass, circuit.cc:1199-1225. */
  for (uint32_t n = 0; n < P; n++) {
    for (uint32_t i = 0; i < npp; i++) {
      uint32_t ptr = n * npp + i;
      if (privacy[ptr]) {
        if (vec_push(&shr[n], ptr)) {
          rc = -1;
          goto done;
        }
        if (locator) locator[ptr] = 0;
      } else {
        if (vec_push(&priv[n], ptr)) {
          rc = -1;
          goto done;
        }
        if (locator) locator[ptr] = 1;
      }
    }
    vec_sort_unique(&ghost[n]);
  }

  /* Second pass, part 2: locations, circuit.cc:1227-1244. */
  for (uint32_t n = 0; n < P; n++) {
    for (uint32_t i = 0; i < wpp; i++) {
      orc_wire_t* w = &wires[(size_t)n * wpp + i];
      w->in_loc = find_location(w->in_ptr, &priv[n], &shr[n], &ghost[n]);
      w->out_loc = find_location(w->out_ptr, &priv[n], &shr[n], &ghost[n]);
      if (w->in_loc < 0 || w->out_loc < 0) rc = -2;
    }
  }

done:
  if (draws) *draws = g_draws;
  if (ghost)
    for (uint32_t n = 0; n < P; n++) free(ghost[n].v);
  if (priv)
    for (uint32_t n = 0; n < P; n++) free(priv[n].v);
  if (shr)
    for (uint32_t n = 0; n < P; n++) free(shr[n].v);
  free(ghost);
  free(priv);
  free(shr);
  free(privacy);
  free(all_nodes_of);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* calc_new_currents                                                          */

static const orc_node_t* get_node(const orc_node_t* pvt, const orc_node_t* shr,
                                  const orc_node_t* ghost, int32_t loc, uint32_t ptr) {
  switch (loc) { /* circuit.cc:80-96 */
    case ORC_PRIVATE_PTR: return &pvt[ptr];
    case ORC_SHARED_PTR: return &shr[ptr];
    case ORC_GHOST_PTR: return &ghost[ptr];
    default: return NULL;
  }
}

int orc_calc_new_currents_f32(int seeding, int steps, orc_wire_t* wires, const orc_node_t* pvt,
                              const orc_node_t* shr, const orc_node_t* ghost,
                              uint32_t first_wire, uint32_t num_wires) {
  int bad = 0;
  for (uint32_t k = 0; k < num_wires; k++) {
    orc_wire_t wire = wires[first_wire + k];
    const orc_node_t* in_node = get_node(pvt, shr, ghost, wire.in_loc, wire.in_ptr);
    const orc_node_t* out_node = get_node(pvt, shr, ghost, wire.out_loc, wire.out_ptr);
    if (!in_node || !out_node) {
      bad++;
      continue;
    }
    /* circuit.cc:376-401 / circuit_gpu.cu:103-129 */
    float dt = ORC_DELTAT;
    float new_v[SEG + 1];
    float new_i[SEG];
    for (int i = 0; i < SEG; i++) new_i[i] = wire.current[i];
    if (seeding == ORC_SEED_CPU_LITERAL) {
      for (int i = 0; i < SEG - 1; i++) new_v[i] = wire.voltage[i];
      new_v[SEG - 1] = 0.f;
      new_v[SEG] = out_node->voltage;
    } else {
      for (int i = 0; i < SEG - 1; i++) new_v[i + 1] = wire.voltage[i];
      new_v[0] = in_node->voltage;
      new_v[SEG] = out_node->voltage;
    }
    for (int j = 0; j < steps; j++) {
      for (int i = 0; i < SEG; i++)
        new_i[i] = ((new_v[i + 1] - new_v[i]) - (wire.inductance * (new_i[i] - wire.current[i]) / dt)) /
                   wire.resistance;
      for (int i = 0; i < SEG - 1; i++)
        new_v[i + 1] = wire.voltage[i] + dt * (new_i[i] - new_i[i + 1]) / wire.capacitance;
    }
    /* circuit.cc:404-409 */
    for (int i = 0; i < SEG; i++) wire.current[i] = new_i[i];
    for (int i = 0; i < SEG - 1; i++) wire.voltage[i] = new_v[i + 1];
    wires[first_wire + k] = wire;
  }
  return bad;
}

int orc_calc_new_currents_f64(int seeding, int steps, orc_wire64_t* wires,
                              const orc_node64_t* nodes, uint32_t first_wire, uint32_t num_wires) {
  const double dt = (double)(float)ORC_DELTAT;
  for (uint32_t k = 0; k < num_wires; k++) {
    orc_wire64_t wire = wires[first_wire + k];
    double new_v[SEG + 1];
    double new_i[SEG];
    for (int i = 0; i < SEG; i++) new_i[i] = wire.current[i];
    if (seeding == ORC_SEED_CPU_LITERAL) {
      for (int i = 0; i < SEG - 1; i++) new_v[i] = wire.voltage[i];
      new_v[SEG - 1] = 0.0;
    } else {
      for (int i = 0; i < SEG - 1; i++) new_v[i + 1] = wire.voltage[i];
      new_v[0] = nodes[wire.in_ptr].voltage;
    }
    new_v[SEG] = nodes[wire.out_ptr].voltage;
    for (int j = 0; j < steps; j++) {
      for (int i = 0; i < SEG; i++)
        new_i[i] = ((new_v[i + 1] - new_v[i]) - (wire.inductance * (new_i[i] - wire.current[i]) / dt)) /
                   wire.resistance;
      for (int i = 0; i < SEG - 1; i++)
        new_v[i + 1] = wire.voltage[i] + dt * (new_i[i] - new_i[i + 1]) / wire.capacitance;
    }
    for (int i = 0; i < SEG; i++) wire.current[i] = new_i[i];
    for (int i = 0; i < SEG - 1; i++) wire.voltage[i] = new_v[i + 1];
    wires[first_wire + k] = wire;
  }
  return 0;
}

int orc_calc_new_currents_f64_limit(orc_wire64_t* wires, const orc_node64_t* nodes,
                                    uint32_t first_wire, uint32_t num_wires) {
  const double dt = (double)(float)ORC_DELTAT;
  enum { NU = 2 * SEG - 1 }; /* I_0..I_9 then v_1..v_9 */
  int singular = 0;
  for (uint32_t k = 0; k < num_wires; k++) {
    orc_wire64_t* w = &wires[first_wire + k];
    const double vin = nodes[w->in_ptr].voltage, vout = nodes[w->out_ptr].voltage;
    const double L = w->inductance, R = w->resistance, C = w->capacitance;
    double A[NU][NU + 1];
    memset(A, 0, sizeof(A));
    for (int i = 0; i < SEG; i++) { /* KVL: (L/dt + R) I_i - v_{i+1} + v_i = (L/dt) I0_i */
      A[i][i] = L / dt + R;
      A[i][NU] = L / dt * w->current[i];
      if (i + 1 <= SEG - 1) A[i][SEG + i] -= 1.0;
      else A[i][NU] += vout;
      if (i >= 1) A[i][SEG + i - 1] += 1.0;
      else A[i][NU] -= vin;
    }
    for (int m = 1; m < SEG; m++) { /* KCL: (C/dt) v_m - I_{m-1} + I_m = (C/dt) V0_m */
      const int r = SEG + m - 1;
      A[r][r] = C / dt;
      A[r][m - 1] = -1.0;
      A[r][m] = 1.0;
      A[r][NU] = C / dt * w->voltage[m - 1];
    }
    for (int c = 0; c < NU; c++) { /* elimination with partial pivoting */
      int piv = c;
      for (int r = c + 1; r < NU; r++)
        if (fabs(A[r][c]) > fabs(A[piv][c])) piv = r;
      if (A[piv][c] == 0.0) {
        singular++;
        break;
      }
      if (piv != c)
        for (int j = 0; j <= NU; j++) {
          double t = A[c][j];
          A[c][j] = A[piv][j];
          A[piv][j] = t;
        }
      for (int r = c + 1; r < NU; r++) {
        const double f = A[r][c] / A[c][c];
        if (f != 0.0)
          for (int j = c; j <= NU; j++) A[r][j] -= f * A[c][j];
      }
    }
    double x[NU];
    for (int r = NU - 1; r >= 0; r--) {
      double acc = A[r][NU];
      for (int j = r + 1; j < NU; j++) acc -= A[r][j] * x[j];
      x[r] = acc / A[r][r];
    }
    for (int i = 0; i < SEG; i++) w->current[i] = x[i];
    for (int m = 0; m < SEG - 1; m++) w->voltage[m] = x[SEG + m];
  }
  return singular;
}

/* ------------------------------------------------------------------------ */
/* distribute_charge                                                          */

static int reduce_node(orc_node_t* pvt, orc_node_t* shr, orc_node_t* ghost, int32_t loc,
                       uint32_t ptr, float value) {
  switch (loc) { /* circuit.cc:98-116 with apply<true> (:32-36) */
    case ORC_PRIVATE_PTR: pvt[ptr].charge += value; return 0;
    case ORC_SHARED_PTR: shr[ptr].charge += value; return 0;
    case ORC_GHOST_PTR: ghost[ptr].charge += value; return 0;
    default: return 1;
  }
}

int orc_distribute_charge_f32(const orc_wire_t* wires, orc_node_t* pvt, orc_node_t* shr,
                              orc_node_t* ghost, uint32_t first_wire, uint32_t num_wires) {
  int bad = 0;
  for (uint32_t k = 0; k < num_wires; k++) {
    const orc_wire_t* wire = &wires[first_wire + k];
    float dt = 1e-6; /* circuit.cc:436 */
    bad += reduce_node(pvt, shr, ghost, wire->in_loc, wire->in_ptr, -dt * wire->current[0]);
    bad += reduce_node(pvt, shr, ghost, wire->out_loc, wire->out_ptr, dt * wire->current[SEG - 1]);
  }
  return bad;
}

static int reduce_fold(orc_node_t* pvt, float* shr_fold, float* ghost_fold, int32_t loc,
                       uint32_t ptr, float value) {
  switch (loc) { /* circuit_gpu.cu:219-243 */
    case ORC_PRIVATE_PTR: pvt[ptr].charge += value; return 0;
    case ORC_SHARED_PTR: shr_fold[ptr] += value; return 0;
    case ORC_GHOST_PTR: ghost_fold[ptr] += value; return 0;
    default: return 1;
  }
}

int orc_distribute_charge_fold_f32(const orc_wire_t* wires, orc_node_t* pvt, float* shr_fold,
                                   float* ghost_fold, uint32_t first_wire, uint32_t num_wires) {
  int bad = 0;
  for (uint32_t k = 0; k < num_wires; k++) {
    const orc_wire_t* wire = &wires[first_wire + k];
    float dt = 1e-6;
    bad += reduce_fold(pvt, shr_fold, ghost_fold, wire->in_loc, wire->in_ptr, -dt * wire->current[0]);
    bad += reduce_fold(pvt, shr_fold, ghost_fold, wire->out_loc, wire->out_ptr,
                       dt * wire->current[SEG - 1]);
  }
  return bad;
}

void orc_fold_apply_f32(orc_node_t* owner, const float* fold, uint32_t first, uint32_t count) {
  for (uint32_t i = first; i < first + count; i++) owner[i].charge += fold[i];
}

int orc_distribute_charge_f64(const orc_wire64_t* wires, orc_node64_t* nodes, uint32_t first_wire,
                              uint32_t num_wires) {
  const double dt = (double)(float)1e-6;
  for (uint32_t k = 0; k < num_wires; k++) {
    const orc_wire64_t* wire = &wires[first_wire + k];
    nodes[wire->in_ptr].charge += -dt * wire->current[0];
    nodes[wire->out_ptr].charge += dt * wire->current[SEG - 1];
  }
  return 0;
}

void orc_distribute_charge_kahan64(const orc_wire_t* wires, uint32_t num_wires, orc_node_t* nodes,
                                   const int32_t* locator, uint32_t num_nodes) {
  double* sum = (double*)malloc((size_t)num_nodes * sizeof(double));
  double* comp = (double*)calloc((size_t)num_nodes, sizeof(double));
  for (uint32_t n = 0; n < num_nodes; n++) sum[n] = (double)nodes[n].charge;
  for (uint32_t k = 0; k < num_wires; k++) {
    const orc_wire_t* wire = &wires[k];
    float dt = 1e-6;
    uint32_t ends[2] = {wire->in_ptr, wire->out_ptr};
    float vals[2] = {-dt * wire->current[0], dt * wire->current[SEG - 1]};
    for (int e = 0; e < 2; e++) {
      uint32_t n = ends[e];
      if (locator[n]) {
        nodes[n].charge += vals[e];
      } else {
        double y = (double)vals[e] - comp[n];
        double t = sum[n] + y;
        comp[n] = (t - sum[n]) - y;
        sum[n] = t;
      }
    }
  }
  for (uint32_t n = 0; n < num_nodes; n++)
    if (!locator[n]) nodes[n].charge = (float)sum[n];
  free(sum);
  free(comp);
}

/* ------------------------------------------------------------------------ */
/* update_voltages                                                            */

void orc_update_voltages_f32(orc_node_t* pvt, orc_node_t* shr, const int32_t* locator,
                             uint32_t first_node, uint32_t num_nodes) {
  for (uint32_t i = first_node; i < first_node + num_nodes; i++) {
    orc_node_t* node = locator[i] ? &pvt[i] : &shr[i];
    /* circuit.cc:1299-1302 */
    node->voltage += node->charge / node->capacitance;
    node->voltage *= (1.f - node->leakage);
    node->charge = 0.f;
  }
}

void orc_update_voltages_f64(orc_node64_t* nodes, uint32_t first_node, uint32_t num_nodes) {
  for (uint32_t i = first_node; i < first_node + num_nodes; i++) {
    orc_node64_t* node = &nodes[i];
    node->voltage += node->charge / node->capacitance;
    node->voltage *= (1.0 - node->leakage);
    node->charge = 0.0;
  }
}

/* ------------------------------------------------------------------------ */
/* Threaded calc_new_currents (the CPU processors of the reference run one   */
/* point task per thread).                                                    */

typedef struct {
  int seeding, steps, prec;
  orc_wire_t* wires32;
  const orc_node_t* nodes32;
  orc_wire64_t* wires64;
  const orc_node64_t* nodes64;
  const orc_piece_t* pieces;
  uint32_t num_pieces;
  uint32_t grain;
  const uint32_t* idx; /* explicit wire list (sample mode) or NULL */
  uint32_t count;
  uint64_t next; /* work counter */
  pthread_mutex_t mu;
} cnc_job;

static int take(cnc_job* j, uint64_t* item) {
  pthread_mutex_lock(&j->mu);
  *item = j->next++;
  pthread_mutex_unlock(&j->mu);
  return 1;
}

static void* cnc_worker(void* arg) {
  cnc_job* j = (cnc_job*)arg;
  uint64_t item;
  for (;;) {
    take(j, &item);
    if (j->idx) {
      uint64_t lo = item * j->grain;
      if (lo >= j->count) break;
      uint64_t hi = lo + j->grain < j->count ? lo + j->grain : j->count;
      for (uint64_t k = lo; k < hi; k++)
        orc_calc_new_currents_f32(j->seeding, j->steps, j->wires32, j->nodes32, j->nodes32,
                                  j->nodes32, j->idx[k], 1);
      continue;
    }
    if (j->grain == 0) {
      if (item >= j->num_pieces) break;
      const orc_piece_t* p = &j->pieces[item];
      if (j->prec == 32)
        orc_calc_new_currents_f32(j->seeding, j->steps, j->wires32, j->nodes32, j->nodes32,
                                  j->nodes32, p->first_wire, p->num_wires);
      else if (j->prec == 65)
        orc_calc_new_currents_f64_limit(j->wires64, j->nodes64, p->first_wire, p->num_wires);
      else
        orc_calc_new_currents_f64(j->seeding, j->steps, j->wires64, j->nodes64, p->first_wire,
                                  p->num_wires);
    } else {
      /* chunks of `grain` wires over the concatenated pieces; a chunk that
         crosses the end of a piece continues in the next one */
      uint64_t lo = item * j->grain, hi = lo + j->grain, base = 0;
      int found = 0;
      for (uint32_t p = 0; p < j->num_pieces && lo < hi; p++) {
        uint64_t nw = j->pieces[p].num_wires;
        if (lo < base + nw) {
          uint64_t off = lo - base;
          uint64_t cnt = nw - off < hi - lo ? nw - off : hi - lo;
          uint32_t fw = j->pieces[p].first_wire + (uint32_t)off;
          if (j->prec == 32)
            orc_calc_new_currents_f32(j->seeding, j->steps, j->wires32, j->nodes32, j->nodes32,
                                      j->nodes32, fw, (uint32_t)cnt);
          else if (j->prec == 65)
            orc_calc_new_currents_f64_limit(j->wires64, j->nodes64, fw, (uint32_t)cnt);
          else
            orc_calc_new_currents_f64(j->seeding, j->steps, j->wires64, j->nodes64, fw,
                                      (uint32_t)cnt);
          lo += cnt;
          found = 1;
        }
        base += nw;
      }
      if (!found) break;
    }
  }
  return NULL;
}

static void run_cnc_threads(cnc_job* j, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  pthread_mutex_init(&j->mu, NULL);
  j->next = 0;
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
  for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, cnc_worker, j);
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  free(th);
  pthread_mutex_destroy(&j->mu);
}

int orc_run_f32(int seeding, int steps, int loops, int reduce, int nthreads, int grain,
                orc_node_t* nodes, uint32_t num_nodes, orc_wire_t* wires, uint32_t num_wires,
                const int32_t* locator, const orc_piece_t* pieces, uint32_t num_pieces) {
  (void)num_wires;
  int bad = 0;
  for (int loop = 0; loop < loops; loop++) {
    cnc_job j;
    memset(&j, 0, sizeof(j));
    j.seeding = seeding;
    j.steps = steps;
    j.prec = 32;
    j.wires32 = wires;
    j.nodes32 = nodes;
    j.pieces = pieces;
    j.num_pieces = num_pieces;
    j.grain = grain < 0 ? 0 : (uint32_t)grain;
    run_cnc_threads(&j, nthreads);
    if (reduce == 1) {
      orc_distribute_charge_kahan64(wires, num_wires, nodes, locator, num_nodes);
    } else {
      for (uint32_t p = 0; p < num_pieces; p++)
        bad += orc_distribute_charge_f32(wires, nodes, nodes, nodes, pieces[p].first_wire,
                                         pieces[p].num_wires);
    }
    for (uint32_t p = 0; p < num_pieces; p++)
      orc_update_voltages_f32(nodes, nodes, locator, pieces[p].first_node, pieces[p].num_nodes);
  }
  return bad;
}

int orc_run_f64(int seeding, int steps, int loops, int nthreads, orc_node64_t* nodes,
                uint32_t num_nodes, orc_wire64_t* wires, uint32_t num_wires,
                const orc_piece_t* pieces, uint32_t num_pieces) {
  (void)num_nodes;
  (void)num_wires;
  for (int loop = 0; loop < loops; loop++) {
    cnc_job j;
    memset(&j, 0, sizeof(j));
    j.seeding = seeding;
    j.steps = steps;
    j.prec = 64;
    j.wires64 = wires;
    j.nodes64 = nodes;
    j.pieces = pieces;
    j.num_pieces = num_pieces;
    j.grain = 256;
    run_cnc_threads(&j, nthreads);
    for (uint32_t p = 0; p < num_pieces; p++)
      orc_distribute_charge_f64(wires, nodes, pieces[p].first_wire, pieces[p].num_wires);
    for (uint32_t p = 0; p < num_pieces; p++)
      orc_update_voltages_f64(nodes, pieces[p].first_node, pieces[p].num_nodes);
  }
  return 0;
}

int orc_run_f64_limit(int loops, int nthreads, orc_node64_t* nodes, uint32_t num_nodes,
                      orc_wire64_t* wires, uint32_t num_wires, const orc_piece_t* pieces,
                      uint32_t num_pieces) {
  (void)num_nodes;
  (void)num_wires;
  for (int loop = 0; loop < loops; loop++) {
    cnc_job j;
    memset(&j, 0, sizeof(j));
    j.prec = 65;
    j.wires64 = wires;
    j.nodes64 = nodes;
    j.pieces = pieces;
    j.num_pieces = num_pieces;
    j.grain = 256;
    run_cnc_threads(&j, nthreads);
    for (uint32_t p = 0; p < num_pieces; p++)
      orc_distribute_charge_f64(wires, nodes, pieces[p].first_wire, pieces[p].num_wires);
    for (uint32_t p = 0; p < num_pieces; p++)
      orc_update_voltages_f64(nodes, pieces[p].first_node, pieces[p].num_nodes);
  }
  return 0;
}

double orc_cnc_sample_f32(int seeding, int steps, int nthreads, orc_wire_t* wires,
                          const orc_node_t* nodes, const uint32_t* wire_idx, uint32_t count) {
  cnc_job j;
  memset(&j, 0, sizeof(j));
  j.seeding = seeding;
  j.steps = steps;
  j.prec = 32;
  j.wires32 = wires;
  j.nodes32 = nodes;
  j.idx = wire_idx;
  j.count = count;
  j.grain = 16;
  struct timespec a, b;
  clock_gettime(CLOCK_MONOTONIC, &a);
  run_cnc_threads(&j, nthreads);
  clock_gettime(CLOCK_MONOTONIC, &b);
  return (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
}

void orc_promote(const orc_node_t* n32, uint32_t num_nodes, const orc_wire_t* w32,
                 uint32_t num_wires, orc_node64_t* n64, orc_wire64_t* w64) {
  for (uint32_t i = 0; i < num_nodes; i++) {
    n64[i].charge = n32[i].charge;
    n64[i].voltage = n32[i].voltage;
    n64[i].capacitance = n32[i].capacitance;
    n64[i].leakage = n32[i].leakage;
  }
  for (uint32_t k = 0; k < num_wires; k++) {
    const orc_wire_t* a = &w32[k];
    orc_wire64_t* b = &w64[k];
    b->in_ptr = a->in_ptr;
    b->out_ptr = a->out_ptr;
    b->in_loc = a->in_loc;
    b->out_loc = a->out_loc;
    b->inductance = a->inductance;
    b->resistance = a->resistance;
    b->capacitance = a->capacitance;
    for (int i = 0; i < SEG; i++) b->current[i] = a->current[i];
    for (int i = 0; i < SEG - 1; i++) b->voltage[i] = a->voltage[i];
  }
}
