// ckt_engine.cpp -- device-resident circuits, tile lists and task launches.
#include "ckt_engine.h"
#include "ckt_index.h"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <string>

namespace ckt {

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

SlabLayout SlabLayout::make(uint32_t N, uint32_t W, uint32_t P) {
  SlabLayout L{};
  L.ws = (uint32_t)align_up(std::max<uint32_t>(W, 1), 64);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  };
  const size_t ws = L.ws, n = std::max<uint32_t>(N, 1), w = std::max<uint32_t>(W, 1);
  L.cur = take(ws * 4 * kSeg);
  L.vol = take(ws * 4 * (kSeg - 1));
  L.ind = take(ws * 4);
  L.res = take(ws * 4);
  L.wcap = take(ws * 4);
  L.in_ptr = take(ws * 4);
  L.out_ptr = take(ws * 4);
  L.in_loc = take(ws);
  L.out_loc = take(ws);
  L.charge = take(n * 4);
  L.voltage = take(n * 4);
  L.ncap = take(n * 4);
  L.leak = take(n * 4);
  L.locator = take(n);
  L.csr_row = take((n + 1) * 4);
  L.csr_ent = take(2 * w * 4);
  L.pieces = take(std::max<uint32_t>(P, 1) * sizeof(ckt_piece_t));
  L.total = off;
  return L;
}

void SlabImage::reset(uint32_t n, uint32_t w, uint32_t p) {
  N = n;
  W = w;
  P = p;
  lay = SlabLayout::make(n, w, p);
  bytes.assign(lay.total, 0);
  pieces.assign(p, ckt_piece_t{0, 0, 0, 0});
}

// Validates the loaded circuit and builds the node -> (wire, end) index in
// global (wire, in-before-out) order.
static int finalize_image(SlabImage& img) {
  const uint32_t N = img.N, W = img.W;
  const SlabLayout& L = img.lay;
  const uint32_t* in_ptr = img.at<uint32_t>(L.in_ptr);
  const uint32_t* out_ptr = img.at<uint32_t>(L.out_ptr);
  const uint8_t* in_loc = img.at<uint8_t>(L.in_loc);
  const uint8_t* out_loc = img.at<uint8_t>(L.out_loc);
  const uint8_t* loc = img.at<uint8_t>(L.locator);
  for (uint32_t w = 0; w < W; w++) {
    if (in_ptr[w] >= N || out_ptr[w] >= N)
      return fail(CKT_ERR_OUT_OF_RANGE, "wire " + std::to_string(w) + " points past the last node");
    if (in_loc[w] > 2 || out_loc[w] > 2)
      return fail(CKT_ERR_INVALID_ARG, "wire " + std::to_string(w) + " has an invalid PointerLocation");
  }
  for (uint32_t k = 0; k < img.P; k++) {
    const ckt_piece_t& p = img.pieces[k];
    if ((uint64_t)p.first_wire + p.num_wires > W || (uint64_t)p.first_node + p.num_nodes > N)
      return fail(CKT_ERR_OUT_OF_RANGE, "piece " + std::to_string(k) + " exceeds the circuit");
    // A private node belongs to one piece's region (circuit.cc:1261): only
    // that piece's wires may reach it.
    for (uint32_t w = p.first_wire; w < p.first_wire + p.num_wires; w++) {
      const uint32_t ends[2] = {in_ptr[w], out_ptr[w]};
      for (uint32_t e : ends)
        if (e - p.first_node >= p.num_nodes && loc[e])
          return fail(CKT_ERR_INVALID_ARG, "wire " + std::to_string(w) + " of piece " +
                                               std::to_string(k) +
                                               " reaches a private node of another piece");
    }
  }
  uint32_t* row = img.at<uint32_t>(L.csr_row);
  uint32_t* ent = img.at<uint32_t>(L.csr_ent);
  std::fill(row, row + N + 1, 0u);
  for (uint32_t w = 0; w < W; w++) {
    row[in_ptr[w] + 1]++;
    row[out_ptr[w] + 1]++;
  }
  for (uint32_t i = 0; i < N; i++) row[i + 1] += row[i];
  std::vector<uint32_t> pos(row, row + N);
  for (uint32_t w = 0; w < W; w++) {
    ent[pos[in_ptr[w]]++] = w << 1;
    ent[pos[out_ptr[w]]++] = (w << 1) | 1u;
  }
  std::memcpy(img.at<ckt_piece_t>(L.pieces), img.pieces.data(), img.P * sizeof(ckt_piece_t));
  return CKT_OK;
}

int image_from_aos(SlabImage& img, const ckt_node_t* nodes, uint32_t N, const ckt_wire_t* wires,
                   uint32_t W, const int32_t* locator, const ckt_piece_t* pieces, uint32_t P) {
  if ((N && !nodes) || (W && !wires) || (P && !pieces))
    return fail(CKT_ERR_INVALID_ARG, "load: null array");
  if (N == 0 || P == 0) return fail(CKT_ERR_INVALID_ARG, "load: need at least one node and one piece");
  if (N >= kMaxElements || W >= kMaxElements)
    return fail(CKT_ERR_OUT_OF_RANGE, "load: more than 2^30 nodes or wires");
  img.reset(N, W, P);
  const SlabLayout& L = img.lay;
  const size_t ws = L.ws;
  float* cur = img.at<float>(L.cur);
  float* vol = img.at<float>(L.vol);
  float *ind = img.at<float>(L.ind), *res = img.at<float>(L.res), *wcap = img.at<float>(L.wcap);
  uint32_t *in_ptr = img.at<uint32_t>(L.in_ptr), *out_ptr = img.at<uint32_t>(L.out_ptr);
  uint8_t *in_loc = img.at<uint8_t>(L.in_loc), *out_loc = img.at<uint8_t>(L.out_loc);
  for (uint32_t w = 0; w < W; w++) {
    const ckt_wire_t& x = wires[w];
    for (int s = 0; s < kSeg; s++) cur[s * ws + w] = x.current[s];
    for (int s = 0; s < kSeg - 1; s++) vol[s * ws + w] = x.voltage[s];
    ind[w] = x.inductance;
    res[w] = x.resistance;
    wcap[w] = x.capacitance;
    in_ptr[w] = x.in_ptr;
    out_ptr[w] = x.out_ptr;
    if (x.in_loc < 0 || x.in_loc > 2 || x.out_loc < 0 || x.out_loc > 2)
      return fail(CKT_ERR_INVALID_ARG, "wire " + std::to_string(w) + " has an invalid PointerLocation");
    in_loc[w] = (uint8_t)x.in_loc;
    out_loc[w] = (uint8_t)x.out_loc;
  }
  float *charge = img.at<float>(L.charge), *voltage = img.at<float>(L.voltage);
  float *ncap = img.at<float>(L.ncap), *leak = img.at<float>(L.leak);
  uint8_t* loc = img.at<uint8_t>(L.locator);
  for (uint32_t i = 0; i < N; i++) {
    charge[i] = nodes[i].charge;
    voltage[i] = nodes[i].voltage;
    ncap[i] = nodes[i].capacitance;
    leak[i] = nodes[i].leakage;
    loc[i] = 1;
  }
  if (locator) {
    for (uint32_t i = 0; i < N; i++) loc[i] = locator[i] ? 1 : 0;
  } else {
    // Without a locator, a node is shared when any wire names it as a
    // shared or ghost end.
    for (uint32_t w = 0; w < W; w++) {
      if (in_loc[w] != CKT_PRIVATE_PTR && in_ptr[w] < N) loc[in_ptr[w]] = 0;
      if (out_loc[w] != CKT_PRIVATE_PTR && out_ptr[w] < N) loc[out_ptr[w]] = 0;
    }
  }
  std::copy(pieces, pieces + P, img.pieces.begin());
  return finalize_image(img);
}

namespace {
struct ImageSink final : GenSink {
  SlabImage* img;
  size_t ws;
  float *voltage, *ncap, *leak, *ind, *res, *wcap;
  uint32_t *in_ptr, *out_ptr;
  uint8_t *in_loc, *out_loc, *loc;
  explicit ImageSink(SlabImage* im) : img(im) {
    const SlabLayout& L = im->lay;
    ws = L.ws;
    voltage = im->at<float>(L.voltage);
    ncap = im->at<float>(L.ncap);
    leak = im->at<float>(L.leak);
    ind = im->at<float>(L.ind);
    res = im->at<float>(L.res);
    wcap = im->at<float>(L.wcap);
    in_ptr = im->at<uint32_t>(L.in_ptr);
    out_ptr = im->at<uint32_t>(L.out_ptr);
    in_loc = im->at<uint8_t>(L.in_loc);
    out_loc = im->at<uint8_t>(L.out_loc);
    loc = im->at<uint8_t>(L.locator);
  }
  void node(uint32_t i, float v, float c, float l) override {
    voltage[i] = v;
    ncap[i] = c;
    leak[i] = l;
  }
  void wire(uint32_t w, uint32_t in, uint32_t out, float L, float R, float C) override {
    in_ptr[w] = in;
    out_ptr[w] = out;
    ind[w] = L;
    res[w] = R;
    wcap[w] = C;
  }
  void ends(uint32_t w, uint32_t* in, uint32_t* out) override {
    *in = in_ptr[w];
    *out = out_ptr[w];
  }
  void locations(uint32_t w, int32_t a, int32_t b) override {
    in_loc[w] = (uint8_t)a;
    out_loc[w] = (uint8_t)b;
  }
  void locator(uint32_t i, int32_t v) override { loc[i] = (uint8_t)v; }
  void piece(uint32_t n, const ckt_piece_t& p) override { img->pieces[n] = p; }
};
}  // namespace

int image_from_gen(SlabImage& img, const ckt_gen_params_t* params) {
  uint32_t N, W, P;
  int st = gen_validate(params, &N, &W, &P);
  if (st) return st;
  img.reset(N, W, P);
  ImageSink sink(&img);
  if ((st = gen_run(params, sink))) return st;
  return finalize_image(img);
}

// ---------------------------------------------------------------------------

static uint64_t fnv(uint64_t h, const void* data, size_t n) {
  const uint8_t* p = (const uint8_t*)data;
  for (size_t i = 0; i < n; i++) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  return h;
}

static uint64_t piece_key(const ckt_piece_t& p) {
  return fnv(1469598103934665603ull, &p, sizeof(p));
}

Engine::~Engine() {
  if (device_ < 0) return;
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  drop_caches();
  for (Circuit* c : slots_) {
    if (c->stream_ordered) (void)hipFreeAsync(c->slab, stream_);
    else (void)hipFree(c->slab);
    delete c;
  }
  if (stream_) (void)hipStreamSynchronize(stream_);
  slots_.clear();
  if (d_table_) (void)hipFree(d_table_);
  if (d_counters_) (void)hipFree(d_counters_);
  if (d_bad_) (void)hipFree(d_bad_);
  if (d_partial_) (void)hipFree(d_partial_);
  if (devgen_scratch_) (void)hipFree(devgen_scratch_);
  for (const PinnedChunk& c : staging_) (void)hipHostFree(c.p);
  for (auto& p : pending_) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (hipEvent_t e : event_pool_) (void)hipEventDestroy(e);
  if (stream_) (void)hipStreamDestroy(stream_);
}

int Engine::init(int device, uint64_t budget_bytes) {
  int is950 = 0;
  size_t lds = 0, hbm = 0;
  int st = ckt_device_info(device, &is950, &num_cus_, &lds, &hbm);
  if (st) return st;
  if (!is950) return fail(CKT_ERR_NO_DEVICE, "device is not gfx950; this library is built for gfx950 only");
  CKT_HIP(hipSetDevice(device));
  device_ = device;
  lds_limit_ = lds;
  budget_ = budget_bytes ? budget_bytes : (uint64_t)(0.9 * (double)hbm);
  CKT_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  // Planned circuits come from the device's default memory pool; keep what
  // they release for reuse instead of returning it at each synchronisation.
  // Without pool support, planned circuits use plain allocations instead.
  hipMemPool_t pool = nullptr;
  uint64_t keep = UINT64_MAX;
  pool_ok_ = hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess &&
             hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) == hipSuccess;
  (void)hipGetLastError();
  CKT_HIP(hipMalloc((void**)&d_counters_, 4 * sizeof(uint32_t)));
  CKT_HIP(hipMalloc((void**)&d_bad_, sizeof(unsigned long long)));
  CKT_HIP(launch_fill_u32(d_counters_, 0u, 4, stream_));
  CKT_HIP(hipStreamSynchronize(stream_));
  return CKT_OK;
}

int Engine::add(SlabImage& img, int32_t loops, uint32_t* id) {
  CKT_HIP(hipSetDevice(device_));
  if (budget_ && in_use_ + img.lay.total > budget_)
    return fail(CKT_ERR_BUDGET, "circuit needs " + std::to_string(img.lay.total) +
                                    " bytes; budget has " + std::to_string(budget_ - in_use_) + " left");
  void* slab = nullptr;
  hipError_t e = hipMalloc(&slab, img.lay.total);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc(circuit)");
  e = hipMemcpy(slab, img.bytes.data(), img.lay.total, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(slab);
    return hip_fail(e, "hipMemcpy(circuit upload)");
  }
  attach(slab, img.lay, img.N, img.W, img.P, img.pieces, loops, id);
  return CKT_OK;
}

int Engine::add_generated_on_device(const ckt_gen_params_t* params, int32_t loops, uint32_t* id) {
  uint32_t N, W, P;
  int st = gen_validate(params, &N, &W, &P);
  if (st) return st;
  CKT_HIP(hipSetDevice(device_));
  const SlabLayout lay = SlabLayout::make(N, W, P);
  if (budget_ && in_use_ + lay.total > budget_)
    return fail(CKT_ERR_BUDGET, "circuit needs " + std::to_string(lay.total) +
                                    " bytes; budget has " + std::to_string(budget_ - in_use_) + " left");
  void* slab = nullptr;
  hipError_t e = hipMalloc(&slab, lay.total);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc(circuit)");
  std::vector<ckt_piece_t> pieces;
  if ((st = devgen_build(params, lay, slab, stream_, pieces))) {
    (void)hipStreamSynchronize(stream_);
    (void)hipFree(slab);
    return st;
  }
  attach(slab, lay, N, W, P, pieces, loops, id);
  return CKT_OK;
}

void Engine::attach(void* slab, const SlabLayout& L, uint32_t N, uint32_t W, uint32_t P,
                    const std::vector<ckt_piece_t>& pieces, int32_t loops, uint32_t* id,
                    bool stream_ordered) {
  Circuit* c = new Circuit();
  c->stream_ordered = stream_ordered;
  c->id = next_id_++;
  c->num_nodes = N;
  c->num_wires = W;
  c->num_pieces = P;
  c->lay = L;
  c->slab = slab;
  c->pieces = pieces;
  c->loops_left = loops;
  uint8_t* b = (uint8_t*)slab;
  DevCircuit& d = c->dev;
  d.cur = (float*)(b + L.cur);
  d.vol = (float*)(b + L.vol);
  d.ind = (float*)(b + L.ind);
  d.res = (float*)(b + L.res);
  d.wcap = (float*)(b + L.wcap);
  d.in_ptr = (uint32_t*)(b + L.in_ptr);
  d.out_ptr = (uint32_t*)(b + L.out_ptr);
  d.in_loc = b + L.in_loc;
  d.out_loc = b + L.out_loc;
  d.charge = (float*)(b + L.charge);
  d.voltage = (float*)(b + L.voltage);
  d.ncap = (float*)(b + L.ncap);
  d.leak = (float*)(b + L.leak);
  d.locator = b + L.locator;
  d.csr_row = (uint32_t*)(b + L.csr_row);
  d.csr_ent = (uint32_t*)(b + L.csr_ent);
  d.pieces = (ckt_piece_t*)(b + L.pieces);
  d.num_wires = W;
  d.ws = L.ws;
  d.num_nodes = N;
  d.num_pieces = P;
  for (uint32_t k = 0; k < P; k++) c->piece_index.emplace(piece_key(pieces[k]), k);
  // Wires partitioned by the pieces -> the CSR modes apply; node ranges
  // disjoint -> the LDS-privatised reduction applies.
  {
    std::vector<ckt_piece_t> ps = pieces;
    std::sort(ps.begin(), ps.end(), [](const ckt_piece_t& a, const ckt_piece_t& x) {
      return a.first_wire < x.first_wire;
    });
    uint64_t at = 0;
    bool ok = true;
    for (auto& p : ps) {
      if (p.num_wires == 0) continue;
      if (p.first_wire != at) ok = false;
      at = (uint64_t)p.first_wire + p.num_wires;
    }
    c->csr_ok = ok && at == W;
    std::sort(ps.begin(), ps.end(), [](const ckt_piece_t& a, const ckt_piece_t& x) {
      return a.first_node < x.first_node;
    });
    uint64_t end = 0;
    ok = true;
    for (auto& p : ps) {
      if (p.num_nodes == 0) continue;
      if (p.first_node < end) ok = false;
      end = (uint64_t)p.first_node + p.num_nodes;
    }
    c->lds_ok = ok;
  }
  slots_.push_back(c);
  by_id_[c->id] = c;
  in_use_ += L.total;
  table_dirty_ = true;
  generation_++;
  if (id) *id = c->id;
}

int Engine::remove(uint32_t id) { return remove_many(&id, 1); }

// A slab from the stream's memory pool is released stream-ordered, after
// everything already queued on the stream, so removing only such circuits
// needs no synchronisation; any other slab is freed after one stream
// synchronisation for the whole set.  The device descriptor table is only
// rewritten by ensure_table, by a copy queued behind every launch already on
// the stream.  Cached tile
// lists and graphs are keyed by the engine generation, which changes here, so
// entries naming the removed circuits can never be looked up again; they stay
// until the cache's size bound evicts them (drop_caches waits for the stream
// first).
int Engine::remove_many(const uint32_t* ids, size_t n) {
  bool sync = false;
  for (size_t k = 0; k < n; k++) {
    auto it = by_id_.find(ids[k]);
    if (it == by_id_.end())
      return fail(CKT_ERR_NOT_FOUND, "no circuit with id " + std::to_string(ids[k]));
    if (!it->second->stream_ordered) sync = true;
  }
  if (n == 0) return CKT_OK;
  CKT_HIP(hipSetDevice(device_));
  if (sync) CKT_HIP(hipStreamSynchronize(stream_));
  for (size_t k = 0; k < n; k++) {
    auto it = by_id_.find(ids[k]);
    if (it == by_id_.end()) continue;  // listed twice
    Circuit* c = it->second;
    if (c->stream_ordered) CKT_HIP(hipFreeAsync(c->slab, stream_));
    else CKT_HIP(hipFree(c->slab));
    in_use_ -= c->lay.total;
    slots_.erase(std::find(slots_.begin(), slots_.end(), c));
    by_id_.erase(it);
    delete c;
  }
  table_dirty_ = true;
  generation_++;
  return CKT_OK;
}

Circuit* Engine::find(uint32_t id) {
  auto it = by_id_.find(id);
  return it == by_id_.end() ? nullptr : it->second;
}

int Engine::slot_of(uint32_t id) const {
  for (size_t i = 0; i < slots_.size(); i++)
    if (slots_[i]->id == id) return (int)i;
  return -1;
}

int Engine::ref_for(uint32_t slot, const ckt_piece_t& p, PieceRef* out) const {
  if (slot >= slots_.size()) return fail(CKT_ERR_NOT_FOUND, "no circuit in slot");
  const Circuit* c = slots_[slot];
  auto it = c->piece_index.find(piece_key(p));
  if (it == c->piece_index.end())
    return fail(CKT_ERR_OUT_OF_RANGE, "piece {first_wire " + std::to_string(p.first_wire) +
                                          ", first_node " + std::to_string(p.first_node) +
                                          "} is not a piece of the loaded circuit");
  out->slot = slot;
  out->piece = it->second;
  return CKT_OK;
}

void Engine::all_refs(std::vector<PieceRef>& refs, bool only_active) const {
  refs.clear();
  for (uint32_t s = 0; s < slots_.size(); s++) {
    if (only_active && slots_[s]->loops_left <= 0) continue;
    for (uint32_t k = 0; k < slots_[s]->num_pieces; k++) refs.push_back(PieceRef{s, k});
  }
}

int Engine::store(uint32_t id, ckt_node_t* nodes, ckt_wire_t* wires) {
  Circuit* c = find(id);
  if (!c) return fail(CKT_ERR_NOT_FOUND, "no circuit with id " + std::to_string(id));
  CKT_HIP(hipSetDevice(device_));
  CKT_HIP(hipStreamSynchronize(stream_));
  const SlabLayout& L = c->lay;
  std::vector<uint8_t> img(L.total);
  CKT_HIP(hipMemcpy(img.data(), c->slab, L.total, hipMemcpyDeviceToHost));
  auto at = [&](size_t off) { return img.data() + off; };
  const size_t ws = L.ws;
  if (wires) {
    const float* cur = (const float*)at(L.cur);
    const float* vol = (const float*)at(L.vol);
    const float *ind = (const float*)at(L.ind), *res = (const float*)at(L.res),
                *wcap = (const float*)at(L.wcap);
    const uint32_t *in_ptr = (const uint32_t*)at(L.in_ptr), *out_ptr = (const uint32_t*)at(L.out_ptr);
    const uint8_t *in_loc = at(L.in_loc), *out_loc = at(L.out_loc);
    for (uint32_t w = 0; w < c->num_wires; w++) {
      ckt_wire_t& x = wires[w];
      std::memset(&x, 0, sizeof(x));
      x.in_ptr = in_ptr[w];
      x.out_ptr = out_ptr[w];
      x.in_loc = in_loc[w];
      x.out_loc = out_loc[w];
      x.inductance = ind[w];
      x.resistance = res[w];
      x.capacitance = wcap[w];
      for (int s = 0; s < kSeg; s++) x.current[s] = cur[s * ws + w];
      for (int s = 0; s < kSeg - 1; s++) x.voltage[s] = vol[s * ws + w];
    }
  }
  if (nodes) {
    const float *charge = (const float*)at(L.charge), *voltage = (const float*)at(L.voltage),
                *ncap = (const float*)at(L.ncap), *leak = (const float*)at(L.leak);
    for (uint32_t i = 0; i < c->num_nodes; i++)
      nodes[i] = ckt_node_t{charge[i], voltage[i], ncap[i], leak[i]};
  }
  return CKT_OK;
}

int Engine::store_topology(uint32_t id, int32_t* locator, uint32_t* csr_row, uint32_t* csr_ent) {
  Circuit* c = find(id);
  if (!c) return fail(CKT_ERR_NOT_FOUND, "no circuit with id " + std::to_string(id));
  CKT_HIP(hipSetDevice(device_));
  CKT_HIP(hipStreamSynchronize(stream_));
  const uint8_t* b = (const uint8_t*)c->slab;
  const SlabLayout& L = c->lay;
  if (locator) {
    std::vector<uint8_t> loc(c->num_nodes);
    CKT_HIP(hipMemcpy(loc.data(), b + L.locator, c->num_nodes, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < c->num_nodes; i++) locator[i] = loc[i];
  }
  if (csr_row)
    CKT_HIP(hipMemcpy(csr_row, b + L.csr_row, ((size_t)c->num_nodes + 1) * 4, hipMemcpyDeviceToHost));
  if (csr_ent && c->num_wires)
    CKT_HIP(hipMemcpy(csr_ent, b + L.csr_ent, (size_t)c->num_wires * 8, hipMemcpyDeviceToHost));
  return CKT_OK;
}

int Engine::copy_voltages_async(uint32_t id, float* dst) {
  Circuit* c = find(id);
  if (!c) return fail(CKT_ERR_NOT_FOUND, "no circuit with id " + std::to_string(id));
  CKT_HIP(hipMemcpyAsync(dst, c->dev.voltage, (size_t)c->num_nodes * 4, hipMemcpyDeviceToHost,
                         stream_));
  return CKT_OK;
}

int Engine::download_voltages(uint32_t id, std::vector<float>& v) {
  Circuit* c = find(id);
  if (!c) return fail(CKT_ERR_NOT_FOUND, "no circuit with id " + std::to_string(id));
  CKT_HIP(hipSetDevice(device_));
  CKT_HIP(hipStreamSynchronize(stream_));
  v.resize(c->num_nodes);
  CKT_HIP(hipMemcpy(v.data(), c->dev.voltage, (size_t)c->num_nodes * 4, hipMemcpyDeviceToHost));
  return CKT_OK;
}

// The table is rewritten by a copy queued on the stream: it runs after every
// launch already queued there (which read the old contents) and before every
// later one.  Graphs are keyed by the engine generation, which changes with
// the table, so no captured graph is replayed against a rewritten table.
// Only growing the buffer waits for the stream.
int Engine::ensure_table() {
  if (!table_dirty_) return CKT_OK;
  std::vector<DevCircuit> tab(slots_.size());
  for (size_t i = 0; i < slots_.size(); i++) tab[i] = slots_[i]->dev;
  if (tab.size() > table_cap_) {
    CKT_HIP(hipStreamSynchronize(stream_));
    if (d_table_) CKT_HIP(hipFree(d_table_));
    d_table_ = nullptr;
    table_cap_ = std::max<size_t>(tab.size(), 2 * table_cap_);
    CKT_HIP(hipMalloc((void**)&d_table_, table_cap_ * sizeof(DevCircuit)));
  }
  if (!tab.empty()) {
    void* staged = nullptr;
    int st = stage(tab.data(), tab.size() * sizeof(DevCircuit), &staged);
    if (st) return st;
    CKT_HIP(hipMemcpyAsync(d_table_, staged, tab.size() * sizeof(DevCircuit),
                           hipMemcpyHostToDevice, stream_));
  }
  table_dirty_ = false;
  return CKT_OK;
}

void Engine::drop_caches() {
  if (tile_cache_.empty() && graphs_.empty()) return;
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
  graphs_.clear();
  for (auto& kv : tile_cache_) {
    if (kv.second.pooled) (void)hipFreeAsync(kv.second.p, stream_);
    else (void)hipFree(kv.second.p);
  }
  tile_cache_.clear();
}

// Tile lists are allocated from the stream's memory pool and filled by a
// copy from pinned staging queued on the stream, so building one for a new
// set of circuits does not wait for the launches already queued.
int Engine::upload_buf(uint64_t key, const void* data, size_t bytes, uint32_t n, DevBuf* out) {
  if (tile_cache_.size() >= 64) drop_caches();
  DevBuf b;
  b.n = n;
  b.pooled = pool_ok_;
  const size_t sz = std::max<size_t>(bytes, 16);
  CKT_HIP(b.pooled ? hipMallocAsync(&b.p, sz, stream_) : hipMalloc(&b.p, sz));
  if (bytes) {
    void* staged = nullptr;
    int st = stage(data, bytes, &staged);
    hipError_t e = st ? hipSuccess : hipMemcpyAsync(b.p, staged, bytes, hipMemcpyHostToDevice, stream_);
    if (st || e != hipSuccess) {
      if (b.pooled) (void)hipFreeAsync(b.p, stream_);
      else (void)hipFree(b.p);
      return st ? st : hip_fail(e, "hipMemcpyAsync(tiles)");
    }
  }
  tile_cache_[key] = b;
  *out = b;
  return CKT_OK;
}

enum TileKind { TK_WIRES = 1, TK_NODES = 2, TK_CSR = 3, TK_PIECES = 4, TK_CNC = 5 };
// update_voltages: 256 lanes x 8 nodes per tile.
constexpr int kUpvTile = 2048;
// Node bytes update_voltages moves per node (4 arrays read, 2 written) and the
// MI355X last-level cache (256 MB of Infinity Cache): a launch over more node
// bytes than that streams them with nontemporal accesses.
constexpr uint64_t kUpvBytesPerNode = 24;
constexpr uint64_t kLastLevelCacheBytes = 256ull << 20;

int Engine::tiles_for(int kind, const std::vector<PieceRef>& refs, int block, DevBuf* out) {
  uint64_t key = fnv(1469598103934665603ull, &kind, sizeof(kind));
  key = fnv(key, &block, sizeof(block));
  if (kind == TK_CNC) {
    const int sched[3] = {wires_per_lane_, tile_span_, tile_order_};
    key = fnv(key, sched, sizeof(sched));
  }
  key = fnv(key, &generation_, sizeof(generation_));
  key = fnv(key, refs.data(), refs.size() * sizeof(PieceRef));
  auto it = tile_cache_.find(key);
  if (it != tile_cache_.end()) {
    *out = it->second;
    return CKT_OK;
  }
  if (kind == TK_PIECES) {
    // `block` is the number of wire subranges per piece; each piece's rows of
    // partial sums follow the previous piece's.
    std::vector<PieceTile> t;
    t.reserve(refs.size());
    uint64_t off = 0;
    for (const PieceRef& r : refs) {
      t.push_back(PieceTile{r.slot, r.piece, (uint32_t)off, 0});
      off += (uint64_t)std::max(block, 1) * slots_[r.slot]->pieces[r.piece].num_nodes;
    }
    return upload_buf(key, t.data(), t.size() * sizeof(PieceTile), (uint32_t)t.size(), out);
  }
  std::vector<Tile> t;
  if (kind == TK_CSR) {
    std::vector<uint32_t> seen;
    for (const PieceRef& r : refs)
      if (std::find(seen.begin(), seen.end(), r.slot) == seen.end()) seen.push_back(r.slot);
    for (uint32_t s : seen) {
      const uint32_t N = slots_[s]->num_nodes;
      for (uint32_t off = 0; off < N; off += (uint32_t)block)
        t.push_back(Tile{s, off, std::min<uint32_t>((uint32_t)block, N - off), 0});
    }
  } else if (kind == TK_CNC) {
    std::vector<Tile> pieces;
    pieces.reserve(refs.size());
    for (const PieceRef& r : refs) {
      const ckt_piece_t& p = slots_[r.slot]->pieces[r.piece];
      pieces.push_back(Tile{r.slot, p.first_wire, p.num_wires, r.piece});
    }
    t = cnc_schedule<Tile>(pieces, (uint32_t)block * (uint32_t)wires_per_lane_, tile_span_, tile_order_);
  } else {
    for (const PieceRef& r : refs) {
      const ckt_piece_t& p = slots_[r.slot]->pieces[r.piece];
      const uint32_t first = kind == TK_WIRES ? p.first_wire : p.first_node;
      const uint32_t num = kind == TK_WIRES ? p.num_wires : p.num_nodes;
      for (uint32_t off = 0; off < num; off += (uint32_t)block)
        t.push_back(Tile{r.slot, first + off, std::min<uint32_t>((uint32_t)block, num - off), r.piece});
    }
  }
  return upload_buf(key, t.data(), t.size() * sizeof(Tile), (uint32_t)t.size(), out);
}

int Engine::timed_begin(int task, hipEvent_t* a) {
  (void)task;
  *a = nullptr;
  if (!timing_) return CKT_OK;
  if (event_pool_.empty()) {
    hipEvent_t e;
    CKT_HIP(hipEventCreate(&e));
    event_pool_.push_back(e);
  }
  *a = event_pool_.back();
  event_pool_.pop_back();
  CKT_HIP(hipEventRecord(*a, stream_));
  return CKT_OK;
}

int Engine::timed_end(int task, hipEvent_t a) {
  if (!timing_ || !a) return CKT_OK;
  hipEvent_t b;
  if (event_pool_.empty()) {
    CKT_HIP(hipEventCreate(&b));
  } else {
    b = event_pool_.back();
    event_pool_.pop_back();
  }
  CKT_HIP(hipEventRecord(b, stream_));
  pending_.push_back(Pending{task, a, b});
  if (pending_.size() > 4096) {
    double ms;
    uint64_t n;
    return kernel_time(-1, &ms, &n);
  }
  return CKT_OK;
}

int Engine::kernel_time(int task, double* ms, uint64_t* launches) {
  CKT_HIP(hipSetDevice(device_));
  CKT_HIP(hipStreamSynchronize(stream_));
  for (auto& p : pending_) {
    float t = 0.f;
    CKT_HIP(hipEventElapsedTime(&t, p.a, p.b));
    acc_ms_[p.task] += t;
    acc_n_[p.task]++;
    event_pool_.push_back(p.a);
    event_pool_.push_back(p.b);
  }
  pending_.clear();
  if (task >= 0 && task < CKT_NUM_TASK_IDS) {
    if (ms) *ms = acc_ms_[task];
    if (launches) *launches = acc_n_[task];
  }
  return CKT_OK;
}

void Engine::reset_timing() {
  double ms;
  uint64_t n;
  (void)kernel_time(-1, &ms, &n);
  for (int i = 0; i < CKT_NUM_TASK_IDS; i++) {
    acc_ms_[i] = 0;
    acc_n_[i] = 0;
  }
}

int Engine::cnc(const std::vector<PieceRef>& refs, int policy, bool fuse_dsc) {
  if (policy < CKT_POLICY_EXACT || policy > CKT_POLICY_SCALED)
    return fail(CKT_ERR_INVALID_ARG, "calc_new_currents: unknown policy");
  int st;
  if ((st = ensure_table())) return st;
  DevBuf t;
  const int block = cnc_block_for(policy);
  if ((st = tiles_for(TK_CNC, refs, block, &t))) return st;
  hipEvent_t a;
  if ((st = timed_begin(CKT_CALC_NEW_CURRENTS, &a))) return st;
  hipError_t e;
  // Launch shape unless set with OPT_CNC_BLOCK / OPT_PERSISTENT /
  // OPT_WAVES_PER_SIMD.  On the 64 x 16384 x 65536 circuit a persistent tile
  // queue of 256-lane workgroups with 4 resident waves per SIMD beats the
  // plain grid by 6% for SCALED and 1-2% for FAST and RECIP.  For EXACT (71
  // VGPRs, at most 7 resident waves) a queue of 64-lane workgroups at 7 waves
  // per SIMD is 1-3% faster than the plain grid of 256-lane workgroups, which
  // it keeps when the workgroup size is set (a queue of 256-lane workgroups
  // is 1-4% slower than that grid).
  const bool persistent = persistent_ >= 0 ? persistent_ != 0
                                           : policy != CKT_POLICY_EXACT || cnc_block_ < 0;
  const int waves = waves_per_simd_ > 0 ? waves_per_simd_ : (policy == CKT_POLICY_EXACT ? 7 : 4);
  if (persistent) {
    const int waves_per_block = block / 64;
    int grid = num_cus_ * 4 * waves / waves_per_block;
    if (grid > (int)t.n) grid = (int)t.n;
    e = launch_cnc_persistent(policy, fuse_dsc, d_table_, (const Tile*)t.p, t.n, block, grid,
                              d_counters_, steps_, stream_);
  } else {
    e = launch_cnc(policy, fuse_dsc, d_table_, (const Tile*)t.p, t.n, block, steps_, stream_);
  }
  if (e != hipSuccess) return hip_fail(e, "calc_new_currents launch");
  return timed_end(CKT_CALC_NEW_CURRENTS, a);
}

// Every tile list one loop's launches use, built before the first launch.
// A synchronous upload issued while a kernel runs waits for CUs the kernel
// holds (for the whole of a calc_new_currents launch), and the host with it.
int Engine::prefetch_tiles(const std::vector<PieceRef>& refs, int policy, int mode) {
  int st;
  if (mode < CKT_REDUCE_ATOMIC || mode > CKT_REDUCE_CSR_KAHAN64)
    return fail(CKT_ERR_INVALID_ARG, "distribute_charge: unknown reduction mode");
  // room for every entry made here, so none is evicted before its launch
  if (tile_cache_.size() + 8 > 64) drop_caches();
  if ((st = ensure_table())) return st;
  DevBuf t;
  if ((st = tiles_for(TK_CNC, refs, cnc_block_for(policy), &t))) return st;
  if (mode == CKT_REDUCE_ATOMIC && !fuse_dsc_) {
    if ((st = tiles_for(TK_WIRES, refs, 256, &t))) return st;
    LdsPlan plan;
    if ((st = lds_plan(refs, &plan))) return st;
    if (plan.lds && (st = tiles_for(TK_PIECES, refs, (int)plan.splits, &t))) return st;
  }
  if (mode != CKT_REDUCE_ATOMIC) {
    if ((st = csr_nodes_ok(refs))) return st;
    if ((st = tiles_for(TK_CSR, refs, 256, &t))) return st;
  }
  return tiles_for(TK_NODES, refs, kUpvTile, &t);
}

int Engine::loop_once(const std::vector<PieceRef>& refs, int policy, int mode) {
  int st;
  if ((st = prefetch_tiles(refs, policy, mode))) return st;
  return launch_loop(refs, policy, mode);
}

int Engine::launch_loop(const std::vector<PieceRef>& refs, int policy, int mode) {
  int st;
  if (fuse_dsc_ && mode == CKT_REDUCE_ATOMIC) {
    if ((st = cnc(refs, policy, true))) return st;
  } else {
    if ((st = cnc(refs, policy))) return st;
    if ((st = dsc(refs, mode))) return st;
  }
  return upv(refs);
}

int Engine::csr_nodes_ok(const std::vector<PieceRef>& refs) const {
  std::map<uint32_t, std::vector<uint32_t>> per;
  for (const PieceRef& r : refs) per[r.slot].push_back(r.piece);
  for (auto& kv : per) {
    const Circuit* c = slots_[kv.first];
    if (!c->csr_ok)
      return fail(CKT_ERR_UNSUPPORTED, "CSR reduction needs pieces that partition the wires");
    std::vector<uint32_t>& v = kv.second;
    std::sort(v.begin(), v.end());
    if (v.size() != c->num_pieces || std::unique(v.begin(), v.end()) != v.end())
      return fail(CKT_ERR_UNSUPPORTED,
                  "CSR reduction runs over every piece of a circuit in one launch");
  }
  return CKT_OK;
}

int Engine::lds_plan(const std::vector<PieceRef>& refs, LdsPlan* plan) {
  LdsPlan p;
  p.lds = dsc_atomic_impl_ != 1 && !refs.empty();
  uint32_t min_wires = UINT32_MAX;
  uint64_t nodes = 0;
  for (const PieceRef& r : refs) {
    const Circuit* c = slots_[r.slot];
    const ckt_piece_t& pc = c->pieces[r.piece];
    p.lds = p.lds && c->lds_ok;
    p.max_nodes = std::max(p.max_nodes, pc.num_nodes);
    min_wires = std::min(min_wires, pc.num_wires);
    nodes += pc.num_nodes;
  }
  p.lds_bytes = dsc_lds_bytes(p.max_nodes);
  p.lds = p.lds && p.lds_bytes <= lds_limit_;
  if (p.lds) {
    // Two workgroups per CU, with no subrange under 2048 wires and at most 16
    // per piece.
    const uint64_t want = ((uint64_t)2 * num_cus_ + refs.size() - 1) / refs.size();
    const uint64_t by_size = std::max<uint32_t>(1, min_wires / 2048);
    p.splits = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>({want, by_size, 16}));
    if ((uint64_t)p.splits * nodes > 0xffffffffull) p.splits = 1;
    // Rows of partial sums even with one subrange: the second kernel's plain
    // read-add-write of every node replaces an atomic per shared node.
    p.partial_floats = (uint64_t)p.splits * nodes <= 0xffffffffull ? (uint64_t)p.splits * nodes : 0;
    if (p.partial_floats > partial_cap_) {
      CKT_HIP(hipStreamSynchronize(stream_));
      if (d_partial_) CKT_HIP(hipFree(d_partial_));
      d_partial_ = nullptr;
      partial_cap_ = 0;
      CKT_HIP(hipMalloc((void**)&d_partial_, p.partial_floats * sizeof(float)));
      partial_cap_ = p.partial_floats;
    }
  }
  *plan = p;
  return CKT_OK;
}

int Engine::dsc(const std::vector<PieceRef>& refs, int mode) {
  int st;
  if (mode < CKT_REDUCE_ATOMIC || mode > CKT_REDUCE_CSR_KAHAN64)
    return fail(CKT_ERR_INVALID_ARG, "distribute_charge: unknown reduction mode");
  if ((st = ensure_table())) return st;
  DevBuf t;
  hipError_t e;
  hipEvent_t a;
  if (mode == CKT_REDUCE_ATOMIC) {
    LdsPlan plan;
    if ((st = lds_plan(refs, &plan))) return st;
    if (dsc_atomic_impl_ == 2 && !plan.lds)
      return fail(CKT_ERR_UNSUPPORTED, "LDS reduction: a piece's node range does not fit in LDS");
    if (plan.lds) {
      if ((st = tiles_for(TK_PIECES, refs, (int)plan.splits, &t))) return st;
      if ((st = timed_begin(CKT_DISTRIBUTE_CHARGE, &a))) return st;
      e = launch_dsc_lds(d_table_, (const PieceTile*)t.p, t.n, plan.splits, plan.max_nodes,
                         plan.partial_floats ? d_partial_ : nullptr, plan.lds_bytes, stream_);
    } else {
      if ((st = tiles_for(TK_WIRES, refs, 256, &t))) return st;
      if ((st = timed_begin(CKT_DISTRIBUTE_CHARGE, &a))) return st;
      e = launch_dsc_atomic(d_table_, (const Tile*)t.p, t.n, 256, stream_);
    }
  } else {
    if ((st = csr_nodes_ok(refs))) return st;
    if ((st = tiles_for(TK_CSR, refs, 256, &t))) return st;
    if ((st = timed_begin(CKT_DISTRIBUTE_CHARGE, &a))) return st;
    e = launch_dsc_csr(mode == CKT_REDUCE_CSR_KAHAN64, d_table_, (const Tile*)t.p, t.n, 256, stream_);
  }
  if (e != hipSuccess) return hip_fail(e, "distribute_charge launch");
  return timed_end(CKT_DISTRIBUTE_CHARGE, a);
}

int Engine::upv(const std::vector<PieceRef>& refs) {
  int st;
  if ((st = ensure_table())) return st;
  DevBuf t;
  if ((st = tiles_for(TK_NODES, refs, kUpvTile, &t))) return st;
  hipEvent_t a;
  if ((st = timed_begin(CKT_UPDATE_VOLTAGES, &a))) return st;
  uint64_t nodes = 0;
  for (const PieceRef& r : refs) nodes += slots_[r.slot]->pieces[r.piece].num_nodes;
  const bool nontemporal = nodes * kUpvBytesPerNode > kLastLevelCacheBytes;
  hipError_t e = launch_upv(d_table_, (const Tile*)t.p, t.n, 256, nontemporal, stream_);
  if (e != hipSuccess) return hip_fail(e, "update_voltages launch");
  return timed_end(CKT_UPDATE_VOLTAGES, a);
}

int Engine::sanity(const std::vector<PieceRef>& refs, uint64_t* bad) {
  int st;
  if ((st = ensure_table())) return st;
  DevBuf t;
  if ((st = tiles_for(TK_WIRES, refs, 256, &t))) return st;
  CKT_HIP(hipMemsetAsync(d_bad_, 0, sizeof(*d_bad_), stream_));
  hipError_t e = launch_sanity(d_table_, (const Tile*)t.p, t.n, 256, d_bad_, stream_);
  if (e != hipSuccess) return hip_fail(e, "sanity_check_wires launch");
  unsigned long long h = 0;
  CKT_HIP(hipMemcpyAsync(&h, d_bad_, sizeof(h), hipMemcpyDeviceToHost, stream_));
  CKT_HIP(hipStreamSynchronize(stream_));
  if (bad) *bad = h;
  return CKT_OK;
}

int Engine::nop() {
  hipError_t e = launch_nop(stream_);
  if (e != hipSuccess) return hip_fail(e, "nop launch");
  return CKT_OK;
}

int Engine::sync() {
  CKT_HIP(hipSetDevice(device_));
  CKT_HIP(hipStreamSynchronize(stream_));
  // Every copy out of the staging chunks has run.
  staging_chunk_ = 0;
  staging_off_ = 0;
  return CKT_OK;
}

// Copies `bytes` into pinned staging memory (new chunks of at least 64 MiB
// as needed) and returns its address, valid until the next sync().
int Engine::stage(const void* src, size_t bytes, void** out) {
  while (staging_chunk_ < staging_.size() &&
         staging_off_ + bytes > staging_[staging_chunk_].cap) {
    staging_chunk_++;
    staging_off_ = 0;
  }
  if (staging_chunk_ == staging_.size()) {
    PinnedChunk c{nullptr, std::max<size_t>(bytes, 64ull << 20)};
    CKT_HIP(hipHostMalloc(&c.p, c.cap, hipHostMallocDefault));
    staging_.push_back(c);
    staging_off_ = 0;
  }
  uint8_t* dst = (uint8_t*)staging_[staging_chunk_].p + staging_off_;
  std::memcpy(dst, src, bytes);
  staging_off_ = align_up(staging_off_ + bytes, 256);
  *out = dst;
  return CKT_OK;
}

int Engine::add_planned(const DevgenPlan& plan, int32_t loops, uint32_t* id) {
  CKT_HIP(hipSetDevice(device_));
  const SlabLayout lay = SlabLayout::make(plan.N, plan.W, plan.P);
  if (budget_ && in_use_ + lay.total > budget_)
    return fail(CKT_ERR_BUDGET, "circuit needs " + std::to_string(lay.total) +
                                    " bytes; budget has " + std::to_string(budget_ - in_use_) + " left");
  if (plan.wire_state.size() != plan.W) return fail(CKT_ERR_INVALID_ARG, "incomplete generator plan");
  const size_t need = devgen_scratch_bytes(plan.N, plan.W);
  if (need == 0) return fail(CKT_ERR_HIP, "device generator scratch size query failed");
  if (need > devgen_scratch_cap_) {
    // Builds already enqueued use the old scratch.
    CKT_HIP(hipStreamSynchronize(stream_));
    if (devgen_scratch_) CKT_HIP(hipFree(devgen_scratch_));
    devgen_scratch_ = nullptr;
    devgen_scratch_cap_ = 0;
    CKT_HIP(hipMalloc(&devgen_scratch_, need));
    devgen_scratch_cap_ = need;
  }
  void* staged = nullptr;
  int st;
  if (plan.W && (st = stage(plan.wire_state.data(), (size_t)plan.W * sizeof(uint64_t), &staged)))
    return st;
  // From the stream's memory pool: allocated and, at removal, released in
  // stream order, without waiting for the device.
  void* slab = nullptr;
  hipError_t e = pool_ok_ ? hipMallocAsync(&slab, lay.total, stream_) : hipMalloc(&slab, lay.total);
  if (e != hipSuccess) return hip_fail(e, "allocating a circuit");
  std::vector<ckt_piece_t> pieces;
  if ((st = devgen_enqueue(plan, (const uint64_t*)staged, lay, slab, devgen_scratch_, stream_,
                           pieces))) {
    (void)hipStreamSynchronize(stream_);
    if (pool_ok_) (void)hipFreeAsync(slab, stream_);
    else (void)hipFree(slab);
    (void)hipStreamSynchronize(stream_);
    return st;
  }
  attach(slab, lay, plan.N, plan.W, plan.P, pieces, loops, id, pool_ok_);
  return CKT_OK;
}

int Engine::run_graph(const std::vector<PieceRef>& refs, int policy, int mode, int times) {
  int st;
  if ((st = ensure_table())) return st;
  uint64_t key = fnv(1469598103934665603ull, refs.data(), refs.size() * sizeof(PieceRef));
  const int64_t cfg[12] = {policy,           mode,       steps_,          cnc_block_,
                           persistent_,      waves_per_simd_, dsc_atomic_impl_, fuse_dsc_,
                           wires_per_lane_,  tile_span_, tile_order_,     (int64_t)generation_};
  key = fnv(key, cfg, sizeof(cfg));
  auto it = graphs_.find(key);
  hipGraphExec_t exec = nullptr;
  if (it == graphs_.end()) {
    // Every buffer the captured launches use is created before capture (an
    // allocation during capture is not allowed).
    if ((st = prefetch_tiles(refs, policy, mode))) return st;
    const bool timing = timing_;
    timing_ = false;
    CKT_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    st = launch_loop(refs, policy, mode);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(stream_, &graph);
    timing_ = timing;
    if (st) {
      if (graph) (void)hipGraphDestroy(graph);
      return st;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    graphs_[key] = exec;
  } else {
    exec = it->second;
  }
  for (int i = 0; i < times; i++) CKT_HIP(hipGraphLaunch(exec, stream_));
  return CKT_OK;
}

int Engine::set_option(int opt, int64_t v) {
  switch (opt) {
    case OPT_STEPS:
      if (v < 0 || v > (1 << 30)) return fail(CKT_ERR_INVALID_ARG, "steps out of range");
      steps_ = (int)v;
      return CKT_OK;
    case OPT_CNC_BLOCK:
      if (v != -1 && v != 64 && v != 128 && v != 256 && v != 512)
        return fail(CKT_ERR_INVALID_ARG, "block must be 64, 128, 256, 512 or -1");
      cnc_block_ = (int)v;
      return CKT_OK;
    case OPT_PERSISTENT: persistent_ = v < 0 ? -1 : v ? 1 : 0; return CKT_OK;
    case OPT_WAVES_PER_SIMD:
      if (v > 8 || v == 0) return fail(CKT_ERR_INVALID_ARG, "waves per SIMD must be 1..8 or -1");
      waves_per_simd_ = v < 0 ? -1 : (int)v;
      return CKT_OK;
    case OPT_DSC_ATOMIC_IMPL:
      if (v < 0 || v > 2) return fail(CKT_ERR_INVALID_ARG, "atomic implementation must be 0, 1 or 2");
      dsc_atomic_impl_ = (int)v;
      return CKT_OK;
    case OPT_FUSE_DSC: fuse_dsc_ = v ? 1 : 0; return CKT_OK;
    case OPT_WIRES_PER_LANE:
      if (v != 1 && v != 2 && v != 4) return fail(CKT_ERR_INVALID_ARG, "wires per lane must be 1, 2 or 4");
      wires_per_lane_ = (int)v;
      return CKT_OK;
    case OPT_TILE_SPAN:
      if (v < 0 || v > 1) return fail(CKT_ERR_INVALID_ARG, "tile span must be 0 or 1");
      tile_span_ = (int)v;
      return CKT_OK;
    case OPT_TILE_ORDER:
      if (v < 0 || v > 2) return fail(CKT_ERR_INVALID_ARG, "tile order must be 0, 1 or 2");
      tile_order_ = (int)v;
      return CKT_OK;
    default: return fail(CKT_ERR_INVALID_ARG, "unknown option");
  }
}

int64_t Engine::option(int opt) const {
  switch (opt) {
    case OPT_STEPS: return steps_;
    case OPT_CNC_BLOCK: return cnc_block_;
    case OPT_PERSISTENT: return persistent_;
    case OPT_WAVES_PER_SIMD: return waves_per_simd_;
    case OPT_DSC_ATOMIC_IMPL: return dsc_atomic_impl_;
    case OPT_FUSE_DSC: return fuse_dsc_;
    case OPT_WIRES_PER_LANE: return wires_per_lane_;
    case OPT_TILE_SPAN: return tile_span_;
    case OPT_TILE_ORDER: return tile_order_;
    default: return -1;
  }
}

}  // namespace ckt
