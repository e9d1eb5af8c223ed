// Serving loop: continuous batching of boolean circuits from many clients.
//
// Each request is a gate circuit (a DAG over wires) encrypted under one client's key.
// At admission the request's wire buffer and gate table are placed in a device arena
// whose size is the server's HBM budget; a request that does not fit waits in FIFO order
// until completed requests release their space. A worker thread keeps up to two kernel
// launches in flight (two streams); every launch gathers, from all admitted requests that
// are not already running, the gates whose inputs are ready (the next layer), up to the
// per-launch gate capacity. Gates read their inputs from and write their outputs to the
// request's wires in HBM, so circuit state never leaves the device between layers.
// There is no reference counterpart: the reference evaluates gates one call at a time.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <thread>

#include "hb_internal.hpp"

using namespace hb;

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// First-fit allocator over one device arena (offsets in bytes, 256-byte granules).
class Arena {
 public:
  void init(size_t bytes) {
    size_ = bytes;
    free_.clear();
    free_[0] = bytes;
  }
  bool alloc(size_t bytes, size_t* off) {
    bytes = (bytes + 255) & ~(size_t)255;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        size_t rest = it->second - bytes, start = it->first + bytes;
        free_.erase(it);
        if (rest) free_[start] = rest;
        used_ += bytes;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t size() const { return size_; }
  size_t used() const { return used_; }

 private:
  std::map<size_t, size_t> free_;
  size_t size_ = 0, used_ = 0;
};

struct Request {
  uint64_t id = 0;
  hb_ek* ek = nullptr;
  uint32_t n1 = 0, n_inputs = 0, n_gates = 0;
  std::vector<uint32_t> outputs;
  std::vector<uint32_t> idx;        // (in0, in1, out) per gate, in schedule order
  std::vector<uint8_t> kind;        // gate kind per scheduled gate
  std::vector<uint32_t> run_end;    // end of each run of one layer and one kind
  std::vector<uint32_t> run_layer;  // layer of each run
  std::vector<uint64_t> inputs;     // host copy until admission
  size_t off = 0, bytes = 0, idx_off = 0;
  uint32_t next = 0;       // first gate not yet launched
  uint32_t completed = 0;  // gates finished
  bool in_flight = false, finished = false;
  // traced requests (vfhe semantics): per gate in schedule order, the (n+1)*2N-word core
  // trace, kept in the arena at trace_off and copied out at completion
  bool traced = false;
  size_t trace_off = 0, trace_gate_bytes = 0;
  std::vector<uint8_t> trace_host;
  std::vector<uint64_t> out;
  double t_submit = 0, t_admit = 0, t_done = 0;
  double t_ready = 0;  // when the gates of the next run got their inputs
};

struct Slot {
  hipStream_t s = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  BootArgs* parts_dev = nullptr;
  BootArgs* parts_host = nullptr;  // pinned
  void* acc = nullptr;
  uint64_t* pair_flags = nullptr;  // pair mode: kPairFlagWords per gate, zeroed once
  uint64_t* pair_status = nullptr; // kPairStatusWords: timeout epoch, ticket counter
  uint64_t pair_tickets = 0;       // workgroups of the pair launches issued so far
  uint32_t pair_epoch = 0;
  bool pair = false;  // the launch in flight runs in pair mode
  bool busy = false;
  std::vector<std::pair<Request*, uint32_t>> work;
  uint32_t gates = 0;
  uint64_t seq = 0;  // launch order
};

}  // namespace

struct hb_server {
  int device = 0;
  uint32_t max_gates = 0;
  Arena arena;
  char* arena_base = nullptr;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  std::thread worker;
  uint64_t next_id = 1;
  std::deque<Request*> queue;
  std::vector<Request*> active;
  std::map<uint64_t, std::unique_ptr<Request>> reqs;
  Slot slots[2];
  uint64_t launch_seq = 0;
  hipStream_t copy = nullptr;
  hb_server_stats_t st{};
  std::string error;
  // gate latency samples: (ms from inputs ready to launch completion, gates)
  std::vector<std::pair<float, uint32_t>> gate_lat;
};

namespace {

constexpr size_t kAccPerGate = kScratchWordsPerGate * 8;
constexpr uint32_t kMaxParts = 4096;

int admit(hb_server* S) {
  while (!S->queue.empty()) {
    Request* r = S->queue.front();
    size_t off;
    if (!S->arena.alloc(r->bytes, &off)) break;
    r->off = off;
    char* base = S->arena_base + off;
    const size_t wires_bytes = (size_t)(r->n_inputs + r->n_gates) * r->n1 * 8;
    r->idx_off = (wires_bytes + 255) & ~(size_t)255;
    HIP_TRY(hipMemcpyAsync(base, r->inputs.data(), r->inputs.size() * 8, hipMemcpyHostToDevice,
                           S->copy));
    HIP_TRY(hipMemcpyAsync(base + r->idx_off, r->idx.data(), r->idx.size() * 4,
                           hipMemcpyHostToDevice, S->copy));
    HIP_TRY(hipStreamSynchronize(S->copy));
    std::vector<uint64_t>().swap(r->inputs);
    r->t_admit = now_ms();
    r->t_ready = r->t_admit;
    S->active.push_back(r);
    S->queue.pop_front();
    S->st.admitted++;
  }
  S->st.bytes_in_use = S->arena.used();
  S->st.queued = S->queue.size();
  return HB_OK;
}

int finish(hb_server* S, Request* r) {
  const char* base = S->arena_base + r->off;
  r->out.resize((size_t)r->outputs.size() * r->n1);
  for (size_t k = 0; k < r->outputs.size(); k++)
    HIP_TRY(hipMemcpyAsync(r->out.data() + k * r->n1, base + (size_t)r->outputs[k] * r->n1 * 8,
                           (size_t)r->n1 * 8, hipMemcpyDeviceToHost, S->copy));
  if (r->traced) {
    r->trace_host.resize(r->trace_gate_bytes * r->n_gates);
    HIP_TRY(hipMemcpyAsync(r->trace_host.data(), base + r->trace_off, r->trace_host.size(),
                           hipMemcpyDeviceToHost, S->copy));
  }
  HIP_TRY(hipStreamSynchronize(S->copy));
  S->arena.release(r->off, r->bytes);
  r->finished = true;
  r->t_done = now_ms();
  S->active.erase(std::find(S->active.begin(), S->active.end(), r));
  S->st.completed++;
  S->st.bytes_in_use = S->arena.used();
  return HB_OK;
}

int complete(hb_server* S, Slot& sl) {
  if (sl.pair) {
    // a pair-mode wait that timed out left this launch's outputs wrong: stop with an error
    // before any request takes them
    uint64_t v = 0;
    HIP_TRY(hipMemcpyAsync(&v, sl.pair_status, sizeof v, hipMemcpyDeviceToHost, sl.s));
    HIP_TRY(hipStreamSynchronize(sl.s));
    sl.pair = false;
    if ((uint32_t)(v >> 32) == sl.pair_epoch) {
      // the slot is retired without finishing its requests: their waiters see the error
      sl.work.clear();
      sl.gates = 0;
      sl.busy = false;
      return fail(HB_ERR_DEVICE, "pair mode: an exchange wait of a serving launch timed out");
    }
  }
  float ms = 0;
  if (hipEventElapsedTime(&ms, sl.ev0, sl.ev1) == hipSuccess) S->st.busy_ms += ms;
  S->st.gates_done += sl.gates;
  const double t = now_ms();
  for (auto& w : sl.work) {
    Request* r = w.first;
    r->completed += w.second;
    r->in_flight = false;
    S->gate_lat.emplace_back((float)(t - r->t_ready), w.second);
    r->t_ready = t;
    if (r->completed == r->n_gates) {
      int rc = finish(S, r);
      if (rc) return rc;
    }
  }
  sl.work.clear();
  sl.gates = 0;
  sl.busy = false;
  return HB_OK;
}

// Gates of the current layer that request r has not launched yet.
uint32_t layer_backlog(const Request* r) {
  if (r->in_flight || r->next >= r->n_gates) return 0;
  size_t ri = std::upper_bound(r->run_end.begin(), r->run_end.end(), r->next) - r->run_end.begin();
  const uint32_t layer = r->run_layer[ri];
  while (ri + 1 < r->run_end.size() && r->run_layer[ri + 1] == layer) ri++;
  return r->run_end[ri] - r->next;
}

// Gathers the next ready run of each idle request into a launch on slot `sl`.
int launch(hb_server* S, Slot& sl) {
  // Latency mode: at most one latency-mode workgroup per CU per launch, since under
  // moderate load many short launches beat few long ones (measured on MI355X, 8 clients,
  // Goldilocks, cap 256 against 2048 gates run one wavefront per gate: 8 circuits/s p99
  // 2.02 s against 2.62 s, 12 circuits/s p50/p99 3.4/5.3 s against 5.5/10.6 s).
  // Throughput mode: once the ready gates fill the batched (4, 2) shape's resident
  // workgroups (two per CU), a launch takes up to max_gates of them in that shape.
  const uint32_t cus = std::max<uint32_t>(1, engine().simd_slots / 4);
  const uint32_t batch_min = 2 * kMixedBatchGpw * cus;
  uint32_t ready = 0;
  if (S->max_gates >= batch_min)
    for (const Request* r : S->active) ready += layer_backlog(r);
  const bool throughput = S->max_gates >= batch_min && ready >= batch_min;
  uint32_t cap = throughput ? S->max_gates : std::min(S->max_gates, cus), used = 0, nparts = 0;
  for (Request* r : S->active) {
    if (r->in_flight || r->next >= r->n_gates || used >= cap) continue;
    // consecutive runs (one kind each) of the current layer, up to the capacity; the
    // next layer waits for these outputs
    size_t ri = std::upper_bound(r->run_end.begin(), r->run_end.end(), r->next) - r->run_end.begin();
    const uint32_t layer = r->run_layer[ri];
    uint32_t pos = r->next;
    uint64_t* wires = reinterpret_cast<uint64_t*>(S->arena_base + r->off);
    const uint32_t* idx = reinterpret_cast<const uint32_t*>(S->arena_base + r->off + r->idx_off);
    while (ri < r->run_end.size() && r->run_layer[ri] == layer && used < cap && nparts < kMaxParts) {
      const uint32_t cnt = std::min(r->run_end[ri] - pos, cap - used);
      BootArgs a = make_args(r->ek, r->kind[pos], cnt, wires, nullptr, nullptr, nullptr, wires,
                             (char*)sl.acc + (size_t)used * kAccPerGate);
      a.idx = idx + 3 * (size_t)pos;
      a.first_block = used;
      if (r->traced) a.trace = S->arena_base + r->off + r->trace_off + (size_t)pos * r->trace_gate_bytes;
      sl.parts_host[nparts++] = a;
      used += cnt;
      pos += cnt;
      if (pos < r->run_end[ri]) break;  // capacity reached inside the run
      ri++;
    }
    const uint32_t take = pos - r->next;
    if (take == 0) continue;
    r->next += take;
    r->in_flight = true;
    sl.work.emplace_back(r, take);
  }
  if (used == 0) return HB_OK;
  bool any_traced = false, narrow = true;
  int sets = 0;
  for (uint32_t k = 0; k < nparts; k++) {
    const BootArgs& a = sl.parts_host[k];
    any_traced |= a.trace != nullptr;
    narrow = narrow && split8_fits(a.br_len, a.br_bits);
  }
  for (const auto& w : sl.work) sets |= 1 << w.first->ek->p.set;
  // pair mode only while the other slot has no kernel in flight: beside another launch's
  // workgroups (whose LDS leaves no room for a pair workgroup on their CUs) the pair
  // workgroups would start as CUs free up, correct (they pair up as they start) but slower
  // than latency mode
  bool other_busy = false;
  for (const Slot& o : S->slots) other_busy = other_busy || (&o != &sl && o.busy);
  const MixedKernel kind = mixed_kernel_for(engine(), sets, used, any_traced, narrow,
                                            !other_busy && sl.pair_flags != nullptr);
  const uint32_t blocks = assign_blocks(sl.parts_host, nparts, mixed_gpw(kind));
  if (kind == MixedKernel::pair) {
    ++sl.pair_epoch;
    for (uint32_t k = 0; k < nparts; k++) {
      sl.parts_host[k].pair_flags = sl.pair_flags + (size_t)sl.parts_host[k].first_block * kPairFlagWords;
      sl.parts_host[k].pair_epoch = sl.pair_epoch;
      sl.parts_host[k].pair_status = sl.pair_status;
      sl.parts_host[k].pair_ticket0 = sl.pair_tickets;
      sl.parts_host[k].pair_wait_ticks = engine().pair_wait_ticks();
    }
  }
  sl.pair = kind == MixedKernel::pair;
  HIP_TRY(hipMemcpyAsync(sl.parts_dev, sl.parts_host, nparts * sizeof(BootArgs),
                         hipMemcpyHostToDevice, sl.s));
  HIP_TRY(hipEventRecord(sl.ev0, sl.s));
  HIP_TRY(launch_mixed_kernel(kind, sl.parts_dev, (int)nparts, blocks, narrow, sets, sl.s,
                              any_traced));
  if (kind == MixedKernel::pair) sl.pair_tickets += 2 * (uint64_t)used;
  HIP_TRY(hipEventRecord(sl.ev1, sl.s));
  sl.busy = true;
  sl.gates = used;
  sl.seq = ++S->launch_seq;
  S->st.launches++;
  if (kind == MixedKernel::batch) S->st.batched_launches++;
  if (kind == MixedKernel::pair) S->st.pair_launches++;
  S->st.max_launch_gates = std::max<uint64_t>(S->st.max_launch_gates, used);
  return HB_OK;
}

}  // namespace

namespace {

void worker_loop(hb_server* S) {
  (void)hipSetDevice(S->device);
  std::unique_lock<std::mutex> lk(S->mu);
  auto set_error = [&](int rc) {
    if (rc && S->error.empty()) S->error = hb_last_error();
  };
  for (;;) {
    // retire finished launches
    for (Slot& sl : S->slots)
      if (sl.busy && hipEventQuery(sl.ev1) == hipSuccess) {
        set_error(complete(S, sl));
        S->cv_done.notify_all();
      }
    set_error(admit(S));
    bool any_busy = S->slots[0].busy || S->slots[1].busy;
    if (S->stop && S->queue.empty() && S->active.empty() && !any_busy) break;
    if (!S->error.empty()) {
      // a device error stops scheduling; waiters see it through hb_server_wait
      S->cv_done.notify_all();
      if (!any_busy) {
        S->cv_work.wait(lk, [&] { return S->stop; });
        break;
      }
    }
    Slot* free_slot = !S->slots[0].busy ? &S->slots[0] : (!S->slots[1].busy ? &S->slots[1] : nullptr);
    if (free_slot && S->error.empty()) {
      set_error(launch(S, *free_slot));
      if (free_slot->busy) continue;  // try to fill the other slot too
    }
    any_busy = S->slots[0].busy || S->slots[1].busy;
    if (any_busy) {
      // wait for the launch that started first
      Slot* w = nullptr;
      for (Slot& sl : S->slots)
        if (sl.busy && (!w || sl.seq < w->seq)) w = &sl;
      hipEvent_t ev = w->ev1;
      lk.unlock();
      (void)hipEventSynchronize(ev);
      lk.lock();
      continue;
    }
    S->cv_work.wait(lk, [&] { return S->stop || !S->queue.empty(); });
  }
  S->cv_done.notify_all();
}

// Layers: inputs are layer 0, a gate's layer is 1 + the max layer of its inputs. Gates are
// scheduled by (layer, kind) so each launch part has one gate kind.
int build_request(const hb_circuit_t* c, uint32_t n1, Request* r) {
  const uint32_t ni = c->n_inputs, ng = c->n_gates;
  if (ni == 0 || ng == 0) return fail(HB_ERR_ARG, "circuit needs inputs and gates");
  if (!c->in0 || !c->in1 || (c->n_outputs && !c->outputs))
    return fail(HB_ERR_ARG, "circuit arrays missing");
  std::vector<uint32_t> level(ni + ng, 0);
  for (uint32_t g = 0; g < ng; g++) {
    const uint32_t w = ni + g;
    if (c->in0[g] >= w || c->in1[g] >= w)
      return fail(HB_ERR_ARG, "gate %u reads wire %u/%u not defined before it", g, c->in0[g], c->in1[g]);
    const uint8_t k = c->kind ? c->kind[g] : (uint8_t)GATE_NAND;
    if (k > GATE_XNOR) return fail(HB_ERR_ARG, "gate %u: kind %u not a two-input gate", g, k);
    level[w] = 1 + std::max(level[c->in0[g]], level[c->in1[g]]);
  }
  for (uint32_t k = 0; k < c->n_outputs; k++)
    if (c->outputs[k] >= ni + ng) return fail(HB_ERR_ARG, "output %u out of range", k);
  std::vector<uint32_t> order(ng);
  for (uint32_t g = 0; g < ng; g++) order[g] = g;
  auto kind_of = [&](uint32_t g) { return c->kind ? c->kind[g] : (uint8_t)GATE_NAND; };
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    if (level[ni + a] != level[ni + b]) return level[ni + a] < level[ni + b];
    return kind_of(a) < kind_of(b);
  });
  r->n1 = n1;
  r->n_inputs = ni;
  r->n_gates = ng;
  r->idx.resize((size_t)3 * ng);
  r->kind.resize(ng);
  for (uint32_t s = 0; s < ng; s++) {
    const uint32_t g = order[s];
    r->idx[3 * (size_t)s] = c->in0[g];
    r->idx[3 * (size_t)s + 1] = c->in1[g];
    r->idx[3 * (size_t)s + 2] = ni + g;
    r->kind[s] = kind_of(g);
    const bool boundary = s + 1 == ng || level[ni + order[s + 1]] != level[ni + g] ||
                          kind_of(order[s + 1]) != r->kind[s];
    if (boundary) {
      r->run_end.push_back(s + 1);
      r->run_layer.push_back(level[ni + g]);
    }
  }
  r->outputs.assign(c->outputs, c->outputs + c->n_outputs);
  const size_t wires_bytes = (size_t)(ni + ng) * n1 * 8;
  r->bytes = ((wires_bytes + 255) & ~(size_t)255) + (size_t)3 * ng * 4;
  return HB_OK;
}

void release_slot(Slot& sl) {
  if (sl.parts_dev) (void)hipFree(sl.parts_dev);
  if (sl.parts_host) (void)hipHostFree(sl.parts_host);
  if (sl.acc) (void)hipFree(sl.acc);
  if (sl.pair_flags) (void)hipFree(sl.pair_flags);
  if (sl.pair_status) (void)hipFree(sl.pair_status);
  if (sl.ev0) (void)hipEventDestroy(sl.ev0);
  if (sl.ev1) (void)hipEventDestroy(sl.ev1);
  if (sl.s) (void)hipStreamDestroy(sl.s);
  sl = Slot{};
}

// Device resources of a new server; on failure the caller releases whatever was made.
int server_alloc(hb_server* S, uint64_t hbm_budget, uint32_t max_gates) {
  HIP_TRY(hipMalloc((void**)&S->arena_base, hbm_budget));
  S->arena.init(hbm_budget);
  for (Slot& sl : S->slots) {
    HIP_TRY(hipStreamCreateWithFlags(&sl.s, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&sl.ev0));
    HIP_TRY(hipEventCreate(&sl.ev1));
    HIP_TRY(hipMalloc((void**)&sl.parts_dev, kMaxParts * sizeof(BootArgs)));
    HIP_TRY(hipHostMalloc((void**)&sl.parts_host, kMaxParts * sizeof(BootArgs)));
    HIP_TRY(hipMalloc(&sl.acc, (size_t)max_gates * kAccPerGate));
    const size_t pair_bytes = (size_t)max_gates * kPairFlagWords * sizeof(uint64_t);
    HIP_TRY(hipMalloc((void**)&sl.pair_flags, pair_bytes));
    HIP_TRY(hipMemset(sl.pair_flags, 0, pair_bytes));  // epochs start at 1
    HIP_TRY(hipMalloc((void**)&sl.pair_status, kPairStatusWords * sizeof(uint64_t)));
    HIP_TRY(hipMemset(sl.pair_status, 0, kPairStatusWords * sizeof(uint64_t)));
  }
  HIP_TRY(hipStreamCreateWithFlags(&S->copy, hipStreamNonBlocking));
  return HB_OK;
}

}  // namespace

extern "C" {

int hb_server_create(int device, uint32_t max_gates, uint64_t hbm_budget, hb_server_t** out) {
  if (!out) return fail(HB_ERR_ARG, "null out");
  if (max_gates == 0 || max_gates > (1u << 20)) return fail(HB_ERR_ARG, "bad max_gates %u", max_gates);
  int rc = ensure_engine(device);
  if (rc) return rc;
  std::unique_ptr<hb_server> S(new hb_server);
  S->device = device;
  S->max_gates = max_gates;
  size_t fr = 0, tot = 0;
  HIP_TRY(hipMemGetInfo(&fr, &tot));
  const size_t slot_bytes = 2 * (size_t)max_gates * kAccPerGate;
  if (hbm_budget == 0) hbm_budget = std::min<uint64_t>((fr > slot_bytes ? fr - slot_bytes : 0) / 2,
                                                       32ull << 30);
  if (hbm_budget + slot_bytes > fr)
    return fail(HB_ERR_CAPACITY, "HBM budget %llu exceeds free device memory %zu",
                (unsigned long long)hbm_budget, fr);
  rc = server_alloc(S.get(), hbm_budget, max_gates);
  if (rc) {
    (void)hb_server_destroy(S.release());  // frees what was allocated; keeps the error
    return rc;
  }
  S->st.budget = hbm_budget;
  hb_server* raw = S.release();
  raw->worker = std::thread(worker_loop, raw);
  *out = raw;
  return HB_OK;
}

int hb_server_destroy(hb_server_t* S) {
  if (!S) return HB_OK;
  {
    std::lock_guard<std::mutex> g(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  if (S->worker.joinable()) S->worker.join();
  for (Slot& sl : S->slots) release_slot(sl);
  if (S->copy) (void)hipStreamDestroy(S->copy);
  if (S->arena_base) (void)hipFree(S->arena_base);
  delete S;
  return HB_OK;
}

int hb_server_submit_traced(hb_server_t* S, hb_ek_t* ek, const hb_circuit_t* c,
                            const uint64_t* inputs, int traced, uint64_t* request_id) {
  if (!S || !ek || !c || !inputs || !request_id) return fail(HB_ERR_ARG, "null argument");
  if (ek->p.N != 1024) return fail(HB_ERR_ARG, "unsupported ring dimension");
  std::unique_ptr<Request> r(new Request);
  const uint32_t n1 = ek->p.n + 1;
  int rc = build_request(c, n1, r.get());
  if (rc) return rc;
  if (traced) {
    r->traced = true;
    r->trace_gate_bytes = (size_t)n1 * 2 * ek->p.N * (ek->p.is_goldilocks ? 8 : 4);
    r->trace_off = (r->bytes + 255) & ~(size_t)255;
    r->bytes = r->trace_off + r->trace_gate_bytes * r->n_gates;
  }
  if (r->n_gates > 0 && r->bytes > S->arena.size())
    return fail(HB_ERR_CAPACITY, "circuit needs %zu bytes, HBM budget is %zu", r->bytes,
                S->arena.size());
  for (uint64_t k = 0; k < (uint64_t)c->n_inputs * n1; k++)
    if (inputs[k] >= ek->p.q) return fail(HB_ERR_ARG, "input word %llu not canonical", (unsigned long long)k);
  r->ek = ek;
  r->inputs.assign(inputs, inputs + (size_t)c->n_inputs * n1);
  r->t_submit = now_ms();
  std::lock_guard<std::mutex> g(S->mu);
  if (S->stop) return fail(HB_ERR_STATE, "server is shutting down");
  r->id = S->next_id++;
  *request_id = r->id;
  S->queue.push_back(r.get());
  S->reqs[r->id] = std::move(r);
  S->st.submitted++;
  S->st.queued = S->queue.size();
  S->cv_work.notify_one();
  return HB_OK;
}

int hb_server_submit(hb_server_t* S, hb_ek_t* ek, const hb_circuit_t* c, const uint64_t* inputs,
                     uint64_t* request_id) {
  return hb_server_submit_traced(S, ek, c, inputs, 0, request_id);
}

int hb_server_poll(hb_server_t* S, uint64_t id, int* done) {
  if (!S || !done) return fail(HB_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> g(S->mu);
  auto it = S->reqs.find(id);
  if (it == S->reqs.end()) return fail(HB_ERR_ARG, "unknown request %llu", (unsigned long long)id);
  *done = it->second->finished ? 1 : 0;
  return S->error.empty() ? HB_OK : fail(HB_ERR_DEVICE, "%s", S->error.c_str());
}

int hb_server_wait_traced(hb_server_t* S, uint64_t id, uint64_t* outputs,
                          uint64_t* core_traces, double* latency_ms) {
  if (!S) return fail(HB_ERR_ARG, "null server");
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->reqs.find(id);
  if (it == S->reqs.end()) return fail(HB_ERR_ARG, "unknown request %llu", (unsigned long long)id);
  Request* r = it->second.get();
  S->cv_done.wait(lk, [&] { return r->finished || !S->error.empty(); });
  if (!r->finished) return fail(HB_ERR_DEVICE, "server error: %s", S->error.c_str());
  if (outputs) std::memcpy(outputs, r->out.data(), r->out.size() * 8);
  if (latency_ms) *latency_ms = r->t_done - r->t_submit;
  if (core_traces && r->traced) {
    // schedule order -> gate order; widen BabyBear words
    const size_t words = r->trace_gate_bytes / (r->ek->p.is_goldilocks ? 8 : 4);
    for (uint32_t s = 0; s < r->n_gates; s++) {
      const uint32_t g = r->idx[3 * (size_t)s + 2] - r->n_inputs;
      const uint8_t* src = r->trace_host.data() + (size_t)s * r->trace_gate_bytes;
      uint64_t* dst = core_traces + (size_t)g * words;
      if (r->ek->p.is_goldilocks) {
        std::memcpy(dst, src, words * 8);
      } else {
        const uint32_t* w = reinterpret_cast<const uint32_t*>(src);
        for (size_t k = 0; k < words; k++) dst[k] = w[k];
      }
    }
  }
  S->reqs.erase(it);
  return HB_OK;
}

int hb_server_wait(hb_server_t* S, uint64_t id, uint64_t* outputs, double* latency_ms) {
  return hb_server_wait_traced(S, id, outputs, nullptr, latency_ms);
}

int hb_server_stats(hb_server_t* S, hb_server_stats_t* out) {
  if (!S || !out) return fail(HB_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> g(S->mu);
  *out = S->st;
  out->bytes_in_use = S->arena.used();
  out->queued = S->queue.size();
  out->active = S->active.size();
  // gate-weighted percentiles of the gate latency
  std::vector<std::pair<float, uint32_t>> v = S->gate_lat;
  std::sort(v.begin(), v.end());
  uint64_t total = 0;
  for (auto& x : v) total += x.second;
  auto pct = [&](double q) -> double {
    if (total == 0) return 0.0;
    const uint64_t target = (uint64_t)std::ceil(q * (double)total);
    uint64_t seen = 0;
    for (auto& x : v) {
      seen += x.second;
      if (seen >= target) return x.first;
    }
    return v.back().first;
  };
  out->gate_p50_ms = pct(0.50);
  out->gate_p99_ms = pct(0.99);
  return HB_OK;
}

}  // extern "C"
