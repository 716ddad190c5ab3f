// Internal interface of libgridsim_hip.so shared by the engine (engine.cpp) and the serving loop (server.cpp): device
// allocations, the context and network handles, the device plan, per-instance HBM layout and the batched launch
// steps. Nothing here crosses the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "gridsim_hip.h"
#include "model.h"
#include "nr_kernels.h"

namespace gs {

hipError_t launch_nr_batch(int lu_fp32, int x_in_lds, int pair, const InstDesc* descs, const DevPlan& plan,
                           int count, const NrParams& prm, size_t lds_bytes, hipStream_t stream);
hipError_t launch_seam(const SeamDesc* descs, const DevPlan& plan, int count, size_t lds_bytes, hipStream_t stream);
hipError_t launch_sincos(const double* x, double* s, double* c, int64_t n, hipStream_t stream);
hipError_t launch_iterative_step(const InstDesc* descs, const DevPlan& plan, const NrParams& prm, double* work,
                                 int* iw, hipStream_t stream);
hipError_t launch_screen(const InstDesc* descs, const DevPlan& plan, int count, gs_screen_summary* out,
                         double* loading, double* flows, hipStream_t stream);
hipError_t launch_group(int lu_fp32, int x_in_lds, int waves, const InstDesc* descs, const DevPlan& plan, int count,
                        const NrParams& prm, size_t lds_bytes, hipStream_t stream);
hipError_t launch_wide_start(const InstDesc* descs, const DevPlan& plan, int count, WideCtl* ctl, hipStream_t stream);
hipError_t launch_wide_step(int lu_fp32, const InstDesc* descs, const DevPlan& plan, int count, const NrParams& prm,
                            WideCtl* ctl, const std::vector<int>& level_chunks, const std::vector<int>& fs_level_pass,
                            const std::vector<int>& bs_level_pass, size_t ws_bytes, const WideSplit& split,
                            hipStream_t stream);
// the split point of a plan (side / events left empty): WideSplit::trunk_level = n_levels when there is no trunk
WideSplit wide_split_of(const DevicePlanHost& d);
hipError_t launch_fp64_probe(double* out, hipStream_t stream);

constexpr size_t kLdsLimit = 64 * 1024;  // dynamic LDS a work-group may take for the solve workspace
constexpr size_t kLdsPerCu = 160 * 1024;  // LDS of one compute unit (gfx950)
constexpr size_t kWavesPerCu = 8;         // the batched kernel's register use allows two waves per SIMD
constexpr int kGroupWavesMax = 8;         // waves of a work-group-per-instance launch (two per SIMD)
constexpr int kGroupWavesDefault = 8;

struct HipError {
  gs_status code;
  std::string what;
};

[[noreturn]] inline void fail(gs_status code, const std::string& what) { throw HipError{code, what}; }

inline void check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    fail(GS_OUT_OF_MEMORY, std::string(what) + ": " + hipGetErrorString(e));
  fail(GS_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

// Device allocation owned by a handle.
struct DevMem {
  void* p = nullptr;
  size_t bytes = 0;
  DevMem() = default;
  DevMem(const DevMem&) = delete;
  DevMem& operator=(const DevMem&) = delete;
  ~DevMem() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void ensure(size_t n) {
    if (n <= bytes && p) return;
    release();
    check(hipMalloc(&p, n ? n : 16), "hipMalloc");
    bytes = n ? n : 16;
  }
  template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

template <typename T>
void upload(DevMem& dst, const std::vector<T>& v, hipStream_t s) {
  dst.ensure(sizeof(T) * v.size());
  if (!v.empty()) check(hipMemcpyAsync(dst.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s), "upload");
}

// The captured Newton step of a level-parallel run (one per device plan; rebuilt when the launch it was captured for
// changes), its control block and the pinned copy the host polls.
struct WideGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t side = nullptr;            // the fork's second branch (captured into the graph; default priority)
  hipEvent_t fork = nullptr, join = nullptr;
  const InstDesc* desc = nullptr;
  int count = 0, fp32 = 0;
  NrParams prm{};
  DevMem ctl;
  WideCtl* host = nullptr;
  WideGraph() = default;
  WideGraph(const WideGraph&) = delete;
  WideGraph& operator=(const WideGraph&) = delete;
  ~WideGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
    if (side) (void)hipStreamDestroy(side);
    if (host) (void)hipHostFree(host);
  }
};

// Host image of the device arrays of one plan (or network pattern): every array at a 256-byte aligned offset of one
// byte block, so the whole plan goes to the device in one allocation and one copy, into a block of its own or into
// memory the caller provides. `bind` points a DevPlan's fields into the block once it lies at `base`.
struct PlanImage {
  std::vector<char> bytes;
  std::vector<std::pair<size_t, size_t>> fields;   // (offset of the DevPlan pointer field, offset in the block)
  size_t add(const void* data, size_t n) {
    const size_t off = (bytes.size() + 255) & ~size_t(255);
    bytes.resize(off + (n ? n : 16), 0);   // an empty array still gets an address of its own
    if (n) std::memcpy(bytes.data() + off, data, n);
    return off;
  }
  template <typename P, typename T>
  void put(P DevPlan::*field, const std::vector<T>& v) {
    const size_t off = add(v.data(), sizeof(T) * v.size());
    DevPlan probe{};
    fields.emplace_back((size_t)(reinterpret_cast<char*>(&(probe.*field)) - reinterpret_cast<char*>(&probe)), off);
  }
  void bind(DevPlan& h, char* base) const {
    for (const auto& [field, off] : fields) {
      char* p = base + off;
      std::memcpy(reinterpret_cast<char*>(&h) + field, &p, sizeof(p));
    }
  }
};

// A plan and its network pattern uploaded to the device: one block for the plan image, one for the branch data the
// screening reads.
struct DevicePlan {
  DevMem block, branch_block;
  DevPlan h{};   // passed to the kernels by value
  DevicePlanHost dph;
  std::unique_ptr<WideGraph> wide;
};

struct SeamPlan;

// A map holding at most `cap` entries: inserting past the cap drops the least recently used one.
template <typename K, typename V>
class LruCache {
 public:
  explicit LruCache(size_t cap) : cap_(cap) {}
  V* find(const K& k) {
    auto it = map_.find(k);
    if (it == map_.end()) return nullptr;
    order_.splice(order_.begin(), order_, it->second.second);
    return &it->second.first;
  }
  V& put(const K& k, V v) {
    auto it = map_.find(k);
    if (it != map_.end()) {
      it->second.first = std::move(v);
      order_.splice(order_.begin(), order_, it->second.second);
      return it->second.first;
    }
    while (!map_.empty() && map_.size() >= cap_) {
      map_.erase(order_.back());
      order_.pop_back();
    }
    order_.push_front(k);
    auto ins = map_.emplace(k, std::make_pair(std::move(v), order_.begin())).first;
    return ins->second.first;
  }
  void clear() { map_.clear(); order_.clear(); }
  size_t size() const { return map_.size(); }

 private:
  size_t cap_;
  std::list<K> order_;   // most recently used first
  std::map<K, std::pair<V, typename std::list<K>::iterator>> map_;
};

// Minimum-degree orders by Jacobian pattern, shared by the fresh symbolic passes of paused instances (the order
// depends on the pattern alone, as CpuSparseLuSolver caches it per pattern, CpuSparseLuSolver.cs:30-39). Keyed by a
// 64-bit FNV-1a hash of (m, rowptr, colidx) with the pattern kept to confirm a hit; at most kOrderCacheCap patterns.
// Thread-safe.
class OrderCache {
 public:
  std::vector<int> order(const std::vector<int>& rp, const std::vector<int>& ci, int m);
  int64_t hits() const { return hits_; }
  int64_t misses() const { return misses_; }

 private:
  struct Entry { std::vector<int> rp, ci, order; };
  std::mutex mu_;
  LruCache<uint64_t, std::shared_ptr<Entry>> map_{256};
  int64_t hits_ = 0, misses_ = 0;
};

// The sparse-solver seam keeps the minimum-degree order of its last kSeamOrderCap patterns and the device plans of its
// last kSeamPlanCap structures (pattern and nonzero positions).
constexpr size_t kSeamOrderCap = 256, kSeamPlanCap = 16;

}  // namespace gs

struct gs_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  std::string err;
  size_t budget = 0;
  double us_per_tick = 0.01;  // device wall clock (wall_clock64) period in microseconds
  // key: n, rowptr, colidx -> minimum-degree order
  gs::LruCache<std::vector<int>, std::vector<int>> seam_orders{gs::kSeamOrderCap};
  // key: n, rowptr, colidx, -1, positions of the nonzero values -> the device plan of the last factorisation
  gs::LruCache<std::vector<int>, std::shared_ptr<gs::SeamPlan>> seam_plans{gs::kSeamPlanCap};
  gs::DevMem seam_buf;        // the seam's per-call device buffers (grown to the largest system, reused)
  char* scratch = nullptr;    // device address of the no-allocation entry point's scratch array on this device
  int64_t seam_plan_hits = 0, seam_plan_builds = 0;
};

namespace gs {
// The device buffers one batched run lays its instances out in: instance blocks, descriptors, states, and the base
// specification followed by any instances' own Y values.
struct RunArena {
  DevMem inst_buf, desc_buf, state_buf, aux_buf;
};
}  // namespace gs

struct gs_net {
  gs_ctx* ctx = nullptr;
  gs::Network net;
  gs::YBus y;
  gs::BusSpec spec;
  bool flat_start = false;
  gs::JacobianPattern jp;
  gs::Plan plan;
  gs::DevicePlan dev;
  // per-call device arenas
  gs::RunArena arena;
  gs::DevMem screen_buf, loading_buf, flow_buf;
  gs::OrderCache orders;      // minimum-degree orders of the patterns its paused instances' symbolic passes met
  double last_total_ms = 0, last_kernel_ms = 0;
  int last_launches = 0;
};

namespace gs {

// Byte offsets of one instance's arrays inside its block of HBM (stride bytes per instance).
struct Layout {
  size_t vm, va, p, q, f, dx, r, jv, tpq, vals, dval, xg, spec, start, stride;
};

// Arithmetic and LDS placement of a launch.
struct BatchRun {
  bool fp32 = true, x_lds = true;
  bool pair = false;   // two instances per work-group sharing the LDS solve vector
  int group_waves = 0; // > 0: a work-group of that many waves per instance (GS_SCHEDULE_GROUP)
  size_t lds_bytes = 0;
};

// Host copies that must outlive the asynchronous uploads of a prepared batch.
struct Staging {
  std::vector<InstDesc> descs;
  std::vector<double> aux;                                  // the network's own specification
  std::vector<std::pair<int, std::vector<double>>> stage;   // per instance: [pspec | qspec | vm0 | va0]
  std::vector<InstState> states;                            // resumed states (uploaded when not starting afresh)
  std::vector<std::pair<double*, std::vector<double>>> own_y;  // instances' own Y values and their device places
};

// Device buffers of one prepared run: `count` instance blocks of Layout::stride bytes each, their descriptors and
// states, and the network's own specification (pspec | qspec, 2 n doubles) for instances that do not scale it.
struct RunBuffers {
  char* inst = nullptr;
  InstDesc* desc = nullptr;
  InstState* state = nullptr;
  const double* base_spec = nullptr;
  double* own_y = nullptr;    // room for own_y_count instances' own Y values (2 (nnz_y + n) doubles each)
  int own_y_count = 0;
};

struct PreparedRun {
  BatchRun run;
  Layout L;
  NrParams prm;
  int count = 0;
  const InstDesc* desc = nullptr;
};

void upload_plan(DevicePlan& dp, const YBus* y, const JacobianPattern& jp, const BusSpec* spec, const Plan& plan,
                 double base_mva, hipStream_t s);
gs_plan_stats plan_stats(const DevicePlanHost& d, const YBus& y, const JacobianPattern& jp);
bool symbolic_pass(const YBus& y, const JacobianPattern& jp, const double* vm, const double* va, Plan& plan,
                   std::string& why, double growth_tol, const std::vector<int>* order = nullptr);
Network instance_network(const Network& base, const gs_instance& in);
int symbolic_threads();
gs_nr_options normalise(const gs_nr_options* o);
Layout instance_layout(const DevPlan& h, bool fp32, bool x_lds);
BatchRun plan_run(const DevPlan& h, const gs_nr_options& o, bool force_fp64);
// An instance's outaged branches (no allocation: this runs per instance on every batch call).
struct OutageSet {
  int n = 0;
  int b[kMaxOutages + 1];
  const int* begin() const { return b; }
  const int* end() const { return b + n; }
};
OutageSet outage_set(const gs_instance& in, int n_branch);
constexpr int kMaxOutagePatches = 4 * (kMaxOutages + 1);   // the Y cells an outage set can touch
void check_instances(int n_branch, int count, const gs_instance* inst, const std::string& where);
int outage_patches(const YBus& y, const Network& net, const OutageSet& outaged, Patch* out);
int own_y_count(const YBus* y, const Network* netw, int count, const gs_instance* inst);
void fill_result(const InstState& s, gs_nr_result& r, int refreshes, double us_per_tick);

// Fills descriptors for `count` instances in the buffers rb and uploads them and the instances' own data on stream s
// (the host copies stay in stg until the stream has consumed them).
PreparedRun prepare_into(const DevPlan& h, const RunBuffers& rb, Staging& stg, int count, const gs_instance* inst,
                         const YBus* y, const Network* netw, const gs_nr_options& o, bool plain,
                         const std::vector<InstState>* resume,
                         const std::vector<std::vector<double>>* spec_override,
                         const std::vector<std::vector<double>>* start_override, hipStream_t s);
// The batched launch of a prepared set: every instance in one kernel, or (schedule > 0) launches of at most
// `schedule` instances one after another on the stream. (The level-parallel schedule polls the device between its
// graph replays and runs through launch_instances only.)
void launch_prepared(const DevPlan& h, const PreparedRun& pr, int schedule, hipStream_t s);
// A paused instance (a static pivot failed its growth check) finished by fresh host symbolic passes and resumed
// launches; vm / va (in and out) hold its state.
// types / qspec (or null): bus types and Q specification other than the network's (a reactive-limit pass).
// The host half of one paused instance's first fresh symbolic pass: its own network (outages out, loads scaled), Y-bus,
// specification, Jacobian pattern and the plan at its paused state (ok false: the LU refuses that state).
struct FreshSymbolic {
  Network own;
  YBus y;
  BusSpec spec;
  JacobianPattern jp;
  Plan plan;
  bool ok = false;
  std::string why;
};
// FreshSymbolic of `count` instances at their states (vm / va per instance) on up to `threads` host threads, the
// orders from `cache`. types / qspec (or null): bus types and Q specification other than the network's.
void fresh_symbolic(const Network& base, bool flat_start, const gs_nr_options& o, int count, const gs_instance* inst,
                    const std::vector<const double*>& vm, const std::vector<const double*>& va,
                    const std::vector<int>* types, const std::vector<double>* qspec, int threads, OrderCache& cache,
                    std::vector<FreshSymbolic>& out);
void finish_paused(gs_net* net, const gs_instance& in, const gs_nr_options& o, bool plain, InstState st,
                   std::vector<double>& vm, std::vector<double>& va, gs_nr_result& result, double& kms, int& launches,
                   std::vector<double>* p_out = nullptr, std::vector<double>* q_out = nullptr,
                   const std::vector<int>* types = nullptr, const std::vector<double>* qspec = nullptr,
                   InstState* st_out = nullptr, FreshSymbolic* first = nullptr);
// The branch data screening reads (end buses, stamps, rate A) and the buses' own voltage limits, uploaded with a
// network.
void upload_branches(DevicePlan& dp, const Network& net, hipStream_t s);
// The host image upload_plan copies to the device (scalar fields of h set, pointer fields recorded for img.bind).
void build_plan_image(PlanImage& img, DevPlan& h, DevicePlanHost& d, const YBus* y, const JacobianPattern& jp,
                      const BusSpec* spec, const Plan& plan, double base_mva);

// Runs f under the context's lock, turning exceptions into status codes and the context's last error.
template <typename F>
gs_status guarded(gs_ctx* ctx, F&& f) {
  if (!ctx) return GS_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  try {
    return f();
  } catch (const HipError& e) {
    ctx->err = e.what;
    return e.code;
  } catch (const std::bad_alloc&) {
    ctx->err = "host allocation failed";
    return GS_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return GS_DEVICE_ERROR;
  }
}

}  // namespace gs
