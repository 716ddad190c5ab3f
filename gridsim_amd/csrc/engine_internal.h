// Internal interface of libgridsim_hip.so shared by the engine (engine.cpp) and the serving loop (server.cpp): device
// allocations, the context and network handles, the device plan, per-instance HBM layout and the batched launch
// steps. Nothing here crosses the C ABI.
#pragma once
#include <hip/hip_runtime.h>

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

// A plan and its network pattern uploaded to the device.
struct DevicePlan {
  DevMem y_rp, y_ij, y_diag, y_gb, slots, aslots, ppos, qpos, j_rp, j_ci, vm0, va0;
  DevMem pinv, q, prow, piv_pos;
  DevMem rs_rec, ch_a_src;
  DevMem fs_rec, bs_rec;
  DevMem rs_chunk_row, fs_pass_step, bs_pass_step;
  DevMem br_ends, br_stamp, br_rate, bus_lo, bus_hi, j_diag;
  DevPlan h{};   // passed to the kernels by value
  DevicePlanHost dph;
  std::unique_ptr<WideGraph> wide;
};

struct SeamPlan;

}  // namespace gs

struct gs_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  std::string err;
  size_t budget = 0;
  double us_per_tick = 0.01;  // device wall clock (wall_clock64) period in microseconds
  std::map<std::pair<std::vector<int>, std::vector<int>>, std::vector<int>> seam_orders;  // pattern -> MD order
  std::map<std::vector<int>, std::shared_ptr<gs::SeamPlan>> seam_plans;                   // nonzero positions -> plan
};

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
  gs::DevMem inst_buf, desc_buf, state_buf, aux_buf;
  gs::DevMem screen_buf, loading_buf, flow_buf;
  double last_total_ms = 0, last_kernel_ms = 0;
  int last_launches = 0;
};

namespace gs {

// Byte offsets of one instance's arrays inside its block of HBM (stride bytes per instance).
struct Layout {
  size_t vm, va, p, q, f, dx, r, jv, tpq, vals, fval, bval, dval, xg, spec, start, stride;
};

// Arithmetic and LDS placement of a launch.
struct BatchRun {
  bool fp32 = true, x_lds = true;
  bool pair = false;   // two instances per work-group sharing the LDS solve vector
  size_t lds_bytes = 0;
};

// Host copies that must outlive the asynchronous uploads of a prepared batch.
struct Staging {
  std::vector<InstDesc> descs;
  std::vector<double> aux;                                  // the network's own specification
  std::vector<std::pair<int, std::vector<double>>> stage;   // per instance: [pspec | qspec | vm0 | va0]
  std::vector<InstState> states;                            // resumed states (uploaded when not starting afresh)
};

// Device buffers of one prepared run: `count` instance blocks of Layout::stride bytes each, their descriptors and
// states, and the network's own specification (pspec | qspec, 2 n doubles) for instances that do not scale it.
struct RunBuffers {
  char* inst = nullptr;
  InstDesc* desc = nullptr;
  InstState* state = nullptr;
  const double* base_spec = nullptr;
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
                   std::string& why);
gs_nr_options normalise(const gs_nr_options* o);
Layout instance_layout(const DevPlan& h, bool fp32, bool x_lds);
BatchRun plan_run(const DevPlan& h, const gs_nr_options& o, bool force_fp64);
int outage_patches(const YBus& y, const Network& net, int branch, Patch* out);
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
void finish_paused(gs_net* net, const gs_instance& in, const gs_nr_options& o, bool plain, InstState st,
                   std::vector<double>& vm, std::vector<double>& va, gs_nr_result& result, double& kms, int& launches,
                   std::vector<double>* p_out = nullptr, std::vector<double>* q_out = nullptr,
                   const std::vector<int>* types = nullptr, const std::vector<double>* qspec = nullptr);
// The branch data screening reads (end buses, stamps, rate A) and the buses' own voltage limits, uploaded with a
// network.
void upload_branches(DevicePlan& dp, const Network& net, std::vector<std::vector<double>>& keep, hipStream_t s);

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
