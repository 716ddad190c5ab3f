// Serving loop of libgridsim_hip.so: continuous batching of power-flow and contingency requests over resident networks.
//
// The reference answers gridsim-rpc/1 "solve" requests one after another on one thread (RpcCommand.cs:460-497 reads a
// line, :668-700 solves it to completion, writes the response, reads the next). Here a dispatcher thread owns up to
// four slots, each with its own stream. A slot is a running batch of one network: a block of instance positions in the
// HBM state arena (taken by memory: the slot halves its size until the arena holds it), each holding one request or
// nothing. Every launch of the batched Newton kernel on the slot runs a few Newton steps of each instance in it (as
// many as keep a launch near a millisecond); between launches the dispatcher publishes the instances that finished,
// and the waiting requests of the same network take the free positions and start afresh in the next launch. A request
// therefore joins the GPU's work at the next Newton-step boundary of a running batch of its network, or in a free slot,
// and leaves as soon as its own solve ends. Each instance is solved by its own wavefront (or work-group), so its result
// does not depend on what runs beside it or when it joined.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <thread>
#include <unordered_map>

#include "engine_internal.h"

using namespace gs;

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

constexpr size_t kAlign = 256;
size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

// First-fit sub-allocator over [0, size) with coalescing free blocks (offsets and sizes multiples of kAlign).
class Arena {
 public:
  void reset(size_t size) {
    free_.clear();
    if (size) free_[0] = size;
    size_ = size;
    used_ = 0;
  }
  // Offset of a block of `bytes`, or SIZE_MAX when no free block is large enough.
  size_t alloc(size_t bytes) {
    bytes = align_up(std::max<size_t>(bytes, 1));
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      size_t off = it->first, rest = it->second - bytes;
      free_.erase(it);
      if (rest) free_[off + bytes] = rest;
      used_ += bytes;
      return off;
    }
    return SIZE_MAX;
  }
  void release(size_t off, size_t bytes) {
    bytes = align_up(std::max<size_t>(bytes, 1));
    used_ -= bytes;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + bytes == next->first) {
      bytes += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += bytes;
        return;
      }
    }
    free_[off] = bytes;
  }
  size_t size() const { return size_; }
  size_t used() const { return used_; }

 private:
  std::map<size_t, size_t> free_;
  size_t size_ = 0, used_ = 0;
};

struct Request {
  uint64_t ticket = 0;
  int net = 0;
  gs_instance inst{};
  std::vector<double> start;  // vm0 | va0 when the request brought its own start point
  std::vector<int32_t> outages;  // the request's outage list (gs_instance::outages points here when launched)
  Clock::time_point submitted;
};

struct Finished {
  gs_status status = GS_OK;
  gs_nr_result result{};
  std::vector<double> vm, va, p, q;   // p, q: per unit
  double base_mva = 100.0;
  double latency_ms = 0;
};

}  // namespace

struct gs_server {
  gs_ctx* ctx = nullptr;
  gs_server_config cfg{};

  struct Net {
    gs_net* net = nullptr;
    gs_nr_options opts{};
    BatchRun run;
    Layout L{};
    DevMem base_spec;   // pspec | qspec of the network as uploaded
    size_t state_span = 0;  // bytes of one instance's vm, va, p, q region copied back
  };
  std::vector<std::unique_ptr<Net>> nets;

  // One slot: a running batch of one network. Its device block holds `cap` positions (instance blocks, then the
  // descriptors, then the states), its pinned block their states and voltage spans. A free position's state says
  // finished, so the kernel returns at once for it.
  struct Slot {
    enum Phase { kIdle, kStepping, kCopying };
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    Phase phase = kIdle;
    int net = -1;
    const Net* nt = nullptr;   // srv->nets[net], taken under the lock (the vector may grow meanwhile)
    int cap = 0, live = 0;
    int steps = 1;             // Newton steps per launch
    std::vector<Request> pos;  // per position: its request (ticket 0: free)
    std::vector<Staging> stg;  // per position: the host copies its last upload reads
    std::vector<int> joining;  // positions given new requests, not yet uploaded
    std::vector<int> leaving;  // positions whose instances ended in the last launch
    std::vector<InstDesc> free_desc;
    std::vector<InstState> free_state;
    size_t dev_off = 0, dev_bytes = 0, host_off = 0, host_bytes = 0;
    PreparedRun pr;
    bool pr_ready = false;
    Clock::time_point launched;
  };
  std::vector<Slot> slots;

  DevMem dev_arena;
  Arena dev_alloc;
  char* pinned = nullptr;
  Arena host_alloc;

  std::mutex mu;
  std::condition_variable work_cv, done_cv;
  std::deque<Request> queue;
  std::unordered_map<uint64_t, Finished> finished;
  std::unordered_map<uint64_t, int> in_flight_tickets;  // ticket -> slot
  uint64_t next_ticket = 1;
  bool stopping = false;
  std::thread worker;

  // statistics since the last reset
  int64_t submitted = 0, completed = 0, rejected = 0, launches = 0, launched_instances = 0, joined = 0;
  std::vector<double> latencies;
  double busy_ms = 0;
  Clock::time_point busy_since;
  int busy_slots = 0;
  size_t state_peak = 0;

  ~gs_server() {
    for (Slot& s : slots) {
      if (s.done) (void)hipEventDestroy(s.done);
      if (s.stream) (void)hipStreamDestroy(s.stream);
    }
    if (pinned) (void)hipHostFree(pinned);
  }
};

namespace {

using Slot = gs_server::Slot;

// A launch should take about this long: long enough that launching costs little beside it, short enough that a
// waiting request soon finds a step boundary.
constexpr double kStepTargetMs = 1.0;
constexpr int kMaxStepsPerLaunch = 8;
// a slot starts with room for at least this many requests (later arrivals of its network join into it)
constexpr int kMinSlotCap = 16;

// Byte sizes of a slot of `count` positions of network nt: device block [instances | descriptors | states] and
// pinned block [states | per position the vm..q span].
size_t device_block(const gs_server::Net& nt, int count) {
  return nt.L.stride * (size_t)count + align_up(sizeof(InstDesc) * (size_t)count) +
         align_up(sizeof(InstState) * (size_t)count);
}
size_t host_block(const gs_server::Net& nt, int count) {
  return align_up(sizeof(InstState) * (size_t)count) + nt.state_span * (size_t)count;
}

char* slot_inst(gs_server* srv, const Slot& s) { return srv->dev_arena.as<char>() + s.dev_off; }
InstDesc* slot_desc(gs_server* srv, const Slot& s) {
  return reinterpret_cast<InstDesc*>(slot_inst(srv, s) + s.nt->L.stride * (size_t)s.cap);
}
InstState* slot_state(gs_server* srv, const Slot& s) {
  return reinterpret_cast<InstState*>(reinterpret_cast<char*>(slot_desc(srv, s)) +
                                      align_up(sizeof(InstDesc) * (size_t)s.cap));
}
const InstState* pinned_state(gs_server* srv, const Slot& s) {
  return reinterpret_cast<const InstState*>(srv->pinned + s.host_off);
}
char* pinned_span(gs_server* srv, const Slot& s, int p) {
  return srv->pinned + s.host_off + align_up(sizeof(InstState) * (size_t)s.cap) + s.nt->state_span * (size_t)p;
}

// A new slot's device state: every position free (its descriptor names its state, the state says finished).
void clear_slot(gs_server* srv, Slot& s) {
  s.free_desc.assign(s.cap, InstDesc{});
  s.free_state.assign(s.cap, InstState{});
  InstState* st = slot_state(srv, s);
  for (int p = 0; p < s.cap; p++) {
    s.free_desc[p].state = st + p;
    s.free_state[p].done = 1;
  }
  check(hipMemcpyAsync(slot_desc(srv, s), s.free_desc.data(), sizeof(InstDesc) * (size_t)s.cap,
                       hipMemcpyHostToDevice, s.stream), "server descriptors");
  check(hipMemcpyAsync(st, s.free_state.data(), sizeof(InstState) * (size_t)s.cap, hipMemcpyHostToDevice, s.stream),
        "server states");
}

// Lays the requests of s.joining out at their positions, each starting afresh at the next launch.
void place_joining(gs_server* srv, Slot& s) {
  const gs_server::Net& nt = *s.nt;
  const DevPlan& h = nt.net->dev.h;
  std::vector<InstState> fresh(1);
  fresh[0].done = kDoneFresh;
  for (int p : s.joining) {
    const Request& r = s.pos[p];
    RunBuffers rb;
    rb.inst = slot_inst(srv, s) + nt.L.stride * (size_t)p;
    rb.desc = slot_desc(srv, s) + p;
    rb.state = slot_state(srv, s) + p;
    rb.base_spec = nt.base_spec.as<double>();
    gs_instance in = r.inst;
    in.outages = r.outages.empty() ? nullptr : r.outages.data();
    in.vm0 = r.start.empty() ? nullptr : r.start.data();
    in.va0 = r.start.empty() ? nullptr : r.start.data() + h.n;
    PreparedRun one = prepare_into(h, rb, s.stg[p], 1, &in, &nt.net->y, &nt.net->net, nt.opts, false, &fresh, nullptr,
                                   nullptr, s.stream);
    if (!s.pr_ready) {
      // the launch over all positions: the same run and parameters, resuming (a fresh instance starts by its state)
      s.pr = one;
      s.pr.count = s.cap;
      s.pr.desc = slot_desc(srv, s);
      s.pr.prm.init = 0;
      s.pr.prm.load_start = 0;
      s.pr.prm.trust_first = 0;
      s.pr_ready = true;
    }
  }
  s.joining.clear();
}

// Uploads the joining requests and launches the slot's next steps, then copies every position's state back.
void launch_steps(gs_server* srv, Slot& s) {
  place_joining(srv, s);
  s.pr.prm.steps_this_launch = s.steps;
  launch_prepared(s.nt->net->dev.h, s.pr, s.nt->opts.schedule, s.stream);
  check(hipMemcpyAsync(srv->pinned + s.host_off, slot_state(srv, s), sizeof(InstState) * (size_t)s.cap,
                       hipMemcpyDeviceToHost, s.stream), "server states back");
  check(hipEventRecord(s.done, s.stream), "server event");
  s.launched = Clock::now();
  s.phase = Slot::kStepping;
}

// After a launch: the positions whose instances ended (finished, or paused for a fresh symbolic pass) get their
// voltages copied back. Returns whether there is anything to copy (the slot then waits for its event again).
bool collect_leaving(gs_server* srv, Slot& s) {
  const InstState* st = pinned_state(srv, s);
  s.leaving.clear();
  for (int p = 0; p < s.cap; p++)
    if (s.pos[p].ticket && st[p].done != 0) s.leaving.push_back(p);
  if (s.leaving.empty()) return false;
  const Layout& L = s.nt->L;
  for (int p : s.leaving)
    check(hipMemcpyAsync(pinned_span(srv, s, p), slot_inst(srv, s) + L.stride * (size_t)p + L.vm, s.nt->state_span,
                         hipMemcpyDeviceToHost, s.stream), "server voltages");
  check(hipEventRecord(s.done, s.stream), "server event");
  s.phase = Slot::kCopying;
  return true;
}

// The results of the leaving positions (their voltages are in pinned memory): finished instances as they are, paused
// ones through fresh symbolic passes. Called without the server lock.
std::vector<std::pair<uint64_t, Finished>> results_of_leaving(gs_server* srv, Slot& s, gs_status status) {
  std::vector<std::pair<uint64_t, Finished>> out;
  const gs_server::Net& nt = *s.nt;
  const int n = nt.net->net.n;
  const InstState* st = pinned_state(srv, s);
  for (int p : s.leaving) {
    Finished f;
    f.status = status;
    f.base_mva = nt.net->net.base;
    if (status == GS_OK) {
      const char* span = pinned_span(srv, s, p);
      auto arr = [&](size_t off) { return reinterpret_cast<const double*>(span + (off - nt.L.vm)); };
      f.vm.assign(arr(nt.L.vm), arr(nt.L.vm) + n);
      f.va.assign(arr(nt.L.va), arr(nt.L.va) + n);
      f.p.assign(arr(nt.L.p), arr(nt.L.p) + n);
      f.q.assign(arr(nt.L.q), arr(nt.L.q) + n);
      if (st[p].done == 1) {
        fill_result(st[p], f.result, 0, srv->ctx->us_per_tick);
      } else {
        try {
          double kms = 0;
          int nl = 0;
          gs_instance in = s.pos[p].inst;
          in.vm0 = in.va0 = nullptr;
          in.outages = s.pos[p].outages.empty() ? nullptr : s.pos[p].outages.data();
          std::vector<double> pp, qq;
          // the fresh symbolic passes and resumed launches run on the context's stream: under the context's lock, as
          // every other call on the context, so they never land inside a capture a client has begun on that stream
          std::lock_guard<std::mutex> ctx_lock(srv->ctx->mu);
          finish_paused(nt.net, in, nt.opts, false, st[p], f.vm, f.va, f.result, kms, nl, &pp, &qq);
          if ((int)pp.size() == n) f.p = pp;
          if ((int)qq.size() == n) f.q = qq;
        } catch (const HipError& e) {
          f.status = e.code;
        }
      }
    }
    out.emplace_back(s.pos[p].ticket, std::move(f));
  }
  return out;
}

// Under the lock: publishes results, frees their positions.
void publish(gs_server* srv, Slot& s, std::vector<std::pair<uint64_t, Finished>>& results) {
  const auto now = Clock::now();
  for (auto& tf : results) {
    const uint64_t ticket = tf.first;
    Finished& f = tf.second;
    for (Request& r : s.pos) {
      if (r.ticket != ticket) continue;
      f.latency_ms = ms_between(r.submitted, now);
      r = Request{};
      s.live--;
      break;
    }
    srv->latencies.push_back(f.latency_ms);
    srv->in_flight_tickets.erase(ticket);
    srv->finished[ticket] = std::move(f);
    srv->completed++;
  }
  s.leaving.clear();
  if (!results.empty()) srv->done_cv.notify_all();
}

// Under the lock: the waiting requests of the slot's network take its free positions (arrival order). A slot stops
// taking its network's requests while an older request of another network waits for a slot and none is idle, so that
// the oldest request is served first.
void take_waiting(gs_server* srv, Slot& s) {
  if (srv->queue.empty() || s.live == s.cap) return;
  if (srv->queue.front().net != s.net) {
    bool idle = false;
    for (const Slot& o : srv->slots) idle |= o.phase == Slot::kIdle;
    if (!idle) return;
  }
  const bool running = s.live > 0;   // the slot's other instances are between Newton steps
  int p = 0;
  for (auto it = srv->queue.begin(); it != srv->queue.end() && s.live < s.cap;) {
    if (it->net != s.net) { ++it; continue; }
    while (s.pos[p].ticket) p++;
    s.pos[p] = std::move(*it);
    it = srv->queue.erase(it);
    srv->in_flight_tickets[s.pos[p].ticket] = (int)(&s - srv->slots.data());
    s.joining.push_back(p);
    s.live++;
    srv->joined += running;
  }
}

// Under the lock: an idle slot takes the network of the oldest waiting request, with room for that network's waiting
// requests (at least kMinSlotCap, at most max_batch), halved until the two arenas hold it. False when nothing fits.
bool open_slot(gs_server* srv, Slot& s) {
  if (srv->queue.empty()) return false;
  const int net = srv->queue.front().net;
  const gs_server::Net& nt = *srv->nets[net];
  int waiting = 0;
  for (const Request& r : srv->queue) waiting += r.net == net;
  int cap = std::min(srv->cfg.max_batch, std::max(waiting, kMinSlotCap));
  for (; cap >= 1; cap = cap > 1 ? cap / 2 : 0) {
    const size_t db = device_block(nt, cap), hb = host_block(nt, cap);
    const size_t doff = srv->dev_alloc.alloc(db);
    if (doff == SIZE_MAX) continue;
    const size_t hoff = srv->host_alloc.alloc(hb);
    if (hoff == SIZE_MAX) {
      srv->dev_alloc.release(doff, db);
      continue;
    }
    s.net = net;
    s.nt = &nt;
    s.cap = cap;
    s.live = 0;
    s.steps = 1;
    s.pos.assign(cap, Request{});
    s.stg.assign(cap, Staging{});
    s.joining.clear();
    s.leaving.clear();
    s.pr_ready = false;
    s.dev_off = doff; s.dev_bytes = db;
    s.host_off = hoff; s.host_bytes = hb;
    srv->state_peak = std::max(srv->state_peak, srv->dev_alloc.used());
    return true;
  }
  return false;  // the arenas are full: wait for a slot to drain
}

void close_slot(gs_server* srv, Slot& s) {
  srv->dev_alloc.release(s.dev_off, s.dev_bytes);
  srv->host_alloc.release(s.host_off, s.host_bytes);
  s.pos.clear();
  s.stg.clear();
  s.net = -1;
  s.nt = nullptr;
  s.phase = Slot::kIdle;
  if (--srv->busy_slots == 0) srv->busy_ms += ms_between(srv->busy_since, Clock::now());
}

// Under the lock: the slot at a step boundary (its last launch and copies done, results published): it takes waiting
// requests and launches its next steps, or closes when nothing is left in it. Unlocks around the launch.
void next_steps(gs_server* srv, Slot& s, std::unique_lock<std::mutex>& lock) {
  take_waiting(srv, s);
  if (s.live == 0) {
    close_slot(srv, s);
    return;
  }
  srv->launches++;
  srv->launched_instances += s.live;
  lock.unlock();
  gs_status st = GS_OK;
  try {
    launch_steps(srv, s);
  } catch (const HipError& e) {
    st = e.code;
  }
  lock.lock();
  if (st != GS_OK) {
    // the launch never reached the device: fail the slot's requests now
    std::vector<std::pair<uint64_t, Finished>> failed;
    for (const Request& r : s.pos)
      if (r.ticket) {
        Finished f;
        f.status = st;
        failed.emplace_back(r.ticket, std::move(f));
      }
    publish(srv, s, failed);
    close_slot(srv, s);
  }
}

void dispatcher(gs_server* srv) {
  (void)hipSetDevice(srv->ctx->device);
  std::unique_lock<std::mutex> lock(srv->mu);
  for (;;) {
    // slots whose launch or copies have completed
    for (Slot& s : srv->slots) {
      if (s.phase == Slot::kIdle) continue;
      const hipError_t q = hipEventQuery(s.done);
      if (q == hipErrorNotReady) continue;
      gs_status status = q == hipSuccess ? GS_OK : GS_DEVICE_ERROR;
      lock.unlock();
      std::vector<std::pair<uint64_t, Finished>> results;
      bool copying = false;
      // a failed launch or copy fails every request the slot holds
      auto fail_all = [&](gs_status st) {
        status = st;
        s.leaving.clear();
        for (int p = 0; p < s.cap; p++)
          if (s.pos[p].ticket) s.leaving.push_back(p);
        results = results_of_leaving(srv, s, st);
      };
      try {
        if (status != GS_OK) {
          fail_all(status);
        } else {
          if (s.phase == Slot::kStepping) {
            // adapt the steps per launch to keep launches near the target length
            const double ms = ms_between(s.launched, Clock::now());
            if (ms < 0.5 * kStepTargetMs && s.steps < kMaxStepsPerLaunch) s.steps *= 2;
            else if (ms > 2.0 * kStepTargetMs && s.steps > 1) s.steps /= 2;
            copying = collect_leaving(srv, s);
          }
          if (!copying) results = results_of_leaving(srv, s, GS_OK);
        }
      } catch (const HipError& e) {
        copying = false;
        fail_all(e.code);
      }
      lock.lock();
      if (copying) continue;
      publish(srv, s, results);
      if (status != GS_OK) {
        close_slot(srv, s);
        continue;
      }
      next_steps(srv, s, lock);
    }
    // idle slots take the oldest waiting request's network
    for (Slot& s : srv->slots) {
      if (s.phase != Slot::kIdle || !open_slot(srv, s)) continue;
      if (srv->busy_slots++ == 0) srv->busy_since = Clock::now();
      s.phase = Slot::kStepping;   // (not idle while it is being set up)
      lock.unlock();
      gs_status st = GS_OK;
      try {
        clear_slot(srv, s);
      } catch (const HipError& e) {
        st = e.code;
      }
      lock.lock();
      if (st != GS_OK) {
        close_slot(srv, s);
        continue;
      }
      next_steps(srv, s, lock);
    }
    if (srv->stopping && srv->queue.empty() && srv->busy_slots == 0) break;
    if (srv->busy_slots == 0) {
      srv->work_cv.wait(lock, [&] { return srv->stopping || !srv->queue.empty(); });
    } else {
      // launches in flight: poll their events, waking early for new requests
      srv->work_cv.wait_for(lock, std::chrono::microseconds(20));
    }
  }
}

double percentile(std::vector<double> v, double p) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  double pos = p / 100.0 * (double)(v.size() - 1);
  size_t lo = (size_t)pos, hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
}

}  // namespace

extern "C" {

GS_API gs_status gs_server_create(gs_ctx* ctx, const gs_server_config* cfg, gs_server** out) {
  if (!out) return GS_INVALID_ARGUMENT;
  *out = nullptr;
  return guarded(ctx, [&]() -> gs_status {
    auto srv = std::make_unique<gs_server>();
    srv->ctx = ctx;
    gs_server_config c = cfg ? *cfg : gs_server_config{};
    if (c.max_batch < 0 || c.streams < 0 || c.streams > 4 || c.queue_capacity < 0)
      fail(GS_INVALID_ARGUMENT, "server config: max_batch >= 0, streams in 0..4, queue_capacity >= 0");
    if (c.max_batch == 0) c.max_batch = 256;
    if (c.streams == 0) c.streams = 4;
    if (c.queue_capacity == 0) c.queue_capacity = 65536;
    if (c.state_bytes == 0) c.state_bytes = size_t(4) << 30;
    if (c.pinned_bytes == 0) c.pinned_bytes = size_t(256) << 20;
    srv->cfg = c;
    check(hipSetDevice(ctx->device), "hipSetDevice");
    srv->dev_arena.ensure(c.state_bytes);
    srv->dev_alloc.reset(c.state_bytes & ~(kAlign - 1));
    check(hipHostMalloc((void**)&srv->pinned, c.pinned_bytes, hipHostMallocDefault), "pinned results");
    srv->host_alloc.reset(c.pinned_bytes & ~(kAlign - 1));
    srv->slots.resize(c.streams);
    for (auto& s : srv->slots) {
      check(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "server stream");
      check(hipEventCreateWithFlags(&s.done, hipEventDisableTiming), "server event");
    }
    gs_server* raw = srv.release();
    raw->worker = std::thread(dispatcher, raw);
    *out = raw;
    return GS_OK;
  });
}

GS_API gs_status gs_server_add_network(gs_server* srv, gs_net* net, const gs_nr_options* opts, int32_t* network_id) {
  if (!srv || !net || !network_id || net->ctx != srv->ctx) return GS_INVALID_ARGUMENT;
  return guarded(srv->ctx, [&]() -> gs_status {
    auto nt = std::make_unique<gs_server::Net>();
    nt->net = net;
    nt->opts = normalise(opts);
    if (nt->opts.schedule == GS_SCHEDULE_LEVELS)
      fail(GS_INVALID_ARGUMENT, "the serving loop launches a wavefront or a work-group per instance, not the "
                                "level-parallel schedule");
    if (nt->opts.flags & (GS_FLAG_Q_LIMITS | GS_FLAG_DC_RETRY))
      fail(GS_INVALID_ARGUMENT, "the serving loop runs neither the reactive-limit passes nor the DC-start retry");
    nt->run = plan_run(net->dev.h, nt->opts, false);
    nt->L = instance_layout(net->dev.h, nt->run.fp32, nt->run.x_lds);
    // vm, va, p, q lie in that order in an instance's block: one strided copy brings all four back
    static_assert(sizeof(double) == 8, "double layout");
    if (!(nt->L.vm < nt->L.va && nt->L.va < nt->L.p && nt->L.p < nt->L.q)) fail(GS_DEVICE_ERROR, "instance layout");
    nt->state_span = nt->L.q - nt->L.vm + 8 * (size_t)net->net.n;
    if (device_block(*nt, 1) > srv->dev_alloc.size() || host_block(*nt, 1) > srv->host_alloc.size())
      fail(GS_OUT_OF_MEMORY, "one instance of this network exceeds the server's state or pinned budget");
    std::vector<double> spec(net->spec.pspec);
    spec.insert(spec.end(), net->spec.qspec.begin(), net->spec.qspec.end());
    upload(nt->base_spec, spec, srv->ctx->stream);
    check(hipStreamSynchronize(srv->ctx->stream), "server spec upload");
    std::lock_guard<std::mutex> lock(srv->mu);
    *network_id = (int32_t)srv->nets.size();
    srv->nets.push_back(std::move(nt));
    return GS_OK;
  });
}

GS_API gs_status gs_server_submit(gs_server* srv, int32_t network_id, const gs_instance* inst, uint64_t* ticket) {
  if (!srv || !inst || !ticket) return GS_INVALID_ARGUMENT;
  Request r;
  {
    std::lock_guard<std::mutex> lock(srv->mu);
    if (network_id < 0 || network_id >= (int32_t)srv->nets.size()) return GS_INVALID_ARGUMENT;
    const gs_net* net = srv->nets[network_id]->net;
    try {
      check_instances(net->net.nb, 1, inst, "");
      // an outage touching more Y cells than a descriptor patches needs an array of its own: not in the serving loop
      if (own_y_count(&net->y, &net->net, 1, inst) > 0) return GS_INVALID_ARGUMENT;
    } catch (const HipError&) {
      return GS_INVALID_ARGUMENT;
    }
    if ((int64_t)srv->queue.size() >= srv->cfg.queue_capacity) {
      srv->rejected++;
      return GS_BUSY;
    }
    r.net = network_id;
    r.inst = *inst;
    if (inst->vm0) {
      const int n = net->net.n;
      r.start.assign(inst->vm0, inst->vm0 + n);
      r.start.insert(r.start.end(), inst->va0, inst->va0 + n);
    }
    r.inst.vm0 = r.inst.va0 = nullptr;
    // the outage list is the caller's: keep a copy
    r.outages.assign(inst->outages, inst->outages + inst->n_outages);
    r.inst.outages = nullptr;
    r.ticket = srv->next_ticket++;
    r.submitted = Clock::now();
    *ticket = r.ticket;
    srv->queue.push_back(std::move(r));
    srv->submitted++;
  }
  srv->work_cv.notify_one();
  return GS_OK;
}

GS_API gs_status gs_server_wait(gs_server* srv, uint64_t ticket, int32_t timeout_ms, gs_nr_result* result,
                                double* vm, double* va, double* p_inj, double* q_inj, double* latency_ms) {
  if (!srv) return GS_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lock(srv->mu);
  auto known = [&] {
    if (srv->finished.count(ticket) || srv->in_flight_tickets.count(ticket)) return true;
    for (const Request& r : srv->queue)
      if (r.ticket == ticket) return true;
    return false;
  };
  if (!known()) return GS_INVALID_ARGUMENT;
  auto ready = [&] { return srv->finished.count(ticket) > 0; };
  if (timeout_ms < 0) srv->done_cv.wait(lock, ready);
  else if (!srv->done_cv.wait_for(lock, std::chrono::milliseconds(timeout_ms), ready)) return GS_PENDING;
  Finished f = std::move(srv->finished[ticket]);
  srv->finished.erase(ticket);
  lock.unlock();
  if (f.status != GS_OK) return f.status;
  if (result) *result = f.result;
  if (vm) std::copy(f.vm.begin(), f.vm.end(), vm);
  if (va) std::copy(f.va.begin(), f.va.end(), va);
  // injections in MW / MVAr, as PowerFlowResult.PInjMw / QInjMvar
  if (p_inj) for (size_t i = 0; i < f.p.size(); i++) p_inj[i] = f.p[i] * f.base_mva;
  if (q_inj) for (size_t i = 0; i < f.q.size(); i++) q_inj[i] = f.q[i] * f.base_mva;
  if (latency_ms) *latency_ms = f.latency_ms;
  return GS_OK;
}

GS_API gs_status gs_server_stats_get(gs_server* srv, gs_server_stats* out) {
  if (!srv || !out) return GS_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(srv->mu);
  *out = gs_server_stats{};
  out->submitted = srv->submitted;
  out->completed = srv->completed;
  out->rejected = srv->rejected;
  out->launches = srv->launches;
  out->queued = (int64_t)srv->queue.size();
  out->in_flight = (int64_t)srv->in_flight_tickets.size();
  out->mean_batch = srv->launches ? (double)srv->launched_instances / (double)srv->launches : 0.0;
  out->latency_p50_ms = percentile(srv->latencies, 50);
  out->latency_p99_ms = percentile(srv->latencies, 99);
  out->latency_max_ms = srv->latencies.empty() ? 0.0 : *std::max_element(srv->latencies.begin(),
                                                                         srv->latencies.end());
  out->busy_ms = srv->busy_ms + (srv->busy_slots ? ms_between(srv->busy_since, Clock::now()) : 0.0);
  out->state_bytes_peak = (int64_t)srv->state_peak;
  out->joined = srv->joined;
  return GS_OK;
}

GS_API gs_status gs_server_stats_reset(gs_server* srv) {
  if (!srv) return GS_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(srv->mu);
  srv->submitted = srv->completed = srv->rejected = srv->launches = srv->launched_instances = srv->joined = 0;
  srv->latencies.clear();
  srv->busy_ms = 0;
  if (srv->busy_slots) srv->busy_since = Clock::now();
  srv->state_peak = srv->dev_alloc.used();
  return GS_OK;
}

GS_API void gs_server_free(gs_server* srv) {
  if (!srv) return;
  {
    std::lock_guard<std::mutex> lock(srv->mu);
    srv->stopping = true;
  }
  srv->work_cv.notify_all();
  if (srv->worker.joinable()) srv->worker.join();
  delete srv;
}

}  // extern "C"
