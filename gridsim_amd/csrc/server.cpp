// Serving loop of libgridsim_hip.so: continuous batching of power-flow and contingency requests over resident networks.
//
// The reference answers gridsim-rpc/1 "solve" requests one after another on one thread (RpcCommand.cs:460-497 reads a
// line, :668-700 solves it to completion, writes the response, reads the next). Here a dispatcher thread owns up to
// four launch slots, each with its own stream. Whenever a slot is free it takes the network of the oldest waiting
// request, gathers that network's waiting requests in arrival order (at most max_batch, and only as many as the HBM
// state arena and the pinned result arena can hold), lays them out in the arenas and launches the batched Newton
// kernel on the slot's stream, with the states and voltages copied back into pinned memory behind it. Requests that
// arrive while launches run are gathered into the next free slot, so their work overlaps with what is already on the
// device. Finished slots are reaped by polling their events; a result is published under its ticket and the arena
// blocks return to the free lists.
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

  // One launch slot: its stream, completion event and the launch it holds.
  struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    int net = -1;
    const Net* nt = nullptr;  // srv->nets[net], taken under the lock (the vector may grow meanwhile)
    std::vector<Request> reqs;
    size_t dev_off = 0, dev_bytes = 0, host_off = 0, host_bytes = 0;
    Staging stg;
    PreparedRun pr;
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
  int64_t submitted = 0, completed = 0, rejected = 0, launches = 0, launched_instances = 0;
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

// Byte sizes of one launch of `count` instances of network nt: device block [instances | descriptors | states] and
// pinned block [states | per instance the vm..va span].
size_t device_block(const gs_server::Net& nt, int count) {
  return nt.L.stride * (size_t)count + align_up(sizeof(InstDesc) * (size_t)count) +
         align_up(sizeof(InstState) * (size_t)count);
}
size_t host_block(const gs_server::Net& nt, int count) {
  return align_up(sizeof(InstState) * (size_t)count) + nt.state_span * (size_t)count;
}

// Lays out and launches the requests of slot s (device and pinned blocks already reserved) on the slot's stream.
void launch_slot(gs_server* srv, gs_server::Slot& s) {
  const gs_server::Net& nt = *s.nt;
  const DevPlan& h = nt.net->dev.h;
  const int count = (int)s.reqs.size();
  char* dbase = srv->dev_arena.as<char>() + s.dev_off;
  RunBuffers rb;
  rb.inst = dbase;
  rb.desc = reinterpret_cast<InstDesc*>(dbase + nt.L.stride * (size_t)count);
  rb.state = reinterpret_cast<InstState*>(dbase + nt.L.stride * (size_t)count +
                                          align_up(sizeof(InstDesc) * (size_t)count));
  rb.base_spec = nt.base_spec.as<double>();
  std::vector<gs_instance> inst(count);
  for (int i = 0; i < count; i++) {
    inst[i] = s.reqs[i].inst;
    inst[i].outages = s.reqs[i].outages.empty() ? nullptr : s.reqs[i].outages.data();
    if (!s.reqs[i].start.empty()) {
      inst[i].vm0 = s.reqs[i].start.data();
      inst[i].va0 = s.reqs[i].start.data() + h.n;
    } else {
      inst[i].vm0 = inst[i].va0 = nullptr;
    }
  }
  s.pr = prepare_into(h, rb, s.stg, count, inst.data(), &nt.net->y, &nt.net->net, nt.opts, false, nullptr, nullptr,
                      nullptr, s.stream);
  launch_prepared(h, s.pr, nt.opts.schedule, s.stream);
  char* hbase = srv->pinned + s.host_off;
  check(hipMemcpyAsync(hbase, rb.state, sizeof(InstState) * (size_t)count, hipMemcpyDeviceToHost, s.stream),
        "server states");
  check(hipMemcpy2DAsync(hbase + align_up(sizeof(InstState) * (size_t)count), nt.state_span, dbase + nt.L.vm,
                         nt.L.stride, nt.state_span, (size_t)count, hipMemcpyDeviceToHost, s.stream),
        "server voltages");
  check(hipEventRecord(s.done, s.stream), "server event");
}

// Results of a finished slot (its event has completed): finished instances from the pinned copies, paused ones
// through fresh symbolic passes. Called without the server lock.
std::vector<std::pair<uint64_t, Finished>> reap_slot(gs_server* srv, gs_server::Slot& s, gs_status launch_status) {
  std::vector<std::pair<uint64_t, Finished>> out;
  const int count = (int)s.reqs.size();
  if (launch_status != GS_OK) {
    for (auto& r : s.reqs) {
      Finished f;
      f.status = launch_status;
      out.emplace_back(r.ticket, std::move(f));
    }
    return out;
  }
  const gs_server::Net& nt = *s.nt;
  const int n = nt.net->net.n;
  const char* hbase = srv->pinned + s.host_off;
  const InstState* st = reinterpret_cast<const InstState*>(hbase);
  const char* volts = hbase + align_up(sizeof(InstState) * (size_t)count);
  auto span = [&](int i, size_t off) {
    return reinterpret_cast<const double*>(volts + nt.state_span * (size_t)i + (off - nt.L.vm));
  };
  for (int i = 0; i < count; i++) {
    Finished f;
    f.base_mva = nt.net->net.base;
    const double *vm = span(i, nt.L.vm), *va = span(i, nt.L.va), *p = span(i, nt.L.p), *q = span(i, nt.L.q);
    f.vm.assign(vm, vm + n);
    f.va.assign(va, va + n);
    f.p.assign(p, p + n);
    f.q.assign(q, q + n);
    if (st[i].done == 1) {
      fill_result(st[i], f.result, 0, srv->ctx->us_per_tick);
    } else {
      try {
        double kms = 0;
        int nl = 0;
        gs_instance in = s.reqs[i].inst;
        in.vm0 = in.va0 = nullptr;
        in.outages = s.reqs[i].outages.empty() ? nullptr : s.reqs[i].outages.data();
        std::vector<double> pp, qq;
        // the fresh symbolic passes and resumed launches run on the context's stream: under the context's lock, as
        // every other call on the context, so they never land inside a capture a client has begun on that stream
        std::lock_guard<std::mutex> ctx_lock(srv->ctx->mu);
        finish_paused(nt.net, in, nt.opts, false, st[i], f.vm, f.va, f.result, kms, nl, &pp, &qq);
        if ((int)pp.size() == n) f.p = pp;
        if ((int)qq.size() == n) f.q = qq;
      } catch (const HipError& e) {
        f.status = e.code;
      }
    }
    out.emplace_back(s.reqs[i].ticket, std::move(f));
  }
  return out;
}

// Picks the next launch for a free slot under the lock: the oldest request's network and that network's waiting
// requests in arrival order, as many as max_batch and the two arenas admit. False when nothing can be launched.
bool admit(gs_server* srv, gs_server::Slot& s) {
  if (srv->queue.empty()) return false;
  const int net = srv->queue.front().net;
  const gs_server::Net& nt = *srv->nets[net];
  int avail = 0;
  for (const Request& r : srv->queue)
    if (r.net == net && ++avail >= srv->cfg.max_batch) break;
  for (int count = avail; count >= 1; count = count > 1 ? std::max(1, count / 2) : 0) {
    size_t db = device_block(nt, count), hb = host_block(nt, count);
    size_t doff = srv->dev_alloc.alloc(db);
    if (doff == SIZE_MAX) continue;
    size_t hoff = srv->host_alloc.alloc(hb);
    if (hoff == SIZE_MAX) {
      srv->dev_alloc.release(doff, db);
      continue;
    }
    s.net = net;
    s.nt = &nt;
    s.dev_off = doff; s.dev_bytes = db;
    s.host_off = hoff; s.host_bytes = hb;
    s.reqs.clear();
    for (auto it = srv->queue.begin(); it != srv->queue.end() && (int)s.reqs.size() < count;) {
      if (it->net == net) {
        s.reqs.push_back(std::move(*it));
        it = srv->queue.erase(it);
      } else {
        ++it;
      }
    }
    srv->state_peak = std::max(srv->state_peak, srv->dev_alloc.used());
    return true;
  }
  return false;  // the arenas are full: wait for a slot to finish
}

void dispatcher(gs_server* srv) {
  (void)hipSetDevice(srv->ctx->device);
  std::unique_lock<std::mutex> lock(srv->mu);
  for (;;) {
    // reap finished slots
    for (gs_server::Slot& s : srv->slots) {
      if (!s.busy) continue;
      hipError_t q = hipEventQuery(s.done);
      if (q == hipErrorNotReady) continue;
      lock.unlock();
      auto results = reap_slot(srv, s, q == hipSuccess ? GS_OK : GS_DEVICE_ERROR);
      auto now = Clock::now();
      lock.lock();
      for (auto& tf : results) {
        const uint64_t ticket = tf.first;
        Finished& f = tf.second;
        auto it = std::find_if(s.reqs.begin(), s.reqs.end(), [&](const Request& r) { return r.ticket == ticket; });
        f.latency_ms = ms_between(it->submitted, now);
        srv->latencies.push_back(f.latency_ms);
        srv->in_flight_tickets.erase(ticket);
        srv->finished[ticket] = std::move(f);
        srv->completed++;
      }
      srv->dev_alloc.release(s.dev_off, s.dev_bytes);
      srv->host_alloc.release(s.host_off, s.host_bytes);
      s.reqs.clear();
      s.busy = false;
      if (--srv->busy_slots == 0) srv->busy_ms += ms_between(srv->busy_since, now);
      srv->done_cv.notify_all();
    }
    // fill free slots
    for (gs_server::Slot& s : srv->slots) {
      if (s.busy || !admit(srv, s)) continue;
      for (const Request& r : s.reqs) srv->in_flight_tickets[r.ticket] = (int)(&s - srv->slots.data());
      s.busy = true;
      if (srv->busy_slots++ == 0) srv->busy_since = Clock::now();
      srv->launches++;
      srv->launched_instances += (int64_t)s.reqs.size();
      lock.unlock();
      gs_status st = GS_OK;
      try {
        launch_slot(srv, s);
      } catch (const HipError& e) {
        st = e.code;
      }
      lock.lock();
      if (st != GS_OK) {
        // the launch never reached the device: fail its requests now
        auto now = Clock::now();
        for (Request& r : s.reqs) {
          Finished f;
          f.status = st;
          f.latency_ms = ms_between(r.submitted, now);
          srv->in_flight_tickets.erase(r.ticket);
          srv->finished[r.ticket] = std::move(f);
          srv->completed++;
        }
        srv->dev_alloc.release(s.dev_off, s.dev_bytes);
        srv->host_alloc.release(s.host_off, s.host_bytes);
        s.reqs.clear();
        s.busy = false;
        if (--srv->busy_slots == 0) srv->busy_ms += ms_between(srv->busy_since, now);
        srv->done_cv.notify_all();
      }
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
  return GS_OK;
}

GS_API gs_status gs_server_stats_reset(gs_server* srv) {
  if (!srv) return GS_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(srv->mu);
  srv->submitted = srv->completed = srv->rejected = srv->launches = srv->launched_instances = 0;
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
