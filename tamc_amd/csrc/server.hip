// server.hip -- continuous-batching request loop for PT-ICM and SA solves (tamc_server_* in
// tamc_gpu.h).
//
// The reference's callers run many PtIcmRunner solves at once from rayon (src/gla.rs:152-162,
// src/bin/tamc-pt-opt.rs:68-71); here one loop thread owns the GPU and serves them all:
//
//  * admission: a submitted request is validated in the caller's thread (plan_solve), queued, and
//    admitted by the loop once its state fits in the device pool (reserved once at create, so the
//    loop never calls hipMalloc/hipFree) and its work groups fit in the free wave slots;
//  * residency: an admitted request's spins, ladder and outputs stay in its pool block until it
//    finishes; nothing is copied back in between;
//  * slices: every slice launches all resident groups, bucketed by (waves, LDS class) into one
//    launch per bucket.  The bucket holding the slowest group runs alone on one stream; the
//    others share the remaining streams, one after another on each, so buckets share the CUs.
//    Group g runs sweeps [t_done, t_done + S_g) with S_g sized from its measured per-sweep time
//    so that every stream takes about `slice_ms` (stretched to a whole number of sweeps of the
//    slowest group); requests that arrive meanwhile join at the next slice boundary;
//  * SA requests: the waves of every resident SA solve run sweeps [done, done + S) of their
//    schedules in one sa_kernel launch per slice (on stream 1, after that stream's PT buckets);
//    between slices their bit planes stay in the pool block, and S is sized from a per-entry
//    cost calibrated on the previous launches;
//  * completion: results of requests that finished in slice k are assembled while slice k+1 runs.
//
// Results depend only on (instance, params, seed): the Philox streams are keyed by sweep, site
// group and replica, and energies are summed in fixed point, so slicing and co-residency change
// nothing.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "pt_host.hpp"
#include "sa_host.hpp"

namespace tamc {
namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// First-fit allocator over one device allocation.
struct DevicePool {
  void *base = nullptr;
  size_t bytes = 0, used = 0;
  std::map<size_t, size_t> free_;  // offset -> size, coalesced
  hipError_t init(size_t b) {
    bytes = b;
    const hipError_t e = hipMalloc(&base, b);
    if (e == hipSuccess) free_[0] = b;
    return e;
  }
  bool alloc(size_t sz, size_t *off) {
    sz = (sz + 255) & ~(size_t)255;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < sz) continue;
      const size_t o = it->first, rest = it->second - sz;
      free_.erase(it);
      if (rest) free_[o + sz] = rest;
      *off = o;
      used += sz;
      return true;
    }
    return false;
  }
  void release(size_t off, size_t sz) {
    sz = (sz + 255) & ~(size_t)255;
    used -= sz;
    auto it = free_.emplace(off, sz).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  ~DevicePool() {
    if (base) (void)hipFree(base);
  }
};

enum class State { Queued, Resident, Finished, Done, Failed, Cancelled };

struct Request {
  uint64_t id = 0;
  SolvePlan S;
  Arena A;
  std::vector<uint32_t *> init_dev;  // per group, device copy of the caller's initial states
  size_t pool_off = 0, pool_sz = 0;
  bool in_pool = false, initialized = false;
  bool init_pending = false;  // initialised on the init stream, joins the slices once that is done
  uint32_t waves = 8, lds = 0, lds_class = 0;
  std::vector<uint32_t> t_done;
  std::vector<double> us_per_sweep;
  State state = State::Queued;
  std::atomic<bool> cancel{false};
  tamc_status err = TAMC_OK;
  std::string msg;
  tamc_pt_result *result = nullptr;
  Clock::time_point t_submit, t_admit, t_ready;
  uint32_t slices = 0;
  const tamc_instance *inst = nullptr;  // holds one of inst->server_uses until the request ends
  // SA requests (tamc_server_submit_sa); their PT plan S stays empty
  bool sa = false;
  tamc_sa_params sa_P{};
  std::vector<float> sa_betas;
  uint32_t sa_n = 0, sa_R = 0, sa_waves = 0, sa_done = 0, sa_slice = 0;
  double sa_cost = 0.0;  // row entries a wave walks per sweep
  const DeviceInstance *sa_dev = nullptr;
  float *d_sa_betas = nullptr, *d_sa_energy = nullptr;
  uint32_t *d_sa_spins = nullptr;
  tamc_sa_result *sa_result = nullptr;
  void drop_instance() {
    if (inst) inst->server_uses.fetch_sub(1);
    inst = nullptr;
  }
  ~Request() {
    drop_instance();
    if (sa_result) tamc_sa_result_free(sa_result);
  }
  // waves, LDS bytes and work groups the request holds while resident
  uint32_t hold_waves() const { return sa ? sa_waves : waves * (uint32_t)S.groups.size(); }
  uint64_t hold_lds() const {
    return sa ? (uint64_t)sa_waves * sa_wave_lds_words(sa_n) * 4u : (uint64_t)lds_class * S.groups.size();
  }
  size_t hold_groups() const { return sa ? (sa_waves + 3u) / 4u : S.groups.size(); }
  bool finished_all() const {
    if (sa) return sa_done >= sa_P.num_betas;
    for (uint32_t gi = 0; gi < S.groups.size(); ++gi)
      if (t_done[gi] < S.T) return false;
    return true;
  }
};

// Pinned host + device staging for one launch bucket.
struct Bucket {
  uint32_t waves = 0, lds = 0;
  std::vector<std::pair<Request *, uint32_t>> members;  // (request, group index)
};

constexpr uint32_t kLdsClasses[3] = {40u * 1024u, 80u * 1024u, 160u * 1024u};
constexpr uint32_t kStreams = 3;

}  // namespace
}  // namespace tamc

using namespace tamc;

struct tamc_server {
  tamc_server_options opt{};
  uint32_t cus = 256, max_groups = 1024;
  DevicePool pool;
  // staging: group descriptors, per-group sweep ranges and launch ticks, per stream
  PtGroup *h_groups[kStreams] = {}, *d_groups[kStreams] = {};
  uint32_t *h_ranges[kStreams] = {}, *d_ranges[kStreams] = {};
  unsigned long long *h_ticks[kStreams] = {}, *d_ticks[kStreams] = {};
  PtGroup *h_init_groups = nullptr, *d_init_groups = nullptr;
  // SA tasks of one slice (stream 1), and the events around their launch
  SaTask *h_sa_tasks = nullptr, *d_sa_tasks = nullptr;
  hipEvent_t ev_sa0 = nullptr, ev_sa1 = nullptr;
  std::vector<std::pair<Request *, uint32_t>> sa_batch;  // (request, sweeps) of the running slice
  double sa_us_per_entry = 0.02;  // calibrated launch time per row entry walked by one wave
  const uint32_t **h_init_ptrs = nullptr;
  const uint32_t **d_init_ptrs = nullptr;
  hipStream_t streams[kStreams] = {};
  hipStream_t init_stream = nullptr;  // uploads and initial states of admitted requests
  bool init_outstanding = false;      // ev_init marks init work not yet waited for
  std::vector<Request *> initializing;
  hipStream_t copy_stream = nullptr;  // result downloads, beside the slices
  hipEvent_t ev_init = nullptr, ev_start[kStreams] = {}, ev_end[kStreams] = {};

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request *> queue;
  std::vector<Request *> resident;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> all;      // tickets the caller holds
  std::unordered_map<uint64_t, std::unique_ptr<Request>> zombies;  // cancelled, still referenced by the loop
  uint64_t next_id = 1;
  bool stop = false;
  uint32_t resident_waves = 0;
  uint64_t lds_in_use = 0;
  tamc_server_stats st{};
  std::thread worker;

  ~tamc_server() {
    for (uint32_t k = 0; k < kStreams; ++k) {
      if (h_groups[k]) (void)hipHostFree(h_groups[k]);
      if (h_ranges[k]) (void)hipHostFree(h_ranges[k]);
      if (h_ticks[k]) (void)hipHostFree(h_ticks[k]);
      if (d_groups[k]) (void)hipFree(d_groups[k]);
      if (d_ranges[k]) (void)hipFree(d_ranges[k]);
      if (d_ticks[k]) (void)hipFree(d_ticks[k]);
      if (streams[k]) (void)hipStreamDestroy(streams[k]);
      if (ev_start[k]) (void)hipEventDestroy(ev_start[k]);
      if (ev_end[k]) (void)hipEventDestroy(ev_end[k]);
    }
    if (h_init_groups) (void)hipHostFree(h_init_groups);
    if (h_sa_tasks) (void)hipHostFree(h_sa_tasks);
    if (d_sa_tasks) (void)hipFree(d_sa_tasks);
    if (ev_sa0) (void)hipEventDestroy(ev_sa0);
    if (ev_sa1) (void)hipEventDestroy(ev_sa1);
    if (h_init_ptrs) (void)hipHostFree((void *)h_init_ptrs);
    if (d_init_groups) (void)hipFree(d_init_groups);
    if (d_init_ptrs) (void)hipFree((void *)d_init_ptrs);
    if (ev_init) (void)hipEventDestroy(ev_init);
    if (init_stream) (void)hipStreamDestroy(init_stream);
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
  }

  tamc_status setup() {
    hipDeviceProp_t prop;
    TAMC_HIP(hipGetDeviceProperties(&prop, 0));
    cus = (uint32_t)std::max(1, prop.multiProcessorCount);
    if (!opt.pool_bytes) opt.pool_bytes = 4ull << 30;
    if (!opt.max_wave_slots) opt.max_wave_slots = 16u * cus;
    if (!(opt.slice_ms > 0.0)) opt.slice_ms = 4.0;
    if (opt.waves_per_group > kMaxWaves) return fail(TAMC_E_INVALID_ARG, "waves_per_group must be 0..16");
    max_groups = std::max(1u, opt.max_wave_slots);  // at least one wave per group
    TAMC_HIP(pool.init(opt.pool_bytes));
    // stream 0 carries the slice's slowest groups and may use every CU; the other streams leave
    // the first eighth of the CUs free, so that those groups never wait for a whole CU to drain of
    // smaller work groups
    std::vector<uint32_t> mask((cus + 31u) / 32u, 0u);
    for (uint32_t cu = cus / 8u; cu < cus; ++cu) mask[cu / 32u] |= 1u << (cu % 32u);
    for (uint32_t k = 0; k < kStreams; ++k) {
      if (k == 0u || cus < 16u)
        TAMC_HIP(hipStreamCreateWithFlags(&streams[k], hipStreamNonBlocking));
      else
        TAMC_HIP(hipExtStreamCreateWithCUMask(&streams[k], (uint32_t)mask.size(), mask.data()));
      TAMC_HIP(hipEventCreate(&ev_start[k]));
      TAMC_HIP(hipEventCreate(&ev_end[k]));
      TAMC_HIP(hipHostMalloc((void **)&h_groups[k], max_groups * sizeof(PtGroup), hipHostMallocDefault));
      TAMC_HIP(hipHostMalloc((void **)&h_ranges[k], max_groups * 2 * sizeof(uint32_t), hipHostMallocDefault));
      TAMC_HIP(hipHostMalloc((void **)&h_ticks[k], max_groups * sizeof(unsigned long long), hipHostMallocDefault));
      TAMC_HIP(hipMalloc((void **)&d_groups[k], max_groups * sizeof(PtGroup)));
      TAMC_HIP(hipMalloc((void **)&d_ranges[k], max_groups * 2 * sizeof(uint32_t)));
      TAMC_HIP(hipMalloc((void **)&d_ticks[k], max_groups * sizeof(unsigned long long)));
    }
    TAMC_HIP(hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking));
    TAMC_HIP(hipEventCreate(&ev_init));
    TAMC_HIP(hipStreamCreateWithFlags(&init_stream, hipStreamNonBlocking));
    TAMC_HIP(hipHostMalloc((void **)&h_init_groups, max_groups * sizeof(PtGroup), hipHostMallocDefault));
    TAMC_HIP(hipHostMalloc((void **)&h_init_ptrs, max_groups * sizeof(void *), hipHostMallocDefault));
    TAMC_HIP(hipMalloc((void **)&d_init_groups, max_groups * sizeof(PtGroup)));
    TAMC_HIP(hipMalloc((void **)&d_init_ptrs, max_groups * sizeof(void *)));
    TAMC_HIP(hipHostMalloc((void **)&h_sa_tasks, opt.max_wave_slots * sizeof(SaTask), hipHostMallocDefault));
    TAMC_HIP(hipMalloc((void **)&d_sa_tasks, opt.max_wave_slots * sizeof(SaTask)));
    TAMC_HIP(hipEventCreate(&ev_sa0));
    TAMC_HIP(hipEventCreate(&ev_sa1));
    return TAMC_OK;
  }

  static uint32_t lds_class_of(uint32_t lds) {
    for (uint32_t c : kLdsClasses)
      if (lds <= c) return c;
    return kLdsClasses[2];
  }

  // Initial per-sweep estimate (us) before a group has been measured: a sweep costs about one
  // task latency per (4 sites x 64 replicas) / waves, plus the ICM passes over the pairs.
  static double model_us_per_sweep(const Request &r, uint32_t gi) {
    const GroupPlan &g = r.S.groups[gi];
    const double tasks = (double)r.S.inst->sched[r.S.P.sweep_order].n_groups * g.G;
    return 8.0 + 5.0 * tasks / r.waves * (g.P ? 1.8 : 1.0);
  }

  // Move queued requests to the device while pool memory and wave slots allow (FIFO: a request
  // that does not fit blocks the ones behind it, so large requests are not starved).
  void admit(std::vector<Request *> &fresh) {
    while (!queue.empty()) {
      Request *r = queue.front();
      if (r->cancel) {
        queue.pop_front();
        finish_cancel(r);
        continue;
      }
      const uint32_t need_waves = r->hold_waves();
      const uint64_t need_lds = r->hold_lds();
      const bool room = resident.empty() || (resident_waves + need_waves <= opt.max_wave_slots &&
                                             lds_in_use + need_lds <= (uint64_t)cus * kLdsClasses[2] &&
                                             resident_group_count() + r->hold_groups() <= max_groups);
      if (!room) break;
      if (!r->in_pool) {
        if (r->A.reqs.empty() && r->sa) {
          r->A.want(&r->d_sa_betas, r->sa_P.num_betas);
          r->A.want(&r->d_sa_spins, (size_t)r->sa_waves * r->sa_n * 2);
          r->A.want(&r->d_sa_energy, (size_t)r->sa_waves * 64);
        } else if (r->A.reqs.empty()) {
          want_buffers(r->S, r->A);
          r->init_dev.assign(r->S.groups.size(), nullptr);
          for (size_t gi = 0; gi < r->S.groups.size(); ++gi)
            if (!r->S.groups[gi].init_bits.empty()) r->A.want(&r->init_dev[gi], r->S.groups[gi].init_bits.size());
        }
        const size_t sz = r->A.total();
        if (sz > pool.bytes) {
          queue.pop_front();
          fail_request(r, TAMC_E_NOMEM, "request state exceeds the server's device pool");
          continue;
        }
        size_t off = 0;
        if (!pool.alloc(sz, &off)) break;  // wait for memory to free up
        r->pool_off = off;
        r->pool_sz = sz;
        r->in_pool = true;
        r->A.place((char *)pool.base + off);
      }
      queue.pop_front();
      r->state = State::Resident;
      r->t_admit = Clock::now();
      r->t_done.assign(r->S.groups.size(), 0u);
      r->us_per_sweep.resize(r->S.groups.size());
      for (uint32_t gi = 0; gi < r->S.groups.size(); ++gi) r->us_per_sweep[gi] = model_us_per_sweep(*r, gi);
      resident.push_back(r);
      resident_waves += need_waves;
      lds_in_use += need_lds;
      fresh.push_back(r);
    }
  }

  size_t resident_group_count() const {
    size_t k = 0;
    for (auto *r : resident) k += r->hold_groups();
    return k;
  }

  void release_resources(Request *r) {
    r->drop_instance();
    if (r->in_pool) {
      pool.release(r->pool_off, r->pool_sz);
      r->in_pool = false;
    }
  }

  void leave_resident(Request *r) {
    resident.erase(std::remove(resident.begin(), resident.end(), r), resident.end());
    resident_waves -= r->hold_waves();
    lds_in_use -= r->hold_lds();
  }

  void fail_request(Request *r, tamc_status e, const std::string &m) {
    release_resources(r);
    r->state = State::Failed;
    r->err = e;
    r->msg = m;
    r->t_ready = Clock::now();
    st.failed++;
    cv_done.notify_all();
  }

  void finish_cancel(Request *r) {
    release_resources(r);
    r->state = State::Cancelled;
    st.cancelled++;
    cv_done.notify_all();
    zombies.erase(r->id);  // deletes r unless the server is being destroyed (then `all` owns it)
  }

  // Device work for the requests admitted since the last slice: initial states and energies.
  // Uploads and initial states of the requests admitted at this boundary run on the init stream,
  // beside the slice that follows; those requests join the slices after it.  The host waits for
  // the previous boundary's init work (normally long done) before reusing its staging buffers
  // and releasing its requests into the slices.
  // waits for the outstanding init work; its requests may then join the slices (or be released)
  tamc_status drain_init() {
    if (!init_outstanding) return TAMC_OK;
    init_outstanding = false;
    for (Request *r : initializing) r->init_pending = false;
    initializing.clear();
    TAMC_HIP(hipEventSynchronize(ev_init));
    return TAMC_OK;
  }

  tamc_status init_fresh(std::vector<Request *> &fresh) {
    if (tamc_status s = drain_init()) return s;
    uint32_t ng = 0;
    bool any_init = false, any_sa = false;
    for (Request *r : fresh) {
      if (r->sa) {  // the schedule's betas; the kernel makes the initial states itself
        TAMC_HIP(hipMemcpyAsync(r->d_sa_betas, r->sa_betas.data(), r->sa_betas.size() * 4, hipMemcpyHostToDevice,
                                init_stream));
        r->initialized = true;
        r->init_pending = true;
        initializing.push_back(r);
        any_sa = true;
        continue;
      }
      tamc_status s = fill_groups(r->S, init_stream);
      if (s) return s;
      for (size_t gi = 0; gi < r->S.groups.size(); ++gi) {
        GroupPlan &g = r->S.groups[gi];
        if (!g.init_bits.empty()) {
          TAMC_HIP(hipMemcpyAsync(r->init_dev[gi], g.init_bits.data(), g.init_bits.size() * 4,
                                  hipMemcpyHostToDevice, init_stream));
          any_init = true;
        }
        h_init_groups[ng] = g.g;
        h_init_ptrs[ng] = r->init_dev[gi];
        ++ng;
      }
      r->initialized = true;
      r->init_pending = true;
      initializing.push_back(r);
    }
    if (!ng && !any_sa) return TAMC_OK;
    if (ng) {
      TAMC_HIP(hipMemcpyAsync(d_init_groups, h_init_groups, ng * sizeof(PtGroup), hipMemcpyHostToDevice, init_stream));
      if (any_init)
        TAMC_HIP(hipMemcpyAsync((void *)d_init_ptrs, (const void *)h_init_ptrs, ng * sizeof(void *),
                                hipMemcpyHostToDevice, init_stream));
      TAMC_HIP(tamc_launch_pt_init(d_init_groups, ng, any_init ? d_init_ptrs : nullptr, init_stream));
    }
    TAMC_HIP(hipEventRecord(ev_init, init_stream));
    init_outstanding = true;
    return TAMC_OK;
  }

  // One slice over the resident groups; returns the slice's device time in ms.
  tamc_status run_slice(double *slice_ms) {
    // buckets by (waves, LDS class); 16-wave groups fill a CU whatever their LDS, so they share
    // one bucket
    std::map<std::pair<uint32_t, uint32_t>, Bucket> buckets;
    for (Request *r : resident)
      for (uint32_t gi = 0; gi < r->S.groups.size(); ++gi) {
        if (r->init_pending || r->t_done[gi] >= r->S.T) continue;
        Bucket &b = buckets[{r->waves, r->waves >= 16u ? 0u : r->lds_class}];
        b.waves = r->waves;
        b.lds = std::max(b.lds, r->lds);
        b.members.push_back({r, gi});
      }
    // The slice lasts about slice_ms, stretched to a whole number of sweeps (rounded to nearest)
    // of its slowest group so that group never idles at the slice's end.  That group's bucket
    // runs alone on stream 0; the other buckets are spread over the other streams, where they run
    // one after another and share the slice's length equally.
    Bucket *crit = nullptr;
    double target_us = opt.slice_ms * 1000.0, slowest = 0.0;
    for (auto &kv : buckets)
      for (auto &mb : kv.second.members) {
        const double us = std::max(1.0, mb.first->us_per_sweep[mb.second]);
        if (us > slowest) {
          slowest = us;
          crit = &kv.second;
        }
      }
    if (crit) target_us = std::max(target_us, slowest * std::max(1.0, std::round(opt.slice_ms * 1000.0 / slowest)));
    std::vector<std::vector<Bucket *>> per_stream(kStreams);
    uint32_t k = 0;
    for (auto &kv : buckets) {
      if (&kv.second == crit) {
        per_stream[0].push_back(crit);
      } else {
        per_stream[1u + k % (kStreams - 1u)].push_back(&kv.second);
        ++k;
      }
    }
    // SA solves: every resident one runs a whole number of sweeps on stream 1, after that stream's
    // buckets, within the same share of the slice as each of them
    sa_batch.clear();
    uint32_t sa_tasks = 0, sa_maxn = 0;
    {
      const double budget_us = target_us / (double)(per_stream[1].size() + 1u);
      for (Request *r : resident) {
        if (!r->sa || r->init_pending || r->sa_done >= r->sa_P.num_betas) continue;
        const double per = std::max(1e-3, sa_us_per_entry * r->sa_cost);
        const double want = std::floor(budget_us / per + 1e-9);
        const uint32_t sweeps = (uint32_t)std::min<double>(r->sa_P.num_betas - r->sa_done, std::max(1.0, want));
        for (uint32_t w = 0; w < r->sa_waves; ++w) {
          SaTask &t = h_sa_tasks[sa_tasks++];
          t = SaTask{};
          t.info = (const uint32_t *)r->sa_dev->sa_info.p;
          t.blocks = (const uint32_t *)r->sa_dev->sa_blocks.p;
          t.betas = r->d_sa_betas;
          t.offset = r->inst->offset;
          t.n = r->sa_n;
          t.K = r->sa_P.num_betas;
          t.R = r->sa_R;
          t.rbase = w * 64u;
          t.k0 = r->sa_done;
          t.k1 = r->sa_done + sweeps;
          t.seed = r->sa_P.seed;
          t.spins_out = r->d_sa_spins + (size_t)w * r->sa_n * 2;
          t.energy_out = r->d_sa_energy + (size_t)w * 64;
        }
        sa_maxn = std::max(sa_maxn, r->sa_n);
        sa_batch.push_back({r, sweeps});
      }
    }
    *slice_ms = 0.0;
    bool launched[kStreams] = {false, false, false};
    std::vector<std::vector<std::pair<Bucket *, uint32_t>>> offsets(kStreams);
    for (uint32_t s = 0; s < kStreams; ++s) {
      uint32_t base = 0;
      const double budget_us = per_stream[s].empty() ? 0.0 : target_us / (double)per_stream[s].size();
      for (Bucket *b : per_stream[s]) {
        offsets[s].push_back({b, base});
        for (size_t m = 0; m < b->members.size(); ++m) {
          Request *r = b->members[m].first;
          const uint32_t gi = b->members[m].second;
          const uint32_t left = r->S.T - r->t_done[gi];
          const double want = std::floor(budget_us / std::max(1.0, r->us_per_sweep[gi]) + 1e-9);
          const uint32_t sweeps = (uint32_t)std::min<double>(left, std::max(1.0, want));
          h_groups[s][base + m] = r->S.groups[gi].g;
          h_ranges[s][2 * (base + m)] = r->t_done[gi];
          h_ranges[s][2 * (base + m) + 1] = r->t_done[gi] + sweeps;
        }
        base += (uint32_t)b->members.size();
      }
      const bool sa_here = s == 1u && sa_tasks > 0u;
      if (!base && !sa_here) continue;
      if (base) {
        TAMC_HIP(hipMemcpyAsync(d_groups[s], h_groups[s], base * sizeof(PtGroup), hipMemcpyHostToDevice, streams[s]));
        TAMC_HIP(hipMemcpyAsync(d_ranges[s], h_ranges[s], base * 2 * sizeof(uint32_t), hipMemcpyHostToDevice,
                                streams[s]));
      }
      if (sa_here)
        TAMC_HIP(hipMemcpyAsync(d_sa_tasks, h_sa_tasks, sa_tasks * sizeof(SaTask), hipMemcpyHostToDevice, streams[s]));
      TAMC_HIP(hipEventRecord(ev_start[s], streams[s]));
      for (auto &bo : offsets[s]) {
        Bucket *b = bo.first;
        TAMC_HIP(tamc_launch_pt_sweeps(d_groups[s] + bo.second, (uint32_t)b->members.size(), 0, 0, b->waves, b->lds, 1u,
                                       d_ranges[s] + 2 * bo.second, d_ticks[s] + bo.second, streams[s], nullptr, 0u));
      }
      if (sa_here) {
        TAMC_HIP(hipEventRecord(ev_sa0, streams[s]));
        if (tamc_status e = sa_launch(d_sa_tasks, sa_tasks, sa_maxn, streams[s])) return e;
        TAMC_HIP(hipEventRecord(ev_sa1, streams[s]));
      }
      TAMC_HIP(hipEventRecord(ev_end[s], streams[s]));
      if (base)
        TAMC_HIP(hipMemcpyAsync(h_ticks[s], d_ticks[s], base * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                streams[s]));
      launched[s] = true;
    }
    slice_launched = launched[0] || launched[1] || launched[2];
    return collect_after_launch(per_stream, offsets, launched, slice_ms);
  }
  bool slice_launched = false;

  std::vector<Request *> finished_prev;

  tamc_status collect_after_launch(std::vector<std::vector<Bucket *>> &per_stream,
                                   std::vector<std::vector<std::pair<Bucket *, uint32_t>>> &offsets,
                                   const bool *launched, double *slice_ms) {
    (void)per_stream;
    // while the slice runs: assemble the requests that finished in the previous slice
    for (Request *r : finished_prev) complete(r);
    finished_prev.clear();
    for (uint32_t s = 0; s < kStreams; ++s)
      if (launched[s]) TAMC_HIP(hipStreamSynchronize(streams[s]));
    for (uint32_t s = 0; s < kStreams; ++s) {
      if (!launched[s]) continue;
      float ms = 0.f;
      // the slice's device time: from the earliest stream start to the latest stream end (the
      // streams overlap, so this is the largest start-to-end distance over all stream pairs)
      for (uint32_t a = 0; a < kStreams; ++a) {
        if (!launched[a]) continue;
        TAMC_HIP(hipEventElapsedTime(&ms, ev_start[a], ev_end[s]));
        *slice_ms = std::max(*slice_ms, (double)ms);
      }
      for (auto &bo : offsets[s]) {
        Bucket *b = bo.first;
        for (size_t m = 0; m < b->members.size(); ++m) {
          Request *r = b->members[m].first;
          const uint32_t gi = b->members[m].second;
          const uint32_t a = h_ranges[s][2 * (bo.second + m)], e = h_ranges[s][2 * (bo.second + m) + 1];
          const uint32_t done = std::min(e, r->S.T) - a;
          const double us = (double)h_ticks[s][bo.second + m] * 0.01;  // 100 MHz ticks
          if (done) {
            const double per = us / done;
            r->us_per_sweep[gi] = r->slices ? 0.5 * (r->us_per_sweep[gi] + per) : per;
          }
          r->t_done[gi] = std::min(e, r->S.T);
        }
      }
    }
    if (!sa_batch.empty()) {
      // the launch lasts as long as its longest wave: recalibrate the time per row entry from it
      float ms = 0.f;
      TAMC_HIP(hipEventElapsedTime(&ms, ev_sa0, ev_sa1));
      double longest = 0.0;
      for (auto &rs : sa_batch) longest = std::max(longest, rs.first->sa_cost * rs.second);
      if (longest > 0.0 && ms > 0.f) sa_us_per_entry = 0.5 * (sa_us_per_entry + (double)ms * 1000.0 / longest);
      for (auto &rs : sa_batch) rs.first->sa_done += rs.second;
    }
    return TAMC_OK;
  }

  void complete_sa(Request *r) {
    const Clock::time_point t0 = Clock::now();
    const size_t np = (size_t)r->sa_waves * r->sa_n * 2, ne = (size_t)r->sa_waves * 64;
    std::vector<uint32_t> planes(np);
    std::vector<float> en(ne);
    tamc_sa_result *res = nullptr;
    if (hipMemcpyAsync(planes.data(), r->d_sa_spins, np * 4, hipMemcpyDeviceToHost, copy_stream) == hipSuccess &&
        hipMemcpyAsync(en.data(), r->d_sa_energy, ne * 4, hipMemcpyDeviceToHost, copy_stream) == hipSuccess &&
        hipStreamSynchronize(copy_stream) == hipSuccess)
      res = sa_assemble(r->sa_n, r->sa_R, planes.data(), en.data());
    std::lock_guard<std::mutex> lk(mu);
    st.complete_ms += ms_between(t0, Clock::now());
    if (!res) {
      fail_request(r, TAMC_E_DEVICE, "result download failed");
      return;
    }
    release_resources(r);
    res->attempts = (uint64_t)r->sa_P.num_betas * r->sa_n * r->sa_R;
    res->timing_us = ms_between(r->t_admit, Clock::now()) * 1000.0;
    r->sa_result = res;
    r->state = State::Done;
    r->t_ready = Clock::now();
    st.completed++;
    st.attempts += res->attempts;
    if (r->cancel) {
      tamc_sa_result_free(res);
      r->sa_result = nullptr;
      finish_cancel(r);
      return;
    }
    cv_done.notify_all();
  }

  void complete(Request *r) {
    if (r->sa) {
      complete_sa(r);
      return;
    }
    const Clock::time_point t0 = Clock::now();
    tamc_pt_result *res = nullptr;
    // one bulk download of the request's pool block, then assembly from the host copy
    std::vector<char> host(r->pool_sz);
    const char *dev = (const char *)pool.base + r->pool_off;
    tamc_status s = TAMC_OK;
    // on a non-blocking stream of its own: the slice streams may be blocking (CU-masked) ones
    if (hipMemcpyAsync(host.data(), dev, r->pool_sz, hipMemcpyDeviceToHost, copy_stream) != hipSuccess ||
        hipStreamSynchronize(copy_stream) != hipSuccess) {
      s = TAMC_E_DEVICE;
    } else {
      const HostMirror m{host.data(), dev, r->pool_sz};
      s = assemble(r->S, &res, &m);
    }
    std::lock_guard<std::mutex> lk(mu);
    st.complete_ms += ms_between(t0, Clock::now());
    if (s) {
      fail_request(r, s, "result download failed");
      return;
    }
    release_resources(r);
    res->timing_us = ms_between(r->t_admit, Clock::now()) * 1000.0;
    r->result = res;
    r->state = State::Done;
    r->t_ready = Clock::now();
    st.completed++;
    st.attempts += res->attempts;
    if (r->cancel) {
      tamc_pt_result_free(res);
      r->result = nullptr;
      finish_cancel(r);
      return;
    }
    cv_done.notify_all();
  }

  void loop() {
    for (;;) {
      std::vector<Request *> fresh;
      {
        std::unique_lock<std::mutex> lk(mu);
        // cancelled requests leave the queue, and the device at this boundary (after any init
        // work still writing their state)
        for (Request *r : resident)
          if (r->cancel && r->init_pending) (void)drain_init();
        for (Request *r : std::vector<Request *>(resident))
          if (r->cancel) {
            leave_resident(r);
            finish_cancel(r);
          }
        for (auto it = queue.begin(); it != queue.end();) {
          if ((*it)->cancel) {
            Request *r = *it;
            it = queue.erase(it);
            finish_cancel(r);
          } else {
            ++it;
          }
        }
        admit(fresh);
        if (resident.empty() && finished_prev.empty()) {
          if (stop) break;
          cv_work.wait(lk, [&] { return stop || !queue.empty(); });
          continue;
        }
      }
      tamc_status s = init_fresh(fresh);
      double slice_ms = 0.0;
      const Clock::time_point t_slice = Clock::now();
      if (!s && !resident.empty()) s = run_slice(&slice_ms);
      const double wall_ms = ms_between(t_slice, Clock::now());
      if (s) {
        // a device error ends every resident request with it
        (void)drain_init();
        std::lock_guard<std::mutex> lk(mu);
        const std::string m = tamc_last_error();
        for (Request *r : std::vector<Request *>(resident)) {
          leave_resident(r);
          fail_request(r, s, m);
        }
        for (Request *r : finished_prev) fail_request(r, s, m);
        finished_prev.clear();
        continue;
      }
      if (resident.empty()) {
        for (Request *r : finished_prev) complete(r);
        finished_prev.clear();
        continue;
      }
      std::lock_guard<std::mutex> lk(mu);
      if (!slice_launched) continue;  // every resident request was still being initialised
      st.slices++;
      st.device_busy_ms += slice_ms;
      st.slice_wall_ms += wall_ms;
      for (Request *r : std::vector<Request *>(resident)) {
        if (r->init_pending) continue;
        r->slices++;
        if (r->finished_all()) {
          leave_resident(r);
          r->state = State::Finished;
          finished_prev.push_back(r);
        }
      }
    }
  }
};

extern "C" {

tamc_status tamc_server_create(const tamc_server_options *opts, tamc_server **out) {
  clear_error();
  if (!out) return fail(TAMC_E_INVALID_ARG, "null argument");
  *out = nullptr;
  tamc_status s = device_init();
  if (s) return s;
  std::unique_ptr<tamc_server> srv(new tamc_server());
  if (opts) srv->opt = *opts;
  if ((s = srv->setup())) return s;
  srv->st.pool_bytes = srv->pool.bytes;
  tamc_server *p = srv.release();
  p->worker = std::thread([p]() { p->loop(); });
  *out = p;
  return TAMC_OK;
}

tamc_status tamc_server_submit(tamc_server *srv, const tamc_instance *inst, const tamc_pt_params *params,
                               const int8_t *init_states, uint64_t *ticket) {
  clear_error();
  if (!srv || !ticket) return fail(TAMC_E_INVALID_ARG, "null argument");
  std::unique_ptr<Request> r(new Request());
  r->t_submit = Clock::now();
  // taken before planning, so that the suscept vectors the plan reads cannot be replaced under it
  if (inst) {
    inst->server_uses.fetch_add(1);
    r->inst = inst;
  }
  r->waves = srv->opt.waves_per_group ? srv->opt.waves_per_group : (inst ? default_waves(inst) : 4u);
  tamc_status s = plan_solve(r->S, inst, params, init_states, r->waves);
  if (s) return s;
  for (auto &g : r->S.groups)
    if (g.spins_global)
      return fail(TAMC_E_UNSUPPORTED, "the server runs chain pairs whose spin planes fit in LDS "
                                      "(about 19,000 sites at 64 replicas); use tamc_pt_run for larger instances");
  if (r->S.groups.size() > srv->max_groups)
    return fail(TAMC_E_UNSUPPORTED, "request has more chain pairs than the server's work-group slots");
  r->S.init_states = nullptr;  // copied into the groups' init_bits
  // the instance is uploaded (once per instance) here, in the caller's thread, not in the loop
  for (auto &g : r->S.groups) {
    if (!device_instance(inst, g.G, &s, r->S.P.sweep_order)) return s;
    if (!device_schedule(inst, g.G, &s, r->S.P.sweep_order)) return s;
  }
  for (auto &g : r->S.groups) r->lds = std::max(r->lds, lds_bytes_for(r->S, g, r->waves));
  r->lds_class = tamc_server::lds_class_of(r->lds);
  std::lock_guard<std::mutex> lk(srv->mu);
  if (srv->stop) return fail(TAMC_E_INVALID_ARG, "server is shutting down");
  r->id = srv->next_id++;
  *ticket = r->id;
  srv->queue.push_back(r.get());
  srv->all[r->id] = std::move(r);
  srv->st.submitted++;
  srv->cv_work.notify_one();
  return TAMC_OK;
}

tamc_status tamc_server_submit_sa(tamc_server *srv, const tamc_instance *inst, const tamc_sa_params *params,
                                  uint64_t *ticket) {
  clear_error();
  if (!srv || !ticket || !params) return fail(TAMC_E_INVALID_ARG, "null argument");
  if (tamc_status s = sa_validate(inst, *params)) return s;
  if (!sa_wave_lds_words(inst->n))
    return fail(TAMC_E_UNSUPPORTED, "the server runs SA instances whose planes fit in LDS (n <= 20224); use "
                                    "tamc_sa_run for larger instances");
  std::unique_ptr<Request> r(new Request());
  r->t_submit = Clock::now();
  inst->server_uses.fetch_add(1);
  r->inst = inst;
  r->sa = true;
  r->sa_P = *params;
  r->sa_betas.assign(params->betas, params->betas + params->num_betas);
  r->sa_P.betas = r->sa_betas.data();
  r->sa_n = inst->n;
  r->sa_R = sa_kept_replicas(*params);
  r->sa_waves = (r->sa_R + 63u) / 64u;
  r->sa_cost = (double)inst->col.size() + inst->n;  // row entries plus per-site work of one sweep
  if (r->sa_waves > srv->opt.max_wave_slots)
    return fail(TAMC_E_UNSUPPORTED, "request has more restart waves than the server's wave slots");
  tamc_status s = TAMC_OK;
  // the instance and its SA rows are uploaded (once per instance) here, in the caller's thread
  if (!(r->sa_dev = device_sa_rows(inst, &s))) return s;
  std::lock_guard<std::mutex> lk(srv->mu);
  if (srv->stop) return fail(TAMC_E_INVALID_ARG, "server is shutting down");
  r->id = srv->next_id++;
  *ticket = r->id;
  srv->queue.push_back(r.get());
  srv->all[r->id] = std::move(r);
  srv->st.submitted++;
  srv->cv_work.notify_one();
  return TAMC_OK;
}

// the common part of tamc_server_poll and tamc_server_poll_sa: waits for the request (of kind sa)
// and hands back its record, or returns TAMC_E_PENDING / the request's error
static tamc_status poll_request(tamc_server *srv, uint64_t ticket, int32_t wait_ms, bool sa,
                                tamc_request_times *times, std::unique_ptr<Request> *taken) {
  std::unique_lock<std::mutex> lk(srv->mu);
  auto it = srv->all.find(ticket);
  if (it == srv->all.end()) return fail(TAMC_E_INVALID_ARG, "unknown or already collected ticket");
  Request *r = it->second.get();
  if (r->sa != sa)
    return fail(TAMC_E_INVALID_ARG, sa ? "ticket is a PT-ICM request: poll it with tamc_server_poll"
                                       : "ticket is an SA request: poll it with tamc_server_poll_sa");
  auto ready = [&] { return r->state == State::Done || r->state == State::Failed; };
  if (!ready() && wait_ms != 0) {
    if (wait_ms < 0) srv->cv_done.wait(lk, ready);
    else srv->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), ready);
  }
  if (!ready()) return TAMC_E_PENDING;
  if (times) {
    times->queued_ms = ms_between(r->t_submit, r->t_admit);
    times->total_ms = ms_between(r->t_submit, r->t_ready);
    times->slices = r->slices;
  }
  *taken = std::move(it->second);
  srv->all.erase(it);
  if ((*taken)->err) return fail((*taken)->err, (*taken)->msg);
  return TAMC_OK;
}

tamc_status tamc_server_poll(tamc_server *srv, uint64_t ticket, int32_t wait_ms, tamc_pt_result **out,
                             tamc_request_times *times) {
  clear_error();
  if (!srv || !out) return fail(TAMC_E_INVALID_ARG, "null argument");
  *out = nullptr;
  std::unique_ptr<Request> r;
  const tamc_status e = poll_request(srv, ticket, wait_ms, false, times, &r);
  if (r && !e) {
    *out = r->result;
    r->result = nullptr;
  }
  return e;
}

tamc_status tamc_server_poll_sa(tamc_server *srv, uint64_t ticket, int32_t wait_ms, tamc_sa_result **out,
                                tamc_request_times *times) {
  clear_error();
  if (!srv || !out) return fail(TAMC_E_INVALID_ARG, "null argument");
  *out = nullptr;
  std::unique_ptr<Request> r;
  const tamc_status e = poll_request(srv, ticket, wait_ms, true, times, &r);
  if (r && !e) {
    *out = r->sa_result;
    r->sa_result = nullptr;
  }
  return e;
}

tamc_status tamc_server_cancel(tamc_server *srv, uint64_t ticket) {
  clear_error();
  if (!srv) return fail(TAMC_E_INVALID_ARG, "null argument");
  std::lock_guard<std::mutex> lk(srv->mu);
  auto it = srv->all.find(ticket);
  if (it == srv->all.end()) return fail(TAMC_E_INVALID_ARG, "unknown or already collected ticket");
  Request *r = it->second.get();
  if (r->state == State::Done || r->state == State::Failed) {
    if (r->result) tamc_pt_result_free(r->result);
    srv->all.erase(it);
    return TAMC_OK;
  }
  // the ticket is released now; the loop drops the request at the next slice boundary
  r->cancel = true;
  srv->zombies[ticket] = std::move(it->second);
  srv->all.erase(it);
  srv->cv_work.notify_one();
  return TAMC_OK;
}

tamc_status tamc_server_get_stats(tamc_server *srv, tamc_server_stats *stats) {
  clear_error();
  if (!srv || !stats) return fail(TAMC_E_INVALID_ARG, "null argument");
  std::lock_guard<std::mutex> lk(srv->mu);
  *stats = srv->st;
  stats->queued = (uint32_t)srv->queue.size();
  stats->resident = (uint32_t)srv->resident.size();
  stats->resident_waves = srv->resident_waves;
  stats->pool_used = srv->pool.used;
  stats->pool_bytes = srv->pool.bytes;
  return TAMC_OK;
}

void tamc_server_destroy(tamc_server *srv) {
  if (!srv) return;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;
    for (auto &kv : srv->all) kv.second->cancel = true;
    srv->cv_work.notify_all();
  }
  if (srv->worker.joinable()) srv->worker.join();
  for (auto &kv : srv->all)
    if (kv.second->result) tamc_pt_result_free(kv.second->result);
  delete srv;
}

}  // extern "C"
