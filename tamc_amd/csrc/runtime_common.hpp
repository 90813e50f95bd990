// runtime_common.hpp -- host-side structures of the runtime: error state, instances (parsed and
// built with the reference's rules), device copies and the sequential-order quad schedule.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tamc_gpu.h"

namespace tamc {

void set_error(const std::string &msg);
void clear_error();
tamc_status fail(tamc_status st, const std::string &msg);
tamc_status device_init();

#define TAMC_HIP(expr)                                                                        \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return ::tamc::fail(TAMC_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_));    \
  } while (0)

// Device allocation that is freed on scope exit.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t b) {
    reset();
    bytes = b;
    return hipMalloc(&p, b ? b : 16);
  }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// Bump allocator over one device allocation.
struct Arena {
  size_t used = 0;
  std::vector<std::pair<void **, size_t>> reqs;
  template <class T>
  void want(T **slot, size_t count) {
    reqs.emplace_back((void **)slot, count * sizeof(T));
  }
  DevBuf buf;
  // bytes the registered buffers take, each rounded up to 256
  size_t total() const {
    size_t t = 0;
    for (auto &r : reqs) t += (r.second + 255) & ~(size_t)255;
    return t;
  }
  // point every registered slot into [base, base + total()) (base 256-aligned, owned elsewhere)
  void place(void *base) {
    size_t off = 0;
    for (auto &r : reqs) {
      *r.first = (char *)base + off;
      off += (r.second + 255) & ~(size_t)255;
    }
    used = off;
  }
  hipError_t commit() {
    hipError_t e = buf.alloc(total());
    if (e != hipSuccess) return e;
    place(buf.p);
    return hipSuccess;
  }
};

enum : uint32_t { kOrderColour = 0, kOrderSequential = 1 };

struct SiteSchedule {
  std::vector<uint32_t> group_sites;   // [n_groups][4], kNone pads the last group of a class
  std::vector<uint32_t> class_groups;  // [n_classes+1] offsets into groups
  uint32_t n_classes = 0, n_groups = 0;
};

struct DeviceSchedule {
  DevBuf lvl_off, tasks, groups;
  // per RNG group, everything its sweep task reads, contiguous so that one task loads it with a
  // few wide scalar loads: sites[4], bias bits[4], then for fixed-stride rows (D > 0)
  // cols[4][D] and J bits[4][D]; padded sites have zeros
  DevBuf records;
  uint32_t n_levels = 0, n_tasks = 0, record_words = 0;
};

struct DeviceInstance {
  DevBuf row_ptr, col, J, bias, g_row_ptr, g_col, chi, ell_col, ell_col16, ell_J, self_J;
  // SA row records: per site {first block, blocks, bias bits, row length}, and blocks of eight
  // row entries -- eight neighbour byte offsets in a wave's spin planes (8 * col), then the eight
  // couplings' bits; the last block of a row is padded with the site itself and +0.0
  DevBuf sa_info, sa_blocks;
  DevBuf icm_tiles;           // [icm_ntiles][64] compact tiles of the coupling graph for ICM growth
  DevBuf icm_trows;           // [icm_ntiles * 64][nbr_stride(ell_D)] their 16-bit rows, in tile order
  uint32_t icm_ntiles = 0;
  uint32_t lat_L = 0;         // L when the graph is a periodic square (lat_dim 2) or cubic (3) lattice
  uint32_t lat_dim = 0;       // in row-major order
  bool has_self = false;
  uint32_t ell_D = 0;
  std::map<uint32_t, std::unique_ptr<DeviceSchedule>> sched;  // keyed by 2 * (words per site G) + order
};

}  // namespace tamc

struct tamc_instance {
  uint32_t n = 0;
  int qubo = 0;
  std::vector<uint32_t> row_ptr, col;  // coupling_vecs (src/ising.rs:261) as CSR
  std::vector<float> J, bias;
  float offset = 0.f;
  std::vector<uint32_t> g_row_ptr, g_col;  // coupling CsMat pattern (ICM graph)
  uint32_t g_nodes = 0;
  std::vector<std::vector<double>> chi;
  // Sweep schedules: classes of mutually uncoupled sites, swept class by class.  Within a class,
  // sites in ascending order form groups of four ("RNG groups", numbered class-major); one Philox
  // call keyed by the group serves its four sites.
  //   [kOrderColour]     greedy colouring in site order: colour(i) = the smallest colour not
  //                      taken by a lower-numbered site coupled to i;
  //   [kOrderSequential] levels: level(i) = 1 + max level of the lower-numbered sites coupled to i
  //                      (0 if none), which reproduces the reference's sequential site order.
  tamc::SiteSchedule sched[2];
  std::mutex mu;
  std::unique_ptr<tamc::DeviceInstance> dev;
  // server requests on this instance that are queued or resident: they hold device pointers to
  // its buffers, so the suscept setters refuse to replace them while this is non-zero
  mutable std::atomic<uint32_t> server_uses{0};
};

namespace tamc {
// Upload the instance (once) and the schedule of the given order for G words per site; returns
// nullptr on error.
DeviceInstance *device_instance(const tamc_instance *inst, uint32_t G, tamc_status *st, uint32_t order = kOrderColour);
const DeviceSchedule *device_schedule(const tamc_instance *inst, uint32_t G, tamc_status *st,
                                      uint32_t order = kOrderColour);
// The instance's device copy with its SA row records if an earlier SA call uploaded them, else
// nullptr; never allocates.
const DeviceInstance *resident_instance(const tamc_instance *inst);
// Upload the instance (once) and its SA row records (once); nullptr on error.
const DeviceInstance *device_sa_rows(const tamc_instance *inst, tamc_status *st);
}  // namespace tamc
