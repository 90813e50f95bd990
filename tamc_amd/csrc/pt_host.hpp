// pt_host.hpp -- host-side planning structures of the production PT-ICM path, shared by the
// one-shot entry points (pt_runtime.hip) and the serving loop (server.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "pt_types.hpp"
#include "runtime_common.hpp"

extern "C" hipError_t tamc_launch_pt_init(const tamc::PtGroup *groups, uint32_t ngroups,
                                          const uint32_t *const *init_bits, hipStream_t s);
extern "C" hipError_t tamc_launch_pt_sweeps(const tamc::PtGroup *groups, uint32_t ngroups, uint32_t t0, uint32_t t1,
                                            uint32_t waves, uint32_t lds_bytes, uint32_t icm_lds, const uint32_t *ranges,
                                            unsigned long long *launch_ticks, hipStream_t s, const uint32_t *jobs,
                                            uint32_t spins_global);
extern "C" hipError_t tamc_pt_sweep_capacity(uint32_t waves, uint32_t lds_bytes, uint32_t *capacity);

namespace tamc {

constexpr uint32_t kLdsLimit = 160u * 1024u;

template <class T>
inline T *host_dup(const std::vector<T> &v) {
  T *p = (T *)malloc(std::max<size_t>(1, v.size() * sizeof(T)));
  if (!v.empty()) memcpy(p, v.data(), v.size() * sizeof(T));
  return p;
}

struct GroupPlan {
  uint32_t chain0, nch, R, G, P;
  bool icm_global = false;
  bool spins_global = false;  // the spin planes stay in global memory (implies icm_global)
  PtGroup g{};  // device pointers filled at allocation
  std::vector<uint32_t> init_bits;
};

struct SolvePlan {
  const tamc_instance *inst = nullptr;
  tamc_pt_params P{};
  std::vector<float> betas;
  const int8_t *init_states = nullptr;
  const tamc_pt_state *resume = nullptr;  // counters, tags and energies to start from (nullable)
  uint32_t n = 0, B = 0, C = 0, T = 0, meas_init = 0, n_meas = 0, n_samp_events = 0, n_samp_slots = 0, nbytes = 0,
           gs_cap = 0, gs_words = 0, npairs = 0, nchi = 0;
  std::vector<GroupPlan> groups;
  float *d_betas = nullptr;
};

// Validate one solve and split it into groups (one per ICM chain pair); waves sizes the LDS check.
tamc_status plan_solve(SolvePlan &S, const tamc_instance *inst, const tamc_pt_params *P, const int8_t *init_states,
                       uint32_t waves);
// After the init kernel: overwrite the groups' ladder state with the imported PTState fields
// (tags, acceptance counters, round trips, diffusion histograms, cached energies).
tamc_status upload_resume(const SolvePlan &S);
// Dynamic LDS bytes the sweep kernel needs for group g at `waves` waves.
uint32_t lds_bytes_for(const SolvePlan &S, const GroupPlan &g, uint32_t waves);
// Register every device buffer of S in the arena (the ground-state trace in Ags and the betas in
// Ab when given).
void want_buffers(SolvePlan &S, Arena &A, Arena *Ags = nullptr, Arena *Ab = nullptr);
// After the arena is placed: upload the betas unless the caller does (on stream s, or
// synchronously when s is null) and fill the groups' PtGroup descriptors.
tamc_status fill_groups(SolvePlan &S, hipStream_t s, bool upload_betas = true);
// Host copy of a device range: result arrays inside it are read from `host`.
struct HostMirror {
  const char *host;
  const char *dev;
  size_t bytes;
};
// Gather one solve's group outputs into a tamc_pt_result laid out like the reference structures
// (m: optional host mirror of the arena holding them).
tamc_status assemble(const SolvePlan &S, tamc_pt_result **out, const HostMirror *m = nullptr);
// Waves per work group covering the instance's typical level width.
uint32_t default_waves(const tamc_instance *inst);

}  // namespace tamc
