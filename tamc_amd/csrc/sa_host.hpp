// sa_host.hpp -- the SA kernel's per-wave task record and launcher, shared by the one-shot SA
// entry points (sa.hip) and the serving loop (server.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "runtime_common.hpp"

namespace tamc {

// One wave: 64 restarts of one SA solve over sweeps [k0, k1) of its schedule.
struct SaTask {
  const uint32_t *info;     // [n][4] per site: first block, blocks, bias bits, row length
  const uint32_t *blocks;   // [blocks][16]: eight neighbour byte offsets (8 * col), eight coupling bits
  const float *betas;       // [K] the whole schedule
  float offset;
  uint32_t n, K, R, rbase;  // rbase: first restart of this wave
  uint32_t k0, k1;          // sweeps of this launch: k0 == 0 starts from the initial states, k0 > 0
                            // continues from spins_out; k1 < K leaves the planes in spins_out for
                            // the next launch, k1 == K ends with the final energies
  uint64_t seed;
  uint32_t *spins_out;      // [n][2] bit-plane halves of this wave's restarts (null: states_out)
  float *energy_out;        // [64]; with states_out: this wave's entries of a [R] array
  int8_t *states_out;       // null, or this wave's rows of a [R][n] int8 array (+1/-1)
};

// Restarts SaRunner::run / run_parallel keep: threads * floor(R / threads) for R > 1
// (src/sa.rs:147,161), R otherwise.
inline uint32_t sa_kept_replicas(const tamc_sa_params &P) {
  return P.num_replicas > 1 ? P.threads * (P.num_replicas / P.threads) : P.num_replicas;
}

// Checks of tamc_sa_run's parameters (SaRunner::new, src/sa.rs:90-108).
tamc_status sa_validate(const tamc_instance *inst, const tamc_sa_params &P);
// 32-bit words of LDS one SA wave uses for an instance of n sites (0 when it does not fit).
uint32_t sa_wave_lds_words(uint32_t n);
// Launch `ntasks` tasks (waves) whose instances have at most maxn sites, on stream s.
tamc_status sa_launch(const SaTask *d_tasks, uint32_t ntasks, uint32_t maxn, hipStream_t s);
// Result of one solve from its waves' final plane words [waves][n][2] and energies [waves][64].
tamc_sa_result *sa_assemble(uint32_t n, uint32_t R, const uint32_t *planes, const float *energies);

}  // namespace tamc
