// percolation.hip -- site percolation (src/percolation.rs:11-75, src/bin/percolate.rs) on the GPU.
// One work group per sample: occupancy draws, connected components of the occupied sites by
// minimum-label propagation with pointer jumping in LDS, the largest component's size, and whether
// an occupied site of negative weight shares a component with one of positive weight (the reference
// looks for a path by depth-first search from the former to the latter).
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "runtime_common.hpp"
#include "tamc_math.hpp"

using namespace tamc;

namespace {

constexpr uint32_t kFree = 0xFFFFFFFFu;      // unoccupied site
constexpr uint32_t kPercLdsSites = 16384u;   // two 32-bit words per site in LDS up to this size; larger
                                             // instances keep them in global scratch
constexpr size_t kPercScratchBytes = 256u << 20;  // global scratch of one launch (samples in batches)

struct PercArgs {
  const uint32_t *row_ptr, *col;  // coupling pattern (no self entries in Ising form)
  const float *weight;            // node weights (the diagonal entries)
  uint32_t n, key0, key1;
  uint64_t p_int;                 // Bernoulli threshold on a 64-bit draw; ~0 = always occupied
  uint32_t *max_cluster;          // [samples]
  uint8_t *percolates;            // [samples]
};

// occupancy of site i in sample k: rand's Bernoulli(p) on one 64-bit draw (v < p_int), the draw
// being words 0 and 1 of Philox4x32-10 keyed by the seed, counter {sample, site, 0, tag}
__device__ __forceinline__ bool occupied(const PercArgs &A, uint32_t k, uint32_t i) {
  if (A.p_int == ~0ull) return true;
  const U4 v = philox4x32_10(k, i, 0u, kTagPercolation, A.key0, A.key1);
  const uint64_t d = ((uint64_t)v.x << 32) | v.y;
  return d < A.p_int;
}

// Work group b handles sample k0 + b.  kG: its labels and counters are the slice [2n) of the global
// scratch at 2nb (an instance too large for LDS); otherwise they are in LDS.
template <bool kG>
__global__ __launch_bounds__(256) void percolation_kernel(PercArgs A, uint32_t k0, uint32_t *scratch) {
  extern __shared__ __align__(16) uint32_t lds[];
  const uint32_t k = k0 + blockIdx.x, n = A.n;
  uint32_t *label = kG ? scratch + 2u * (size_t)n * blockIdx.x : lds, *aux = label + n;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    label[i] = occupied(A, k, i) ? i : kFree;
    aux[i] = 0u;
  }
  __syncthreads();
  // components: labels only decrease, to the smallest site of the component
  for (;;) {
    int changed = 0;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
      uint32_t l = label[i];
      if (l == kFree) continue;
      const uint32_t l0 = l;
      for (uint32_t e = A.row_ptr[i]; e < A.row_ptr[i + 1]; ++e) {
        const uint32_t lj = label[A.col[e]];
        if (lj < l) l = lj;  // kFree is larger than every label
      }
      l = min(l, label[l]);  // one pointer jump
      if (l < l0) {
        atomicMin(&label[i], l);
        atomicMin(&label[l0], l);  // hook the old root as well
        changed = 1;
      }
    }
    if (!__syncthreads_or(changed)) break;
  }
  // component sizes and the components holding an occupied negative-weight site
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    const uint32_t l = label[i];
    if (l == kFree) continue;
    atomicAdd(&aux[l], 1u);
    if (A.weight[i] < 0.0f) atomicOr(&aux[l], 0x80000000u);
  }
  __syncthreads();
  __shared__ uint32_t best, perc;
  if (threadIdx.x == 0) {
    best = 0u;
    perc = 0u;
  }
  __syncthreads();
  uint32_t mine = 0u, p = 0u;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    const uint32_t l = label[i];
    if (l == kFree) continue;
    if (l == i) mine = max(mine, aux[i] & 0x7FFFFFFFu);
    if (A.weight[i] > 0.0f && (aux[l] & 0x80000000u)) p = 1u;
  }
  atomicMax(&best, mine);
  if (p) atomicOr(&perc, 1u);
  __syncthreads();
  if (threadIdx.x == 0) {
    A.max_cluster[k] = best;
    A.percolates[k] = (uint8_t)perc;
  }
}

}  // namespace

extern "C" tamc_status tamc_percolate(const tamc_instance *inst, double p, uint32_t num_samples, uint64_t seed,
                                      uint32_t *max_cluster, uint8_t *percolates) {
  clear_error();
  if (!inst || (num_samples && (!max_cluster || !percolates))) return fail(TAMC_E_INVALID_ARG, "null argument");
  if (!(p >= 0.0 && p <= 1.0)) return fail(TAMC_E_INVALID_ARG, "Bernoulli probability must be in [0, 1]");
  if (inst->qubo) return fail(TAMC_E_INVALID_ARG, "percolation takes an instance read in Ising form");
  tamc_status st = device_init();
  if (st) return st;
  if (num_samples == 0 || inst->n == 0) {
    for (uint32_t k = 0; k < num_samples; ++k) {
      max_cluster[k] = 0u;
      percolates[k] = 0u;
    }
    return TAMC_OK;
  }
  DeviceInstance *D = device_instance(inst, 1u, &st, kOrderColour);
  if (!D) return st;
  DevBuf out_c, out_p;
  TAMC_HIP(out_c.alloc((size_t)num_samples * 4));
  TAMC_HIP(out_p.alloc(num_samples));
  PercArgs A;
  A.row_ptr = (const uint32_t *)D->row_ptr.p;
  A.col = (const uint32_t *)D->col.p;
  A.weight = (const float *)D->bias.p;
  A.n = inst->n;
  A.key0 = (uint32_t)seed;
  A.key1 = (uint32_t)(seed >> 32);
  // rand 0.8 Bernoulli::new: p_int = p * 2^64, with p = 1 always true
  A.p_int = (p >= 1.0) ? ~0ull : (uint64_t)(p * 18446744073709551616.0);
  A.max_cluster = (uint32_t *)out_c.p;
  A.percolates = (uint8_t *)out_p.p;
  if (inst->n <= kPercLdsSites) {
    const size_t lds = (size_t)inst->n * 8u;
    if (lds > 65536u)
      TAMC_HIP(hipFuncSetAttribute((const void *)percolation_kernel<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
    hipLaunchKernelGGL(percolation_kernel<false>, dim3(num_samples), dim3(256), lds, 0, A, 0u, (uint32_t *)nullptr);
    TAMC_HIP(hipGetLastError());
  } else {
    // labels and counters in global scratch, as many samples per launch as it holds
    const size_t per = (size_t)inst->n * 8u;
    const uint32_t batch = (uint32_t)std::max<size_t>(1, std::min<size_t>(num_samples, kPercScratchBytes / per));
    DevBuf scr;
    TAMC_HIP(scr.alloc(per * batch));
    for (uint32_t k0 = 0; k0 < num_samples; k0 += batch) {
      const uint32_t nb = std::min(batch, num_samples - k0);
      hipLaunchKernelGGL(percolation_kernel<true>, dim3(nb), dim3(256), 0, 0, A, k0, (uint32_t *)scr.p);
      TAMC_HIP(hipGetLastError());
    }
  }
  TAMC_HIP(hipMemcpy(max_cluster, out_c.p, (size_t)num_samples * 4, hipMemcpyDeviceToHost));
  TAMC_HIP(hipMemcpy(percolates, out_p.p, num_samples, hipMemcpyDeviceToHost));
  return TAMC_OK;
}
