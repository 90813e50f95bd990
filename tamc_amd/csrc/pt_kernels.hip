// pt_kernels.hip -- gfx950 kernels of the production PT-ICM path.
//
// One work group owns one PtGroup (an instance's chain pair) for a range of sweeps.  Everything a
// sweep touches lives in LDS: bit-plane spins (one bit per replica), tracked energies, the slot
// permutation, and the ICM scratch.  Per sweep t (reference order, src/pt.rs:446-450):
//   1. Houdayer ICM on every beta slot >= lo of the chain pair     (src/pt.rs:31-77, 521-540)
//   2. replica exchange on each chain, one lane per chain          (tamc-core/src/pt.rs:120-144)
//   3. Metropolis sweep of every replica                           (src/ising.rs:225-233)
//   4. measurements and ground-state tracking after warm-up        (src/pt.rs:542-583)
//
// The sweep visits sites level by level: a site's level is one more than the highest level of the
// lower-numbered sites it is coupled to, so every site sees exactly the neighbour spins it would see
// in the reference's sequential 0..N-1 order, and spin trajectories equal the sequential
// algorithm's given the same uniforms.  Sites of a level are uncoupled; in ascending order they
// form groups of four, and the uniform for (sweep t, site, replica rho) is word `slot` of
// Philox4x32-10(ctr = {t, group, rho, 0}, key = seed): a function of the instance, not of the
// tiling or of batch placement.
#include <hip/hip_runtime.h>

#include <mutex>
#include <type_traits>

#include "pt_types.hpp"
#include "tamc_math.hpp"

#define TAMC_EXC_DECL static __device__ const
#include "generated/expf_exceptions.inc"

namespace tamc {

// Read-only device data is accessed through the constant address space: wave-uniform indices then
// become scalar (SMEM) loads, divergent ones ordinary global loads.
#define TAMC_AS1 __attribute__((address_space(1)))
#define TAMC_AS3 __attribute__((address_space(3)))
#define TAMC_AS4 __attribute__((address_space(4)))
template <class T>
__device__ __forceinline__ const TAMC_AS4 T *c4(const T *p) {
  return (const TAMC_AS4 T *)p;
}

__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint32_t popc64(uint64_t x) { return (uint32_t)__popcll(x); }
// number of set bits of m below this lane
__device__ __forceinline__ uint32_t mbcnt64(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ ExpfExceptions exc_table() { return ExpfExceptions{kExpfExcX, kExpfExcY, kExpfExcCount}; }

// Relative error band of the hardware estimate p = v_exp_f32(x * log2(e)) of expf(x): rounding
// x * log2(e) costs |x| 2^-23 and v_exp_f32 2^-23, together below 2^-16.3 for |x| <= 105 (expf
// underflows to 0 below -104).  The band 2^-14 leaves a factor 4.9 of margin.  The bound holds for
// normal results only; a denormal or flushed estimate (P = p * 2^24 < 2^-102) can only matter for
// k = 0, and the accept threshold fma(P, 1 - band, -2^-100) is <= 0 there, so such a case never
// takes the fast accept: it is rejected (k >= 1 > P) or goes to the exact path.
constexpr float kBand = 0x1p-14f;
__device__ __forceinline__ float accept_below(float P) { return __fmaf_rn(P, 1.0f - kBand, -0x1p-100f); }
__device__ __forceinline__ float reject_above(float P) { return P * (1.0f + kBand); }

// u < expf(x) for u = (w >> 8) * 2^-24, decided with the hardware exp2 unless u falls within the
// band around the estimate; inside the band the exact double-precision path decides.
__device__ __forceinline__ bool accept_u(float x, uint32_t w) {
  const uint32_t k = w >> 8;
  const float p = __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
  const float P = p * 16777216.0f;
  const float kf = (float)k;
  if (kf < accept_below(P)) return true;
  if (kf > reject_above(P)) return false;
  return accept_exact(x, k, exc_table());
}

// Tracked-energy increments are summed in 2^-32 fixed point: integer addition is associative, so
// the total does not depend on how quads are spread over waves or groups over a batch; for +-J and
// QUBO couplings every increment is exact and the result equals the sequential f32 sum bit for bit.
__device__ __forceinline__ long long to_fix(float x) { return (long long)rint((double)x * 4294967296.0); }
__device__ __forceinline__ float add_fix(float e, long long s) {
  return e + (float)((double)s * 2.3283064365386962890625e-10);
}

// spin bit (1 <=> -1) of replica r at site i in a bit-plane array of 32-bit halves
__device__ __forceinline__ uint32_t sbit(const uint32_t *sp, uint32_t G, uint32_t i, uint32_t r) {
  return (sp[(i * G + (r >> 6)) * 2u + ((r >> 5) & 1u)] >> (r & 31u)) & 1u;
}

#ifndef TAMC_SWEEP_KERNEL_ANY
// --------------------------------------------------------------------------------------------
// Device-side initialisation of a group: bit planes from init states or Philox, energies by
// energy_ref in the reference's order (src/ising.rs:438-454), identity permutation, fresh tags.
// One work group per PtGroup; lane = physical replica.
// init_bits: nullable [n][G][2] bit planes prepared by the host from caller init states.
__global__ void pt_init_kernel(const PtGroup *groups, const uint32_t *const *init_bits) {
  const TAMC_AS4 PtGroup &X = *c4(groups + blockIdx.x);
  const uint32_t n = X.n, G = X.G, R = X.R, B = X.B;
  const uint32_t key0 = (uint32_t)X.seed, key1 = (uint32_t)(X.seed >> 32);
  const uint32_t *ib = init_bits ? init_bits[blockIdx.x] : nullptr;
  // spins: thread per (site, g, half)
  for (uint32_t idx = threadIdx.x; idx < n * G * 2u; idx += blockDim.x) {
    uint32_t v;
    if (ib) {
      v = ib[idx];
    } else {
      const uint32_t i = idx / (G * 2u), rem = idx % (G * 2u);
      const uint32_t g = rem >> 1, half = rem & 1u;
      v = 0;
      for (uint32_t l = 0; l < 32u; ++l) {
        const uint32_t r = g * 64u + half * 32u + l;
        if (r < R) {
          const uint32_t rho = X.chain0 * B + r;
          const U4 o = philox4x32_10(0u, i >> 2, rho, kTagInit, key0, key1);
          // reference: spin +1 iff the top bit is set; bit plane stores 1 for -1
          v |= (((pick(o, i) >> 31) ^ 1u) & 1u) << l;
        }
      }
    }
    X.spins[idx] = v;
  }
  __syncthreads();
  __threadfence();
  for (uint32_t r = threadIdx.x; r < R; r += blockDim.x) {
    float total = X.offset;
    for (uint32_t i = 0; i < n; ++i) {
      float h = 0.0f;
      const float si = sbit(X.spins, G, i, r) ? -1.0f : 1.0f;
      h += c4(X.bias)[i];
      for (uint32_t e = c4(X.row_ptr)[i]; e < c4(X.row_ptr)[i + 1]; ++e) {
        const float sj = sbit(X.spins, G, c4(X.col)[e], r) ? -1.0f : 1.0f;
        h += __fmul_rn(c4(X.J)[e], sj) / 2.0f;
      }
      total = __fadd_rn(total, __fmul_rn(h, si));
    }
    X.energy[r] = total;
    X.drift[r] = 0;
  }
  if (threadIdx.x < kPhaseSlots && X.phase_ticks) X.phase_ticks[threadIdx.x] = 0ull;
  for (uint32_t k = threadIdx.x; k < X.nch * B; k += blockDim.x) {
    const uint32_t b = k % B;
    X.perm[k] = k;  // chain c slot b holds replica c*B+b
    X.tags[k] = (b == B - 1) ? 2u : ((b == 0) ? 1u : 0u);  // MaxBeta / MinBeta / None (PTState::new)
    X.nacc[k] = 0;
    X.hist[2 * k] = 0;
    X.hist[2 * k + 1] = 0;
  }
  if (threadIdx.x < X.nch) X.rt[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    X.gs_count[0] = 0;
    X.gs_min[0] = 0.0f;
    X.icm_moves[0] = 0ull;
  }
}

#endif  // TAMC_SWEEP_KERNEL_ANY

// --------------------------------------------------------------------------------------------
struct Ctx {
  const TAMC_AS4 PtGroup *X;
  uint32_t *sp;      // LDS bit planes
  float *energy, *beta_of, *dEx;
  long long *partial, *drift;
  uint32_t *perm, *pair_a, *pair_b, *pair_of, *seed_blk, *misc;
  uint32_t *moves;   // LDS count of ICM moves in this launch
  uint8_t *tags;
  uint32_t *nacc, *hist, *rt;
  uint32_t *qw, *cw, *blk, *uw;
  const uint16_t *nbr;  // ICM neighbour lists [n][D] (LDS copy or global)
  bool nbr_in_lds;
  long long *dEa;
  unsigned long long *ticks;  // LDS phase timers [kPhaseSlots] and the last timestamp
  uint32_t *fscr;             // ICM labelling: [p] a site of pair p's partial cluster
  unsigned long long *fmask;  // ICM flip masks by nibble of a cluster word: [g][k][v] = replicas, in
                              // plane word g, of the pairs 4k + j with bit j of v set
  uint32_t n, G, R, B, P, nch, lo, W, wave, lane, tid, nthreads, key0, key1, flags, nblk;
  uint32_t members, member;
  bool icm_lds;
};

// Phase timer of wave 0: adds the time since the previous lap to slot ph.
__device__ __forceinline__ void lap(const Ctx &c, int ph) {
  if (c.tid == 0u) {
    const unsigned long long now = wall_clock64();
    c.ticks[ph] += now - c.ticks[kPhaseSlots];
    c.ticks[kPhaseSlots] = now;
  }
}

// Teams (the work groups that share a group's sweep) meet after each colour class and each ICM
// flip.  The member's count of meetings is kept in misc[15] (misc[12..14] are the labelling's
// round flags).  Waits are bounded by kTeamSpinLimit polls.
constexpr uint32_t kTeamSpinLimit = 1u << 20;

// The words the members exchange are written and read with agent-scope atomic stores and loads,
// which are coherent across the XCDs' L2 caches without a fence: a fence at agent scope would
// write back and invalidate the whole L2 of the XCD, which every other group's sweep records
// share.  Each wave waits until its stores have completed before the work group arrives.
__device__ __forceinline__ void box_store(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long box_load(const unsigned long long *p) {
  return __hip_atomic_load(const_cast<unsigned long long *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Values cross between members as tagged slots: a 64-bit value takes two 64-bit slots, each holding
// 32 bits of the value beside the 32-bit epoch of the meeting that published it (the members count
// their meetings alike, from the count kept in team_ctr between launches).  A reader polls the
// slots it needs until both carry the meeting's epoch, so a meeting costs one round trip to memory
// and no counter.  A slot is published again only at a later meeting of the same kind, which the
// publisher reaches only after every member has published at an intervening meeting, that is,
// after every member has read the slot; ICM merges of consecutive chunks alternate between two
// buffers for this reason.
__device__ __forceinline__ void tag_put(unsigned long long *p, uint64_t v, uint32_t e) {
  box_store(p, (v & 0xFFFFFFFFull) | ((uint64_t)e << 32));
  box_store(p + 1, (v >> 32) | ((uint64_t)e << 32));
}

// The values published at meeting e in up to four tagged values (value j is read when bit j of
// `want` is set), their loads issued together: values already published cost one round trip.
// Waits are bounded: after kTeamSpinLimit polls the reader raises the team's abort flag, and once
// the flag is up no reader waits again (the values are then meaningless and the host reports the
// run as failed).
__device__ __forceinline__ void tag_get4(const unsigned long long *const src[4], uint32_t want, uint32_t e,
                                         uint32_t *abort_flag, uint64_t v[4]) {
#pragma unroll
  for (int j = 0; j < 4; ++j) v[j] = 0ull;
  for (uint32_t spins = 0; want != 0u; ++spins) {
    uint64_t a[4], b[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      a[j] = b[j] = 0ull;
      if ((want >> j) & 1u) {
        a[j] = box_load(src[j]);
        b[j] = box_load(src[j] + 1);
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (((want >> j) & 1u) && (uint32_t)(a[j] >> 32) == e && (uint32_t)(b[j] >> 32) == e) {
        v[j] = (a[j] & 0xFFFFFFFFull) | (b[j] << 32);
        want &= ~(1u << j);
      }
    }
    if (want == 0u) return;
    if ((spins & 15u) == 15u) {
      if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) return;
      if (spins > kTeamSpinLimit) {
        __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
      }
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// Sum over the K members of the value each published at meeting e in sums[m * stride], in member
// order, with the member's own value `mine` taken locally.
__device__ __forceinline__ long long tag_sum(const unsigned long long *sums, size_t stride, uint32_t K,
                                             uint32_t member, long long mine, uint32_t e, uint32_t *abort_flag) {
  long long s = 0;
  for (uint32_t m0 = 0; m0 < K; m0 += 4u) {
    const unsigned long long *src[4];
    uint32_t want = 0;
#pragma unroll
    for (uint32_t j = 0; j < 4u; ++j) {
      src[j] = sums + 2u * (m0 + j) * stride;
      if (m0 + j < K && m0 + j != member) want |= 1u << j;
    }
    uint64_t v[4];
    tag_get4(src, want, e, abort_flag, v);
#pragma unroll
    for (uint32_t j = 0; j < 4u; ++j)
      if (m0 + j < K) s += (m0 + j == member) ? mine : (long long)v[j];
  }
  return s;
}

// Inclusive prefix sum over the 64 lanes with DPP row shifts and row broadcasts.  Every lane of
// the wave must be active.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return v;
}

// value of v in lane `src` (0..63); every lane of the wave must be active
// Inclusive running maximum over the 64 lanes (same DPP pattern as wave_incl_scan; 0 is neutral).
// OR of v over the wave, in every lane
__device__ __forceinline__ uint32_t wave_or(uint32_t v) {
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// OR of a predicate over the work group with one barrier (__syncthreads_or takes two).  Round k
// of a loop writes slot k % 3 and clears slot (k + 2) % 3, which round k - 1 read before round k's
// barrier and round k + 2 writes after round k + 1's; slots 0 and 1 start at zero.
__device__ __forceinline__ bool block_any(uint32_t tid, uint32_t lane, uint32_t *slots, uint32_t k, bool v) {
  const uint64_t b = __builtin_amdgcn_ballot_w64(v);
  if (lane == 0u && b != 0ull) slots[k % 3u] = 1u;
  __syncthreads();
  const bool any = slots[k % 3u] != 0u;
  if (tid == 0u) slots[(k + 2u) % 3u] = 0u;
  return any;
}

__device__ __forceinline__ uint32_t wave_incl_max(uint32_t v) {
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false));  // row_shr:1
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false));  // row_shr:2
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false));  // row_shr:4
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false));  // row_shr:8
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));  // row_bcast:15
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return v;
}

// Minimum over the 64 lanes, in every lane (every lane of the wave must be active): the same DPP
// pattern as wave_incl_scan with +inf (or all ones) for lanes that have no source, then lane 63.
__device__ __forceinline__ float wave_min_f(float x) {
  const int inf = 0x7F800000;
  int v = __float_as_int(x);
  v = __float_as_int(fminf(__int_as_float(v), __int_as_float(__builtin_amdgcn_update_dpp(inf, v, 0x111, 0xf, 0xf, false))));
  v = __float_as_int(fminf(__int_as_float(v), __int_as_float(__builtin_amdgcn_update_dpp(inf, v, 0x112, 0xf, 0xf, false))));
  v = __float_as_int(fminf(__int_as_float(v), __int_as_float(__builtin_amdgcn_update_dpp(inf, v, 0x114, 0xf, 0xf, false))));
  v = __float_as_int(fminf(__int_as_float(v), __int_as_float(__builtin_amdgcn_update_dpp(inf, v, 0x118, 0xf, 0xf, false))));
  v = __float_as_int(fminf(__int_as_float(v), __int_as_float(__builtin_amdgcn_update_dpp(inf, v, 0x142, 0xa, 0xf, false))));
  v = __float_as_int(fminf(__int_as_float(v), __int_as_float(__builtin_amdgcn_update_dpp(inf, v, 0x143, 0xc, 0xf, false))));
  return __int_as_float(__builtin_amdgcn_readlane(v, 63));
}
__device__ __forceinline__ uint32_t wave_min_u(uint32_t x) {
  const int none = -1;
  int v = (int)x;
  v = (int)min((uint32_t)v, (uint32_t)__builtin_amdgcn_update_dpp(none, v, 0x111, 0xf, 0xf, false));
  v = (int)min((uint32_t)v, (uint32_t)__builtin_amdgcn_update_dpp(none, v, 0x112, 0xf, 0xf, false));
  v = (int)min((uint32_t)v, (uint32_t)__builtin_amdgcn_update_dpp(none, v, 0x114, 0xf, 0xf, false));
  v = (int)min((uint32_t)v, (uint32_t)__builtin_amdgcn_update_dpp(none, v, 0x118, 0xf, 0xf, false));
  v = (int)min((uint32_t)v, (uint32_t)__builtin_amdgcn_update_dpp(none, v, 0x142, 0xa, 0xf, false));
  v = (int)min((uint32_t)v, (uint32_t)__builtin_amdgcn_update_dpp(none, v, 0x143, 0xc, 0xf, false));
  return (uint32_t)__builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ uint32_t lane_value(uint32_t v, uint32_t src) {
  return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)v);
}

// index of the r-th set bit (r < popcount(x)) of x
__device__ __forceinline__ uint32_t select_bit(uint32_t x, uint32_t r) {
  uint32_t p = 0, c;
  c = (uint32_t)__popc(x & 0xFFFFu);
  if (r >= c) { r -= c; x >>= 16; p += 16; }
  c = (uint32_t)__popc(x & 0xFFu);
  if (r >= c) { r -= c; x >>= 8; p += 8; }
  c = (uint32_t)__popc(x & 0xFu);
  if (r >= c) { r -= c; x >>= 4; p += 4; }
  c = (uint32_t)__popc(x & 0x3u);
  if (r >= c) { r -= c; x >>= 2; p += 2; }
  if (r >= (x & 1u)) p += 1;
  return p;
}

// ICM scratch -- overlap words qw, cluster words cw, per-block counts blk -- and the u16
// neighbour lists, with address spaces fixed at compile time: the scratch is in LDS (kLds) or in
// HBM for groups too large for LDS, and the neighbour lists are an LDS copy or the HBM table.
// Typed pointers make every access a ds_* or global_* instruction instead of a flat one (flat
// accesses wait on both the LDS and the vector-memory counters).
template <bool kLds>
struct IcmMem {
  using U32 = std::conditional_t<kLds, TAMC_AS3 uint32_t, TAMC_AS1 uint32_t>;
  U32 *qw, *cw, *blk;
  const TAMC_AS3 uint16_t *nbr_lds;  // null when the lists are not in LDS
  const TAMC_AS4 uint16_t *nbr_glb;  // null when the instance has no fixed-stride lists
  __device__ __forceinline__ uint32_t nbr(size_t k) const { return nbr_lds ? nbr_lds[k] : nbr_glb[k]; }
};

// ICM step (d): swap the cluster spins of both replicas of every pair (= flip both) and sum the
// first replica's energy change
//   dE_a(p) = sum_{i in cluster p} -2 s_a(i) (h_i + sum_{j not in cluster p} J_ij s_a(j));
// the second replica's change is -dE_a (the move conserves E_a + E_b).  A wave takes 64 sites at
// a time, lane = site.  A lane in some cluster loads its site's data once (bias, row, couplings,
// the neighbours' cluster words and, with one plane word per site, the plane words of the site
// and its neighbours), then steps through its pairs one per iteration, adding each term into the
// pair's fixed-point total.  Then the lane applies the site's flip mask.  Pair p only reads
// replica a_p at neighbours outside cluster p and only flips replicas a_p, b_p at sites inside
// it, so a block is flipped as soon as its terms are done, and plane words loaded before other
// blocks' flips still hold every bit a term reads.
template <int D, bool kLds>
__device__ __forceinline__ void icm_flip(Ctx &c, const IcmMem<kLds> &M) {
  const TAMC_AS4 PtGroup &X = *c.X;
  const uint32_t P = c.P, n = c.n, G = c.G;
  constexpr int kD = D > 0 ? D : 1;
  // the terms of the pairs in `pmask` for the sites of block blk
  auto terms = [&](uint32_t blk, uint32_t cwi, uint32_t pmask) {
    const uint32_t i = blk * 64u + c.lane;
    uint32_t rem = cwi & pmask;
    if (!ballot(rem != 0u)) return;
    float bias = 0.0f, J[kD];
    uint32_t nj[kD], cwj[kD];
    uint64_t wi = 0ull, wj[kD];
    if (rem) {
      bias = c4(X.bias)[i];
      if constexpr (D > 0) {
        if constexpr (D <= 4) {
          const uint64_t w = M.nbr_lds ? reinterpret_cast<const TAMC_AS3 uint64_t *>(M.nbr_lds)[i]
                                       : reinterpret_cast<const TAMC_AS4 uint64_t *>(M.nbr_glb)[i];
#pragma unroll
          for (int k = 0; k < D; ++k) nj[k] = (uint32_t)(w >> (16 * k)) & 0xFFFFu;
        } else {
#pragma unroll
          for (int k = 0; k < D; ++k) nj[k] = M.nbr((size_t)i * nbr_stride(D) + k);
        }
#pragma unroll
        for (int k = 0; k < D; ++k) {
          J[k] = c4(X.ell_J)[(size_t)i * D + k];
          cwj[k] = M.cw[nj[k]];
          if (G == 1u) wj[k] = *reinterpret_cast<const uint64_t *>(c.sp + 2u * nj[k]);
        }
      }
      if (G == 1u) wi = *reinterpret_cast<const uint64_t *>(c.sp + 2u * i);
    }
    while (rem) {
      const uint32_t p = (uint32_t)__builtin_ctz(rem);
      rem &= rem - 1u;
      const uint32_t a = c.pair_a[p];
      float hh = bias;
      uint32_t si;
      if constexpr (D > 0) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
          if (!((cwj[k] >> p) & 1u)) {
            const uint32_t sj = (G == 1u) ? (uint32_t)(wj[k] >> a) & 1u : sbit(c.sp, G, nj[k], a);
            hh += sj ? -J[k] : J[k];
          }
        }
      } else {
        const uint32_t e0 = c4(X.row_ptr)[i], e1 = c4(X.row_ptr)[i + 1];
        for (uint32_t e = e0; e < e1; ++e) {
          const uint32_t j = c4(X.col)[e];
          if (!((M.cw[j] >> p) & 1u)) {
            const float Je = c4(X.J)[e];
            hh += sbit(c.sp, G, j, a) ? -Je : Je;
          }
        }
      }
      si = (G == 1u) ? (uint32_t)(wi >> a) & 1u : sbit(c.sp, G, i, a);
      const float term = si ? 2.0f * hh : -2.0f * hh;
      atomicAdd((unsigned long long *)&c.dEa[p], (unsigned long long)to_fix(term));
    }
  };
  // lane = site: flip replicas a_p, b_p for every pair p whose cluster holds the site
  auto flip_block = [&](uint32_t blk, uint32_t cwi) {
    const uint32_t i_lane = blk * 64u + c.lane;
    if (G == 1u) {
      unsigned long long mg = 0ull;
      for (uint32_t k = 0; 4u * k < P; ++k) mg |= c.fmask[16u * k + ((cwi >> (4u * k)) & 15u)];
      if (mg) {
        unsigned long long *w = reinterpret_cast<unsigned long long *>(c.sp) + i_lane;
        *w ^= mg;
      }
      return;
    }
    for (uint32_t g = 0; g < G; ++g) {
      uint64_t mg = 0ull;
      for (uint32_t k = 0; 4u * k < P; ++k) mg |= c.fmask[(g * (kMaxPairs / 4u) + k) * 16u + ((cwi >> (4u * k)) & 15u)];
      if (mg) {
        c.sp[(i_lane * G + g) * 2u] ^= (uint32_t)mg;
        c.sp[(i_lane * G + g) * 2u + 1u] ^= (uint32_t)(mg >> 32);
      }
    }
  };
  // the work group's blocks: all, or in a team the member's blocks member, member + K, ...
  const uint32_t K = c.members;
  const uint32_t nbm = c.nblk > c.member ? (c.nblk - c.member + K - 1u) / K : 0u;
  if (nbm >= c.W) {
    // a wave owns whole blocks: terms, then flips, block after block
    for (uint32_t j = c.wave; j < nbm; j += c.W) {
      const uint32_t blk = c.member + j * K;
      const uint32_t i_lane = blk * 64u + c.lane;
      const uint32_t cwi = (i_lane < n) ? M.cw[i_lane] : 0u;
      if (!ballot(cwi != 0u)) continue;
      terms(blk, cwi, 0xFFFFFFFFu);
      flip_block(blk, cwi);
    }
  } else {
    // fewer blocks than waves: every wave takes every block, for the pairs p = wave (mod W); the
    // flips wait until every term is done
    uint32_t pmask = 0u;
    for (uint32_t p = c.wave; p < P; p += c.W) pmask |= 1u << p;
    for (uint32_t j = 0; j < nbm; ++j) {
      const uint32_t blk = c.member + j * K;
      const uint32_t i_lane = blk * 64u + c.lane;
      const uint32_t cwi = (i_lane < n) ? M.cw[i_lane] : 0u;
      terms(blk, cwi, pmask);
    }
    __syncthreads();
    for (uint32_t j = c.wave; j < nbm; j += c.W) {
      const uint32_t blk = c.member + j * K;
      const uint32_t i_lane = blk * 64u + c.lane;
      const uint32_t cwi = (i_lane < n) ? M.cw[i_lane] : 0u;
      if (ballot(cwi != 0u)) flip_block(blk, cwi);
    }
  }
  // In a team the member flips the other members' blocks too, without their terms: every member
  // holds every cluster, and a flip of pair q touches replicas a_q and b_q only, so it changes no
  // bit that a term of another pair reads (and the member's own sites are flipped by itself).
  if (K > 1u) {
    for (uint32_t blk = c.wave; blk < c.nblk; blk += c.W) {
      if (blk % K == c.member) continue;
      const uint32_t i_lane = blk * 64u + c.lane;
      const uint32_t cwi = (i_lane < n) ? M.cw[i_lane] : 0u;
      if (ballot(cwi != 0u)) flip_block(blk, cwi);
    }
  }
}

// In a team every member sums the terms of its own 64-site blocks (and flips every block): the
// members exchange only their per-pair sums.  Two buffers alternate by ICM chunk, so a member that
// runs ahead into the next chunk does not overwrite sums another member is still reading.
__device__ __forceinline__ void icm_team_merge(Ctx &c) {
  const TAMC_AS4 PtGroup &X = *c.X;
  const uint32_t K = c.members, P = c.P;
  const uint32_t par = ((c.lo - X.lo) / kMaxPairs) & 1u;
  unsigned long long *Dp = X.team_box + team_pair_slot(c.n, c.G, c.R, K, par);
  uint32_t *abort_flag = X.team_ctr + 1;
  __syncthreads();
  const uint32_t e = c.misc[15] + 1u;
  for (uint32_t p = c.tid; p < P; p += c.nthreads)
    tag_put(Dp + 2u * ((size_t)c.member * kMaxPairs + p), (unsigned long long)c.dEa[p], e);
  // then sums all the members' terms in member order, so that every member gets the same totals
  for (uint32_t p = c.tid; p < P; p += c.nthreads)
    c.dEa[p] = tag_sum(Dp + 2u * p, kMaxPairs, K, c.member, c.dEa[p], e, abort_flag);
  __syncthreads();
  if (c.tid == 0u) c.misc[15] = e;
}

// ICM step (c) for the pairs whose clusters still grow after kLabelAfterPasses passes (long
// clusters take one bit-parallel round per hop): each such pair's cluster is completed as the
// connected component, in the pair's q<0 sites, of the sites already in it.
//  * The q words are no longer read, so they hold the labels of two pairs at a time (16 bits
//    each, 0xFFFF outside the pair's subgraph); the cluster-word bit of a straggling pair holds
//    the site's q bit until the pair's component is known, and one site of each partial cluster
//    is kept.
//  * Components come from label hooking with shortcutting (Shiloach-Vishkin style).  Labels only
//    decrease and always name a site of the same component.  A hook round makes every subgraph
//    edge whose ends carry different labels lower the larger label's site (a root) to the smaller
//    label; a shortcut round then points every label at its root.  When a hook round finds
//    nothing to do, labels are constant on components, and the cluster is every member site with
//    the label of the kept site.
constexpr uint32_t kLabelAfterPasses = 10u, kMaxLabelPairs = 16u, kNoLabel = 0xFFFFu;

// lowers the 16-bit half `h` (0 low, 1 high) of an LDS word to at most l
template <class U32P>
__device__ __forceinline__ void min_half(U32P w, uint32_t h, uint32_t l) {
  const uint32_t sh = 16u * h, mask = 0xFFFFu << sh;
  uint32_t old = *w;
  while (((old >> sh) & 0xFFFFu) > l) {
    const uint32_t seen = atomicCAS((uint32_t *)w, old, (old & ~mask) | (l << sh));
    if (seen == old) break;
    old = seen;
  }
}

// spt: sites per thread (site tid + s * nthreads for s < spt); S only sets how many sites' rows are
// requested together.
template <uint32_t S, uint32_t RW, bool kLds>
__device__ __forceinline__ void icm_label_pairs(Ctx &c, const IcmMem<kLds> &M, uint32_t n, uint32_t pairs,
                                                uint32_t spt) {
  constexpr uint32_t kLabelChunk = (RW == 1u && S >= 8u) ? 8u : 4u;
  auto row = [&](uint32_t i, uint32_t k) -> uint64_t {
    return M.nbr_lds ? reinterpret_cast<const TAMC_AS3 uint64_t *>(M.nbr_lds)[RW * i + k]
                     : reinterpret_cast<const TAMC_AS4 uint64_t *>(M.nbr_glb)[RW * i + k];
  };
  uint32_t *keep = c.fscr;  // [p]: a site of pair p's partial cluster
  uint32_t *any_slots = c.misc + 12, rnd = 0u;
  if (c.tid < 32u) keep[c.tid] = 0xFFFFFFFFu;
  if (c.tid < 2u) any_slots[c.tid] = 0u;
  __syncthreads();
  for (uint32_t i = c.tid; i < n; i += c.nthreads) {
    const uint32_t cw = M.cw[i], q = M.qw[i];
    for (uint32_t m = cw & pairs; m; m &= m - 1u) atomicMin(&keep[__builtin_ctz(m)], i);
    M.cw[i] = (cw & ~pairs) | (q & pairs);
  }
  __syncthreads();
  for (uint32_t m = pairs; m;) {
    const uint32_t pa = (uint32_t)__builtin_ctz(m);
    m &= m - 1u;
    const uint32_t pb = m ? (uint32_t)__builtin_ctz(m) : 32u;  // 32: no second pair
    if (m) m &= m - 1u;
    const uint32_t ma = 1u << pa, mb = pb < 32u ? 1u << pb : 0u;
    for (uint32_t i = c.tid; i < n; i += c.nthreads) {
      const uint32_t cw = M.cw[i];
      M.qw[i] = ((cw & ma) ? i : kNoLabel) | (((cw & mb) ? i : kNoLabel) << 16);
    }
    __syncthreads();
    for (;;) {
      if (c.tid == 0u) c.ticks[kPhIcmPasses] += 1u;
      int hooked = 0;
      // sites kLabelChunk at a time, their rows requested together
#pragma unroll 1
      for (uint32_t s0 = 0; s0 < spt; s0 += kLabelChunk) {
        uint64_t nb[kLabelChunk][RW];
#pragma unroll
        for (uint32_t t = 0; t < kLabelChunk; ++t) {
          const uint32_t i = c.tid + (s0 + t) * c.nthreads, ii = i < n ? i : 0u;
#pragma unroll
          for (uint32_t r = 0; r < RW; ++r) nb[t][r] = row(ii, r);
        }
#pragma unroll
        for (uint32_t t = 0; t < kLabelChunk; ++t) {
          const uint32_t i = c.tid + (s0 + t) * c.nthreads;
          const uint32_t w = i < n ? M.qw[i] : 0xFFFFFFFFu;
          if (w != 0xFFFFFFFFu) {
            const uint32_t la = w & 0xFFFFu, lb = w >> 16;
            uint32_t mina = la, minb = lb;
#pragma unroll
            for (uint32_t r = 0; r < RW; ++r)
#pragma unroll
              for (uint32_t e = 0; e < 4u; ++e) {
                const uint32_t x = M.qw[(nb[t][r] >> (16u * e)) & 0xFFFFu];
                mina = min(mina, x & 0xFFFFu);
                minb = min(minb, x >> 16);
              }
            // only a member hooks (a non-member's label 0xFFFF names no site)
            if (la != kNoLabel && mina < la) {
              min_half(&M.qw[la], 0u, mina);
              hooked = 1;
            }
            if (lb != kNoLabel && minb < lb) {
              min_half(&M.qw[lb], 1u, minb);
              hooked = 1;
            }
          }
        }
      }
      if (!block_any(c.tid, c.lane, any_slots, rnd++, hooked != 0)) break;
      for (uint32_t i = c.tid; i < n; i += c.nthreads) {
        const uint32_t w = M.qw[i];
        if (w == 0xFFFFFFFFu) continue;
        uint32_t la = w & 0xFFFFu, lb = w >> 16, up;
        if (la != kNoLabel)
          while ((up = M.qw[la] & 0xFFFFu) != la) la = up;
        if (lb != kNoLabel)
          while ((up = M.qw[lb] >> 16) != lb) lb = up;
        const uint32_t nw = la | (lb << 16);
        if (nw != w) M.qw[i] = nw;
      }
      __syncthreads();
    }
    const uint32_t ra = M.qw[keep[pa]] & 0xFFFFu;
    const uint32_t rb = mb ? M.qw[keep[pb]] >> 16 : kNoLabel;
    for (uint32_t i = c.tid; i < n; i += c.nthreads) {
      const uint32_t w = M.qw[i], cw = M.cw[i];
      uint32_t out = cw & ~(ma | mb);
      if ((cw & ma) && (w & 0xFFFFu) == ra) out |= ma;
      if ((cw & mb) && (w >> 16) == rb) out |= mb;
      if (out != cw) M.cw[i] = out;
    }
    __syncthreads();
  }
}

// ICM step (c) with every thread's sites in registers: thread tid owns sites tid + s * nthreads
// (s < S) and is the only writer of their cluster words, so it keeps the cluster word of each in a
// register (up to 8 sites; with 16 the words are re-read beside q, which keeps the kernel's
// register budget for the other phases), and for S == 4 with rows of width <= 4 also q and the
// padded neighbour row (otherwise
// these are re-read each round, beside each other).  RW is the row length in 64-bit words (four
// u16 entries each; padding entries hold the site itself and are not read).  With re-read rows a
// round takes the sites four at a time: it gathers the neighbours' cluster words of the unfinished
// ones, so those loads are in flight together, then joins and writes back; with rows in registers
// every site is joined and written before the next is read.  Four rounds per barrier.
template <uint32_t S, uint32_t RW, bool kLds>
__device__ __forceinline__ void icm_grow_registers(Ctx &c, const IcmMem<kLds> &M, uint32_t n, uint32_t D) {
  constexpr bool kRegs = S == 4u && RW == 1u;
  constexpr uint32_t kR = kRegs ? S : 1u;
  constexpr uint32_t kChunk = kRegs ? 1u : 4u;
  constexpr uint32_t kWin = kRegs ? 1u : (S * RW < 12u ? S * RW : 12u);
  constexpr bool kCvRegs = S <= 8u;  // with 16 sites the cluster words are re-read from LDS
  constexpr uint32_t kC = kCvRegs ? S : 1u;
  // q is read-only here: a constant view lets its loads move past the cluster-word stores
  using QPtr = std::conditional_t<kLds, const TAMC_AS3 uint32_t *, const TAMC_AS4 uint32_t *>;
  const QPtr qw = (QPtr)M.qw;
  uint32_t cv[kC], qv[kR];
  uint64_t rw[kR];
  auto row = [&](uint32_t i, uint32_t k) -> uint64_t {
    return M.nbr_lds ? reinterpret_cast<const TAMC_AS3 uint64_t *>(M.nbr_lds)[RW * i + k]
                     : reinterpret_cast<const TAMC_AS4 uint64_t *>(M.nbr_glb)[RW * i + k];
  };
  // cluster words of the neighbours in one row word (its first `cnt` entries, cnt >= 3)
  auto gather = [&](uint64_t w, uint32_t cnt) -> uint32_t {
    uint32_t x = M.cw[w & 0xFFFFu] | M.cw[(w >> 16) & 0xFFFFu] | M.cw[(w >> 32) & 0xFFFFu];
    if (cnt >= 4u) x |= M.cw[w >> 48];
    return x;
  };
#pragma unroll
  for (uint32_t s = 0; s < S; ++s) {
    const uint32_t i = c.tid + s * c.nthreads;
    uint32_t q = 0u, cw = 0u;
    uint64_t r = 0ull;
    if (i < n) {
      cw = M.cw[i];
      if (kRegs) {
        q = qw[i];
        r = row(i, 0u);
      }
    }
    if (kCvRegs) cv[s % kC] = cw;
    if (kRegs) {
      qv[s % kR] = q;
      rw[s % kR] = r;
    }
  }
  // pass k ORs the pairs that gained sites into misc[8 + k % 3]; the slot of pass k - 1 is
  // cleared after pass k's barrier (it is next written in pass k + 2)
  if (c.tid < 3u) c.misc[8u + c.tid] = 0u;
  __syncthreads();
  for (uint32_t pass = 0;; ++pass) {
    if (c.tid == 0u) c.ticks[kPhIcmPasses] += 1u;
    uint32_t gain = 0u;
#pragma unroll 1
    for (uint32_t rep = 0; rep < 4u; ++rep) {
      uint32_t base = c.tid;
      asm volatile("" : "+v"(base));  // keeps the site addresses out of registers between rounds
      // re-read rows stream through a window of kWin row words: the first kWin are requested at the
      // start of the round and each slot is refilled with the word kWin further on as soon as it is
      // read, so the loads stay ahead of their use; sites past n read row 0 and are never joined
      // (q = 0)
      uint64_t pre[kWin];
      auto fetch = [&](uint32_t wd) {  // row word wd of the round: site wd / RW, word wd % RW
        const uint32_t i = base + (wd / RW) * c.nthreads;
        pre[wd % kWin] = row(i < n ? i : 0u, wd % RW);
      };
      if (!kRegs) {
#pragma unroll
        for (uint32_t wd = 0; wd < kWin; ++wd) fetch(wd);
      }
#pragma unroll
      for (uint32_t s0 = 0; s0 < S; s0 += kChunk) {
        uint32_t nb[kChunk], qk[kChunk], ck[kChunk];
#pragma unroll
        for (uint32_t k = 0; k < kChunk; ++k) {
          const uint32_t s = s0 + k, i = base + s * c.nthreads;
          uint32_t q = 0u;
          uint64_t w0 = 0ull, w1 = 0ull;
          if (kRegs) {
            q = qv[s % kR];
            w0 = rw[s % kR];
          } else {
            if (i < n) q = qw[i];
            w0 = pre[(s * RW) % kWin];
            if (RW == 2u) w1 = pre[(s * RW + 1u) % kWin];
#pragma unroll
            for (uint32_t k2 = 0; k2 < RW; ++k2)
              if (s * RW + k2 + kWin < S * RW) fetch(s * RW + k2 + kWin);
          }
          const uint32_t cur = kCvRegs ? cv[s % kC] : (i < n ? M.cw[i] : 0u);
          uint32_t x = 0u;
          if (cur != q) {  // something left to join (never for q == 0 and padding)
            if (RW == 1u) {
              x = gather(w0, D);
            } else {
              x = gather(w0, 4u) | M.cw[w1 & 0xFFFFu];
              if (D >= 6u) x |= M.cw[(w1 >> 16) & 0xFFFFu];
              if (D >= 7u) x |= M.cw[(w1 >> 32) & 0xFFFFu];
              if (D == 8u) x |= M.cw[w1 >> 48];
            }
          }
          nb[k] = x;
          qk[k] = q;
          ck[k] = cur;
        }
#pragma unroll
        for (uint32_t k = 0; k < kChunk; ++k) {
          const uint32_t s = s0 + k;
          const uint32_t nc = ck[k] | (qk[k] & nb[k]);
          if (nc != ck[k]) {
            M.cw[base + s * c.nthreads] = nc;
            gain |= nc ^ ck[k];
          }
          if (kCvRegs) cv[s % kC] = nc;
        }
      }
    }
    const uint32_t g = wave_or(gain);
    uint32_t *slot = c.misc + 8u + pass % 3u;
    if (c.lane == 0u && g) atomicOr(slot, g);
    __syncthreads();
    const uint32_t grew = *slot;
    if (c.tid == 0u) c.misc[8u + (pass + 2u) % 3u] = 0u;
    if (grew == 0u) break;
  }
}

// ICM step (c) over compact tiles: the instance's sites are grouped into tiles of at most 64 that
// are balls of the coupling graph (icm_tiles, padded with kNone), and a wave takes a tile at a
// time, lane = site.  Within the tile the wave joins sites until nothing in the tile changes (or
// kTileIters rounds), reading the cluster words its own lanes have just written, so a cluster
// crosses a whole tile in one visit instead of one hop per round.  Every site belongs to one tile
// and a tile to one wave per pass, so each cluster word has a single writer; cluster words only
// gain bits, and the passes repeat until one changes nothing, which reaches the same fixpoint as
// any other visiting order.
constexpr uint32_t kTileIters = 24u;

template <uint32_t S, uint32_t RW, bool kLds>
__device__ __forceinline__ void icm_grow_tiles(Ctx &c, const IcmMem<kLds> &M, uint32_t n, uint32_t D) {
  const TAMC_AS4 PtGroup &X = *c.X;
  const uint32_t nt = X.icm_ntiles;
  // a tile's rows: the site-order copy in LDS when the lists live there, else the tile-order copy
  // (one coalesced load, independent of the site id)
  auto rows = [&](uint32_t pos, uint32_t i, uint64_t &w0, uint64_t &w1) {
    if (M.nbr_lds) {
      const uint32_t ii = i < n ? i : 0u;
      w0 = reinterpret_cast<const TAMC_AS3 uint64_t *>(M.nbr_lds)[RW * ii];
      if (RW == 2u) w1 = reinterpret_cast<const TAMC_AS3 uint64_t *>(M.nbr_lds)[RW * ii + 1u];
    } else {
      w0 = reinterpret_cast<const TAMC_AS4 uint64_t *>(c4(X.icm_trows))[RW * pos];
      if (RW == 2u) w1 = reinterpret_cast<const TAMC_AS4 uint64_t *>(c4(X.icm_trows))[RW * pos + 1u];
    }
  };
  if (c.tid < 3u) c.misc[8u + c.tid] = 0u;
  __syncthreads();
  for (uint32_t pass = 0;; ++pass) {
    if (c.tid == 0u) c.ticks[kPhIcmPasses] += 1u;
    uint32_t gain = 0u;
    // the next tile's site ids and rows are requested before the current tile is worked on
    uint32_t tl = c.wave, i = kNone;
    uint64_t w0 = 0ull, w1 = 0ull;
    if (tl < nt) {
      i = c4(X.icm_tiles)[tl * 64u + c.lane];
      rows(tl * 64u + c.lane, i, w0, w1);
    }
    while (tl < nt) {
      const uint32_t tn = tl + c.W;
      uint32_t ni = kNone;
      uint64_t n0 = 0ull, n1 = 0ull;
      if (tn < nt) {
        ni = c4(X.icm_tiles)[tn * 64u + c.lane];
        rows(tn * 64u + c.lane, ni, n0, n1);
      }
      const bool in = i < n;
      uint32_t q = 0u, cw = 0u;
      if (in) {
        q = M.qw[i];
        cw = M.cw[i];
      }
      if (ballot(cw != q)) {  // else every site of the tile joined, or none can
#pragma unroll 1
        for (uint32_t it = 0; it < kTileIters; ++it) {
          uint32_t x = 0u;
          if (cw != q) {
            x = M.cw[w0 & 0xFFFFu] | M.cw[(w0 >> 16) & 0xFFFFu] | M.cw[(w0 >> 32) & 0xFFFFu];
            if (D >= 4u) x |= M.cw[w0 >> 48];
            if (RW == 2u) {
              x |= M.cw[w1 & 0xFFFFu];
              if (D >= 6u) x |= M.cw[(w1 >> 16) & 0xFFFFu];
              if (D >= 7u) x |= M.cw[(w1 >> 32) & 0xFFFFu];
              if (D == 8u) x |= M.cw[w1 >> 48];
            }
          }
          const uint32_t nc = cw | (q & x);
          const bool ch = nc != cw;
          if (ch) {
            M.cw[i] = nc;
            gain |= nc ^ cw;
            cw = nc;
          }
          if (!ballot(ch)) break;
        }
      }
      tl = tn;
      i = ni;
      w0 = n0;
      w1 = n1;
    }
    const uint32_t g = wave_or(gain);
    uint32_t *slot = c.misc + 8u + pass % 3u;
    if (c.lane == 0u && g) atomicOr(slot, g);
    __syncthreads();
    const uint32_t grew = *slot;
    if (c.tid == 0u) c.misc[8u + (pass + 2u) % 3u] = 0u;
    if (grew == 0u) break;
  }
}

// ICM step (c) on periodic lattices with row-major sites, the L x L square lattice cut into 8 x 8
// tiles or the L x L x L cubic lattice cut into 4 x 4 x 4 tiles (the last tile along an axis holds
// the remaining L - E (T - 1) sites, E the tile edge, T = ceil(L / E) tiles per axis).  For each
// tile and pair the pair's q and cluster bits become 64-bit bitboards: bit 8 y + x of tile (ty, tx)
// is site (8 ty + y) L + 8 tx + x, bit 16 z + 4 y + x of tile (tz, ty, tx) is site
// (4 tz + z) L^2 + (4 ty + y) L + 4 tx + x; positions past the lattice's edge hold no q bit.  The
// bitboards live in bb (LDS): q of tile t and pair p at word 2 (t P + p), the cluster bits
// 2 P T^dim words later.  A thread takes one pair and a run of consecutive tiles and, per tile, joins sites
// by occluded fills along each axis until nothing changes, the adjacent tiles' faces coming in as
// a halo; passes over all (tile, pair) repeat until one changes nothing, which reaches the same
// fixpoint as the other orders.  The cluster words are then rebuilt from the bitboards (the q
// words are not read again in the sweep's ICM).
template <int kDim>
__host__ __device__ __forceinline__ uint32_t lattice_tiles(uint32_t L) {
  constexpr uint32_t E = kDim == 2 ? 8u : 4u;
  const uint32_t T = (L + E - 1u) / E;
  return kDim == 2 ? T * T : T * T * T;
}

template <int kDim, class SrcU32>
__device__ __forceinline__ void icm_grow_lattice(Ctx &c, SrcU32 *qw, SrcU32 *cw, uint32_t L,
                                                 TAMC_AS3 uint64_t *bb) {
  constexpr uint32_t E = kDim == 2 ? 8u : 4u;  // tile edge
  const uint32_t T = (L + E - 1u) / E, nt = lattice_tiles<kDim>(L), P = c.P, last = L - E * (T - 1u);
  const uint32_t L2 = L * L;
  TAMC_AS3 uint64_t *bq = bb, *bcl = bb + (size_t)nt * P;
  auto ext = [&](uint32_t tc) { return tc + 1u == T ? last : E; };  // sites of a tile along an axis
  // lane l's site in tile (tz, ty, tx), or kNone past the lattice's edge
  auto lane_site = [&](uint32_t tz, uint32_t ty, uint32_t tx, uint32_t l) -> uint32_t {
    if constexpr (kDim == 2) {
      const uint32_t y = l >> 3, x = l & 7u;
      return (y < ext(ty) && x < ext(tx)) ? (E * ty + y) * L + E * tx + x : kNone;
    } else {
      const uint32_t z = l >> 4, y = (l >> 2) & 3u, x = l & 3u;
      return (z < ext(tz) && y < ext(ty) && x < ext(tx)) ? (E * tz + z) * L2 + (E * ty + y) * L + E * tx + x : kNone;
    }
  };
  auto coords = [&](uint32_t t, uint32_t &tz, uint32_t &ty, uint32_t &tx) {
    tx = t % T;
    ty = (t / T) % T;
    tz = kDim == 2 ? 0u : t / (T * T);
  };
  // words -> bitboards, a wave per tile, four tiles' words requested together (from HBM they take
  // a round trip each)
  auto words = [&](uint32_t t, uint32_t &q, uint32_t &cl) {
    q = cl = 0u;
    if (t >= nt) return;
    uint32_t tz, ty, tx;
    coords(t, tz, ty, tx);
    const uint32_t i = lane_site(tz, ty, tx, c.lane);
    if (i != kNone) {
      q = qw[i];
      cl = cw[i];
    }
  };
  for (uint32_t t4 = c.wave; t4 < nt; t4 += 4u * c.W) {
    uint32_t q4[4], c4w[4];
#pragma unroll
    for (uint32_t j = 0; j < 4u; ++j) words(t4 + j * c.W, q4[j], c4w[j]);
#pragma unroll
    for (uint32_t j = 0; j < 4u; ++j) {
      const uint32_t t = t4 + j * c.W;
      if (t >= nt) break;
      uint64_t myq = 0ull, myc = 0ull;
      for (uint32_t p = 0; p < P; ++p) {
        const uint64_t vq = ballot(((q4[j] >> p) & 1u) != 0u), vc = ballot(((c4w[j] >> p) & 1u) != 0u);
        if (c.lane == p) {
          myq = vq;
          myc = vc;
        }
      }
      if (c.lane < P) {
        bq[(size_t)t * P + c.lane] = myq;
        bcl[(size_t)t * P + c.lane] = myc;
      }
    }
  }
  if (c.tid < 3u) c.misc[8u + c.tid] = 0u;
  __syncthreads();
  // thread tid takes pair pp = tid % P and a run of m consecutive tiles (raster order: a strip of a
  // tile row), m = nt / S rounded up, S = nthreads / P threads per pair; a pass makes kLatRounds
  // rounds over the run between barriers, forwards and backwards in turn, so that a cluster crosses
  // the run, and what other threads wrote in earlier rounds, without a barrier
  constexpr uint32_t kLatRounds = 2u;
  const uint32_t S = c.nthreads / P, pp = c.tid % P, t0 = c.tid / P;
  const uint32_t m = (nt + S - 1u) / S, tb = t0 * m, te = (t0 < S && tb < nt) ? (tb + m < nt ? tb + m : nt) : tb;
  const uint32_t sx = P, sy = T * P, sz = T * T * P;
  for (uint32_t pass = 0;; ++pass) {
    if (c.tid == 0u) c.ticks[kPhIcmPasses] += 1u;
    uint32_t gain = 0u;
    for (uint32_t rep = 0; rep < kLatRounds; ++rep) {
      const bool back = ((pass * kLatRounds + rep) & 1u) != 0u, first = pass == 0u && rep == 0u;
      uint32_t tz, ty, tx;
      coords(back ? te - 1u : tb, tz, ty, tx);
      // this tile's bitboards of pair pp: index (tile in raster order) P + pp, one tile on per step
      uint32_t o = (back ? te - 1u : tb) * P + pp;
      for (uint32_t k = tb; k < te; ++k, o = back ? o - P : o + P) {
        const uint32_t cx = tx, cy = ty, cz = tz;
        if (back) {  // step to the previous tile in raster order
          if (tx) {
            --tx;
          } else {
            tx = T - 1u;
            if (kDim == 2 || ty) {
              --ty;
            } else {
              ty = T - 1u;
              --tz;
            }
          }
        } else if (++tx == T) {
          tx = 0u;
          if (++ty == T && kDim == 3) {
            ty = 0u;
            ++tz;
          }
        }
        const uint64_t q = bq[o], cl = bcl[o];
        if (q == cl) continue;  // every q<0 site of the tile joined (or there is none)
        // the adjacent tiles' bitboards (index steps sx, sy, sz along the axes, periodic) and their
        // extents; the halo: their last and first layers along each axis, moved to this tile's
        // first and last layers
        const uint32_t ol = cx ? o - sx : o + (T - 1u) * sx, or_ = cx + 1u < T ? o + sx : o - (T - 1u) * sx;
        const uint32_t ou = cy ? o - sy : o + (T - 1u) * sy, od = cy + 1u < T ? o + sy : o - (T - 1u) * sy;
        const uint32_t el = cx ? E : last, eu = cy ? E : last;  // the tile before along x / y
        uint64_t h;
        if constexpr (kDim == 2) {
          constexpr uint64_t kX0 = 0x0101010101010101ull;
          h = ((bcl[ol] >> (el - 1u)) & kX0) | ((bcl[or_] & kX0) << (ext(cx) - 1u)) |
              ((bcl[ou] >> (8u * (eu - 1u))) & 0xFFull) | ((bcl[od] & 0xFFull) << (8u * (ext(cy) - 1u)));
        } else {
          constexpr uint64_t kX0 = 0x1111111111111111ull, kY0 = 0x000F000F000F000Full;
          const uint32_t oz = cz ? o - sz : o + (T - 1u) * sz, ozd = cz + 1u < T ? o + sz : o - (T - 1u) * sz;
          const uint32_t ez = cz ? E : last;
          h = ((bcl[ol] >> (el - 1u)) & kX0) | ((bcl[or_] & kX0) << (ext(cx) - 1u)) |
              ((bcl[ou] >> (4u * (eu - 1u))) & kY0) | ((bcl[od] & kY0) << (4u * (ext(cy) - 1u))) |
              ((bcl[oz] >> (16u * (ez - 1u))) & 0xFFFFull) | ((bcl[ozd] & 0xFFFFull) << (16u * (ext(cz) - 1u)));
        }
        uint64_t x = cl | (q & h);
        // after the first round a tile's bits are closed under these joins (only its own thread
        // writes them, at the end of a fill), so without new halo bits nothing changes
        if (x == cl && !first) continue;
        for (uint32_t it = 0; it < 64u; ++it) {  // each round joins at least one more site
          uint64_t y = x, r;
          // along x both ways (a site at x = 0 does not take the row before it, one at the last x
          // not the row after it), then along y and z: each an occluded fill of shift steps; no
          // q bit lies past the lattice's edge, so nothing joins there
          if constexpr (kDim == 2) {
            constexpr uint64_t kX0 = 0x0101010101010101ull, kX7 = 0x8080808080808080ull;
            r = q & ~kX0;
            y |= r & (y << 1);
            r &= r << 1;
            y |= r & (y << 2);
            r &= r << 2;
            y |= r & (y << 4);
            r = q & ~kX7;
            y |= r & (y >> 1);
            r &= r >> 1;
            y |= r & (y >> 2);
            r &= r >> 2;
            y |= r & (y >> 4);
            r = q;
            y |= r & (y << 8);
            r &= r << 8;
            y |= r & (y << 16);
            r &= r << 16;
            y |= r & (y << 32);
            r = q;
            y |= r & (y >> 8);
            r &= r >> 8;
            y |= r & (y >> 16);
            r &= r >> 16;
            y |= r & (y >> 32);
          } else {
            constexpr uint64_t kX0 = 0x1111111111111111ull, kX3 = 0x8888888888888888ull;
            constexpr uint64_t kY0 = 0x000F000F000F000Full, kY3 = 0xF000F000F000F000ull;
            r = q & ~kX0;
            y |= r & (y << 1);
            r &= r << 1;
            y |= r & (y << 2);
            r = q & ~kX3;
            y |= r & (y >> 1);
            r &= r >> 1;
            y |= r & (y >> 2);
            r = q & ~kY0;
            y |= r & (y << 4);
            r &= r << 4;
            y |= r & (y << 8);
            r = q & ~kY3;
            y |= r & (y >> 4);
            r &= r >> 4;
            y |= r & (y >> 8);
            r = q;
            y |= r & (y << 16);
            r &= r << 16;
            y |= r & (y << 32);
            r = q;
            y |= r & (y >> 16);
            r &= r >> 16;
            y |= r & (y >> 32);
          }
          if (y == x) break;
          x = y;
        }
        if (x != cl) {
          bcl[o] = x;
          gain |= 1u << pp;
        }
      }
    }
    const uint32_t g = wave_or(gain);
    uint32_t *slot = c.misc + 8u + pass % 3u;
    if (c.lane == 0u && g) atomicOr(slot, g);
    __syncthreads();
    const uint32_t grew = *slot;
    if (c.tid == 0u) c.misc[8u + (pass + 2u) % 3u] = 0u;
    if (grew == 0u) break;
  }
  // bitboards -> cluster words
  for (uint32_t t = c.wave; t < nt; t += c.W) {
    uint32_t tz, ty, tx;
    coords(t, tz, ty, tx);
    const uint64_t myc = c.lane < P ? bcl[(size_t)t * P + c.lane] : 0ull;
    uint32_t w = 0u;
    for (uint32_t p = 0; p < P; ++p) {
      const uint64_t b = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(myc >> 32), (int)p) << 32) |
                         (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)myc, (int)p);
      w |= (uint32_t)((b >> c.lane) & 1ull) << p;
    }
    const uint32_t i = lane_site(tz, ty, tx, c.lane);
    if (i != kNone) cw[i] = w;
  }
  __syncthreads();
}

// ICM step (c): grow every pair's cluster through its q<0 sites.  A site joins pair p's cluster
// when its q bit p is set and a neighbour is in it.  Passes update in place until a pass changes
// nothing; each pass makes several rounds over the sites between barriers (four when the sites
// live in registers, two otherwise): updates made by other waves during one round are already
// visible in the next, so clusters grow by more than one hop per barrier.  The cluster fixpoint
// does not depend on the visiting order.
template <bool kLds>
__device__ __forceinline__ void icm_grow(Ctx &c, const IcmMem<kLds> &M) {
  const TAMC_AS4 PtGroup &X = *c.X;
  const uint32_t n = c.n, D = c.nbr ? X.ell_D : 0u;
  if ((c.flags & kFlagIcmTiles) && X.icm_tiles != nullptr && D != 0u && n <= 16u * c.nthreads) {
    const uint32_t S = n <= 4u * c.nthreads ? 4u : (n <= 8u * c.nthreads ? 8u : 16u);
    if (D <= 4u) {
      if (S == 4u) icm_grow_tiles<4, 1>(c, M, n, D);
      else if (S == 8u) icm_grow_tiles<8, 1>(c, M, n, D);
      else icm_grow_tiles<16, 1>(c, M, n, D);
    } else {
      if (S == 4u) icm_grow_tiles<4, 2>(c, M, n, D);
      else if (S == 8u) icm_grow_tiles<8, 2>(c, M, n, D);
      else icm_grow_tiles<16, 2>(c, M, n, D);
    }
  } else if (D != 0u && D <= 4u && n <= 4u * c.nthreads) {
    icm_grow_registers<4, 1>(c, M, n, D);
  } else if (D != 0u && D <= 4u && n <= 8u * c.nthreads) {
    icm_grow_registers<8, 1>(c, M, n, D);
  } else if (D != 0u && D <= 4u && n <= 16u * c.nthreads) {
    icm_grow_registers<16, 1>(c, M, n, D);
  } else if (D > 4u && D <= 8u && n <= 4u * c.nthreads) {
    icm_grow_registers<4, 2>(c, M, n, D);
  } else if (D > 4u && D <= 8u && n <= 8u * c.nthreads) {
    icm_grow_registers<8, 2>(c, M, n, D);
  } else if (D > 4u && D <= 8u && n <= 16u * c.nthreads) {
    icm_grow_registers<16, 2>(c, M, n, D);
  } else {
    // sites in memory, any number of them.  Pass k ORs the pairs that gained sites into
    // misc[8 + k % 3] (the slot of pass k + 2 is cleared after the barrier).  On graphs of at most
    // four couplings per site (two-dimensional lattices, where clusters near percolation are long
    // and take hundreds of passes, each a walk over every site in memory), the pairs that still
    // grow after kLabelAfterPasses passes, if few, are finished by component labelling, which
    // needs the 16-bit neighbour lists (and 16-bit site labels).  Wider rows (three-dimensional
    // lattices) finish sooner by growth alone.
    if (c.tid < 3u) c.misc[8u + c.tid] = 0u;
    __syncthreads();
    uint32_t label_pairs = 0u;
    const bool can_label = D != 0u && D <= 4u && n < kNoLabel;
    for (uint32_t rnd = 0;; ++rnd) {
      if (c.tid == 0u) c.ticks[kPhIcmPasses] += 1u;
      uint32_t gain = 0u;
      for (uint32_t rep = 0; rep < 2u; ++rep)
      for (uint32_t i = c.tid; i < n; i += c.nthreads) {
        const uint32_t q = M.qw[i];
        const uint32_t cur = M.cw[i];
        if (cur == q) continue;  // no q<0 bit left to join (also q == 0)
        uint32_t nb = 0;
        if (D != 0u && D <= 4u) {
          // one 8-byte load of the (padded) row
          const uint64_t w = M.nbr_lds ? reinterpret_cast<const TAMC_AS3 uint64_t *>(M.nbr_lds)[i]
                                       : reinterpret_cast<const TAMC_AS4 uint64_t *>(M.nbr_glb)[i];
          nb = M.cw[w & 0xFFFFu] | M.cw[(w >> 16) & 0xFFFFu] | M.cw[(w >> 32) & 0xFFFFu] | M.cw[w >> 48];
        } else if (D != 0u) {
          uint64_t w0, w1;
          if (M.nbr_lds) {
            w0 = reinterpret_cast<const TAMC_AS3 uint64_t *>(M.nbr_lds)[2u * i];
            w1 = reinterpret_cast<const TAMC_AS3 uint64_t *>(M.nbr_lds)[2u * i + 1u];
          } else {
            w0 = reinterpret_cast<const TAMC_AS4 uint64_t *>(M.nbr_glb)[2u * i];
            w1 = reinterpret_cast<const TAMC_AS4 uint64_t *>(M.nbr_glb)[2u * i + 1u];
          }
          nb = M.cw[w0 & 0xFFFFu] | M.cw[(w0 >> 16) & 0xFFFFu] | M.cw[(w0 >> 32) & 0xFFFFu] | M.cw[w0 >> 48] |
               M.cw[w1 & 0xFFFFu] | M.cw[(w1 >> 16) & 0xFFFFu] | M.cw[(w1 >> 32) & 0xFFFFu] | M.cw[w1 >> 48];
        } else if (i < X.g_nodes) {
          const uint32_t e0 = c4(X.g_row_ptr)[i], e1 = c4(X.g_row_ptr)[i + 1];
          for (uint32_t e = e0; e < e1; ++e) nb |= M.cw[c4(X.g_col)[e]];
        }
        const uint32_t nc = cur | (q & nb);
        if (nc != cur) {
          M.cw[i] = nc;
          gain |= nc ^ cur;
        }
      }
      const uint32_t g = wave_or(gain);
      uint32_t *slot = c.misc + 8u + rnd % 3u;
      if (c.lane == 0u && g) atomicOr(slot, g);
      __syncthreads();
      const uint32_t grew = *slot;
      if (c.tid == 0u) c.misc[8u + (rnd + 2u) % 3u] = 0u;
      if (grew == 0u) break;
      if (can_label && rnd + 1u >= kLabelAfterPasses && (uint32_t)__popc(grew) <= kMaxLabelPairs) {
        label_pairs = grew;
        break;
      }
    }
    if (label_pairs) {
      const uint32_t spt = (n + c.nthreads - 1u) / c.nthreads;
      icm_label_pairs<16, 1>(c, M, n, label_pairs, (spt + 7u) & ~7u);
    }
  }
}

// 1. Houdayer ICM for all pairs of the group at once (src/pt.rs:31-77).  Pair p couples slot lo+p of
// the two chains; bit p of the per-site words qw/cw belongs to pair p.
template <bool kLds, bool kSpinsGlobal = false>
__device__ __forceinline__ void icm_phase(Ctx &c, uint32_t t) {
  const TAMC_AS4 PtGroup &X = *c.X;
  using U32 = typename IcmMem<kLds>::U32;
  IcmMem<kLds> M;
  M.qw = (U32 *)c.qw;
  M.cw = (U32 *)c.cw;
  M.blk = (U32 *)c.blk;
  M.nbr_lds = c.nbr_in_lds ? (const TAMC_AS3 uint16_t *)c.nbr : nullptr;
  M.nbr_glb = c4(X.ell_col16);
  const uint32_t P = c.P, n = c.n, G = c.G, D = c.nbr ? X.ell_D : 0u;
  if (c.tid < P) {
    const uint32_t b = c.lo + c.tid;
    c.pair_a[c.tid] = c.perm[b];
    c.pair_b[c.tid] = c.perm[c.B + b];
    c.dEa[c.tid] = 0;
  }
  __syncthreads();
  // flip masks: entry (g, k, v) holds the replicas in plane word g of the pairs 4k + j with bit j
  // of v set
  for (uint32_t idx = c.tid; idx < G * (kMaxPairs / 4u) * 16u; idx += c.nthreads) {
    const uint32_t g = idx / ((kMaxPairs / 4u) * 16u), k = (idx >> 4) % (kMaxPairs / 4u), v = idx & 15u;
    if (4u * k >= P) continue;
    unsigned long long m = 0ull;
    for (uint32_t j = 0; j < 4u; ++j) {
      const uint32_t p = 4u * k + j;
      if (p >= P || !((v >> j) & 1u)) continue;
      const uint32_t a = c.pair_a[p], b = c.pair_b[p];
      if ((a >> 6) == g) m |= 1ull << (a & 63u);
      if ((b >> 6) == g) m |= 1ull << (b & 63u);
    }
    c.fmask[idx] = m;  // read only after the growth loop's barriers
  }
  // (a) overlap words: bit p of qw[i] <=> the replicas of pair p disagree at site i; blk[p][b] =
  //     number of such sites in 64-site block b.  Bit p of q is bit a_p ^ bit b_p of the site's
  //     planes.  With one plane word per site the word is held in a register and shifted per pair.
  //     With two to kMaxTableWords words, q is the XOR, over the bytes of the plane words, of a
  //     table entry per byte: tab[g][j][v] = XOR of pm(r) over the bits k of v, r = 64 g + 8 j + k,
  //     where pm(r) = 1 << p when r is replica a_p or b_p of a pair of the chunk (else 0); the
  //     tables (8 KB per plane word) live in the partial-sum area, which only the sweep and the
  //     measurements use.  With more words each pair's two bits are read from the planes.
  const bool tables = G > 1u && G <= kMaxTableWords;
  uint32_t *tab = reinterpret_cast<uint32_t *>(c.partial);
  if (tables) {
    uint32_t *pm = c.pair_of;
    for (uint32_t r = c.tid; r < c.R; r += c.nthreads) pm[r] = 0u;
    __syncthreads();
    if (c.tid < P) {
      pm[c.pair_a[c.tid]] = 1u << c.tid;
      pm[c.pair_b[c.tid]] = 1u << c.tid;
    }
    __syncthreads();
    for (uint32_t idx = c.tid; idx < G * 2048u; idx += c.nthreads) {
      const uint32_t r0 = idx >> 5 & ~7u, v = idx & 255u;  // r0 = 64 g + 8 j
      uint32_t t = 0u;
#pragma unroll
      for (uint32_t k = 0; k < 8u; ++k)
        if (((v >> k) & 1u) && r0 + k < c.R) t ^= pm[r0 + k];
      tab[idx] = t;
    }
    __syncthreads();
  }
  for (uint32_t blk = c.wave; blk < c.nblk; blk += c.W) {
    const uint32_t i = blk * 64u + c.lane;
    const bool valid = i < n;
    uint32_t q = 0;
    if (valid && G == 1u) {
      // the pairs' replicas four at a time (one 16-byte LDS read each for a and b; entries past P
      // are in bounds and unused), so a block waits for P / 4 LDS round trips instead of P
      const uint64_t w = *reinterpret_cast<const uint64_t *>(c.sp + 2u * i);
      typedef uint32_t U4v __attribute__((ext_vector_type(4)));
      for (uint32_t k = 0; 4u * k < P; ++k) {
        const U4v av = reinterpret_cast<const U4v *>(c.pair_a)[k], bv = reinterpret_cast<const U4v *>(c.pair_b)[k];
#pragma unroll
        for (uint32_t j = 0; j < 4u; ++j)
          if (4u * k + j < P) q |= (uint32_t)(((w >> av[j]) ^ (w >> bv[j])) & 1ull) << (4u * k + j);
      }
    } else if (valid && tables) {
      for (uint32_t g = 0; g < G; ++g) {
        const uint32_t lo = c.sp[(i * G + g) * 2u], hi = c.sp[(i * G + g) * 2u + 1u];
        const uint32_t *tg = tab + g * 2048u;
        q ^= tg[lo & 255u] ^ tg[256u + ((lo >> 8) & 255u)] ^ tg[512u + ((lo >> 16) & 255u)] ^ tg[768u + (lo >> 24)];
        q ^= tg[1024u + (hi & 255u)] ^ tg[1280u + ((hi >> 8) & 255u)] ^ tg[1536u + ((hi >> 16) & 255u)] ^
             tg[1792u + (hi >> 24)];
      }
    } else if (valid) {
      for (uint32_t p = 0; p < P; ++p) q |= (sbit(c.sp, G, i, c.pair_a[p]) ^ sbit(c.sp, G, i, c.pair_b[p])) << p;
    }
    if (valid) {
      M.qw[i] = q;
      M.cw[i] = 0u;
    }
    // per-pair counts of the block, four pairs at a time: bits 4k..4k+3 of q spread into the low
    // bits of four bytes (the multiply places copies of the nibble at bits 0, 7, 14 and 21, which
    // cannot carry into each other), summed over the wave (each byte <= 64); lane p keeps byte
    // p % 4 of word p / 4
    uint32_t cnt = 0;
    for (uint32_t k = 0; 4u * k < P; ++k) {
      const uint32_t v = __umul24((q >> (4u * k)) & 0xFu, 0x00204081u) & 0x01010101u;
      const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan(v), 63);
      if ((c.lane >> 2) == k) cnt = (tot >> (8u * (c.lane & 3u))) & 0xFFu;
    }
    if (c.lane < P) M.blk[c.lane * c.nblk + blk] = cnt;
  }
  __syncthreads();
  lap(c, kPhIcmOverlap);
  // (b) seed choice, one wave per pair: k = floor(w * count / 2^32) into the ascending list of the
  //     pair's q<0 sites.  Lane j holds block j's count (64 blocks at a time); a wave scan finds the
  //     block holding the k-th site and a ballot over that block the site, which joins cw.
  for (uint32_t p = c.wave; p < P; p += c.W) {
    uint32_t total = 0;
    for (uint32_t c0 = 0; c0 < c.nblk; c0 += 64u) {
      const uint32_t j = c0 + c.lane;
      const uint32_t v = (j < c.nblk) ? M.blk[p * c.nblk + j] : 0u;
      total += (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan(v), 63);
    }
    uint32_t sb = kNone;
    if (total) {
      uint32_t k;
      if (X.icm_k) {
        k = c4(X.icm_k)[c.lo - X.lo + p];
        if (k >= total) k = total - 1u;
      } else {
        const U4 v = philox4x32_10(t, c.lo + p, X.pair_global, kTagIcm, c.key0, c.key1);
        k = (uint32_t)(((uint64_t)v.x * total) >> 32);
      }
      uint32_t cum = 0;
      for (uint32_t c0 = 0; c0 < c.nblk; c0 += 64u) {
        const uint32_t j = c0 + c.lane;
        const uint32_t v = (j < c.nblk) ? M.blk[p * c.nblk + j] : 0u;
        const uint32_t incl = wave_incl_scan(v);
        const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        if (k < cum + tot) {
          const uint32_t r = k - cum;
          const uint32_t jj = (uint32_t)__builtin_ctzll(ballot(incl > r));
          const uint32_t rank = r - ((uint32_t)__builtin_amdgcn_readlane((int)incl, (int)jj) -
                                     (uint32_t)__builtin_amdgcn_readlane((int)v, (int)jj));
          sb = c0 + jj;
          const uint32_t i = sb * 64u + c.lane;
          const bool qb = i < n && ((M.qw[i] >> p) & 1u);
          const uint64_t m = ballot(qb);
          if (qb && mbcnt64(m) == rank) atomicOr((uint32_t *)&M.cw[i], 1u << p);
          break;
        }
        cum += tot;
      }
      if (c.lane == 0) atomicAdd(c.moves, 1u);
    }
    if (c.lane == 0) c.seed_blk[p] = sb;
  }
  lap(c, kPhIcmSeed);
  __syncthreads();
  // square and cubic lattices: bitboards in the LDS of the parked spin planes (which are not read
  // while clusters grow), when they fit
  const bool lattice = !kSpinsGlobal && X.lat_L != 0u &&
                       4u * (X.lat_dim == 3u ? lattice_tiles<3>(X.lat_L) : lattice_tiles<2>(X.lat_L)) * P <= n * G * 2u;
  if (lattice) {
    const uint32_t nw = n * G * 2u;
    uint32_t *park = X.g_park ? X.g_park : X.spins;
    for (uint32_t k = c.tid; k < nw; k += c.nthreads) park[k] = c.sp[k];
    __syncthreads();
    TAMC_AS3 uint64_t *bb = reinterpret_cast<TAMC_AS3 uint64_t *>((TAMC_AS3 uint32_t *)c.sp);
    if (X.lat_dim == 3u) icm_grow_lattice<3>(c, M.qw, M.cw, X.lat_L, bb);
    else icm_grow_lattice<2>(c, M.qw, M.cw, X.lat_L, bb);
    for (uint32_t k = c.tid; k < nw; k += c.nthreads) c.sp[k] = park[k];
    __syncthreads();
  } else if constexpr (kLds || kSpinsGlobal) {
    // (with the spin planes in global memory there is no LDS to borrow: the clusters grow in the
    // global scratch)
    icm_grow(c, M);
  } else {
    // the spins are not read while clusters grow: park the planes in the group's spin array and
    // grow in the LDS they leave (q and cluster words take 8 bytes per site, the planes 8 G)
    const uint32_t nw = n * G * 2u;
    uint32_t *park = X.g_park ? X.g_park : X.spins;
    for (uint32_t k = c.tid; k < nw; k += c.nthreads) park[k] = c.sp[k];
    __syncthreads();
    IcmMem<true> S;
    S.qw = (TAMC_AS3 uint32_t *)c.sp;
    S.cw = (TAMC_AS3 uint32_t *)(c.sp + n);
    S.blk = nullptr;
    S.nbr_lds = M.nbr_lds;
    S.nbr_glb = M.nbr_glb;
    for (uint32_t i = c.tid; i < n; i += c.nthreads) {
      S.qw[i] = M.qw[i];
      S.cw[i] = M.cw[i];
    }
    __syncthreads();
    icm_grow(c, S);
    __syncthreads();
    for (uint32_t i = c.tid; i < n; i += c.nthreads) M.cw[i] = S.cw[i];
    __syncthreads();
    for (uint32_t k = c.tid; k < nw; k += c.nthreads) c.sp[k] = park[k];
    __syncthreads();
  }
  lap(c, kPhIcmGrow);
  // (d) swap the cluster spins of both replicas (= flip both) and sum the energy change
  //     (the fixed-width terms read the 16-bit neighbour lists, which instances of 65,536 sites or
  //     more do not have: those walk the coupling rows)
  switch (c.nbr ? X.ell_D : 0u) {
    case 3u: icm_flip<3>(c, M); break;
    case 4u: icm_flip<4>(c, M); break;
    case 6u: icm_flip<6>(c, M); break;
    case 8u: icm_flip<8>(c, M); break;
    default: icm_flip<0>(c, M);
  }
  if (c.members > 1u) icm_team_merge(c);
  __syncthreads();
  // moved pairs: energies become energy_ref of the new states, as the reference recomputes them
  if (c.tid < P && c.seed_blk[c.tid] != kNone) {
    const uint32_t a = c.pair_a[c.tid], b = c.pair_b[c.tid];
    const long long d = c.dEa[c.tid];
    c.energy[a] = add_fix(c.energy[a], d - c.drift[a]);
    c.energy[b] = add_fix(c.energy[b], -d - c.drift[b]);
    c.drift[a] = 0;
    c.drift[b] = 0;
  }
  __syncthreads();
}

// 2. replica exchange (tamc-core/src/pt.rs:126-144; threaded variant tamc-core/src/parallel/pt.rs:202-215)
//    and round trips (tamc-core/src/pt.rs:54-84).  One wave per chain, lane j = ladder step j
//    (slots j, j+1).  The threaded variant decides every step independently; the serial one
//    carries dE into the next step after a swap, so its decisions are a scan over the steps,
//    done on values read from the lanes' registers.  The walk's permutation then follows from the
//    decisions in parallel: when step i rejects, the walker restarts at slot i+1, so slot j
//    receives slot j+1 on acceptance and otherwise the slot where the walk last restarted.
__device__ __forceinline__ void exchange_serial_walk(Ctx &c, uint32_t t, uint32_t ch) {
  const TAMC_AS4 PtGroup &X = *c.X;
  const uint32_t B = c.B, B1 = B - 1u;
  uint32_t *perm = c.perm + ch * B;
  uint8_t *tags = c.tags + ch * B;
  const bool threaded = (c.flags & kFlagThreaded) != 0u;
  uint32_t cur = perm[0];
  uint8_t curtag = tags[0];
  bool prev_ok = false;
  float prev = 0.0f;
  for (uint32_t j = 0; j < B1; ++j) {
    const float dEj = c.energy[perm[j + 1]] - c.energy[perm[j]];
    const U4 v = philox4x32_10(t, j >> 2, X.chain0 + ch, kTagExchange, c.key0, c.key1);
    // serial variant: after a swap at j-1, dE[j] += dE[j-1] (the value used at j-1)
    const float eff = (!threaded && prev_ok) ? dEj + prev : dEj;
    const float dlt = (c4(X.betas)[j + 1] - c4(X.betas)[j]) * eff;
    bool ok = dlt >= 0.0f;
    if (!ok) ok = accept_u(dlt, pick(v, j));
    const uint32_t nxt = perm[j + 1];
    const uint8_t nxttag = tags[j + 1];
    if (ok) {
      perm[j] = nxt;
      tags[j] = nxttag;
      c.nacc[ch * B + j] += 1u;
    } else {
      perm[j] = cur;
      tags[j] = curtag;
      cur = nxt;
      curtag = nxttag;
    }
    prev_ok = ok;
    prev = eff;
  }
  perm[B1] = cur;
  tags[B1] = curtag;
}

__device__ __forceinline__ void round_trips(Ctx &c, uint32_t ch) {
  // update_round_trips; tags: 0 None 1 Min 2 Max 3 MinRefl 4 MaxRefl
  const uint32_t B = c.B, B1 = B - 1u;
  uint8_t *tags = c.tags + ch * B;
  uint8_t lo = tags[0];
  if (lo == 0) lo = 1;
  else if (lo == 4) {
    c.rt[ch] += 1u;
    lo = 1;
  } else if (lo == 2) lo = 3;
  tags[0] = lo;
  uint8_t hi = tags[B1];
  if (hi == 0) hi = 2;
  else if (hi == 3) {
    c.rt[ch] += 1u;
    hi = 2;
  } else if (hi == 1) hi = 4;
  tags[B1] = hi;
}

__device__ __forceinline__ void exchange_phase(Ctx &c, uint32_t t) {
  const TAMC_AS4 PtGroup &X = *c.X;
  const uint32_t B = c.B, B1 = B - 1u;
  const bool threaded = (c.flags & kFlagThreaded) != 0u;
  for (uint32_t ch = c.wave; ch < c.nch; ch += c.W) {
    if (B > 64u) {  // ladders longer than a wave: one lane walks them
      if (c.lane == 0) {
        exchange_serial_walk(c, t, ch);
        round_trips(c, ch);
      }
      continue;
    }
    uint32_t *perm = c.perm + ch * B;
    uint8_t *tags = c.tags + ch * B;
    const uint32_t j = c.lane;
    const bool step = j < B1;
    float dE = 0.f, dbeta = 0.f;
    uint32_t w = 0u;
    if (step) {
      dE = c.energy[perm[j + 1]] - c.energy[perm[j]];
      dbeta = c4(X.betas)[j + 1] - c4(X.betas)[j];
      const U4 v = philox4x32_10(t, j >> 2, X.chain0 + ch, kTagExchange, c.key0, c.key1);
      w = pick(v, j);
    }
    uint64_t okmask;
    if (threaded) {
      const float dlt = dbeta * dE;
      const bool ok = step && (dlt >= 0.0f || accept_u(dlt, w));
      okmask = ballot(ok);
    } else {
      // every step decided at once as if the step before it rejected (its dE unchanged); the walk
      // then decides again only the steps that follow an accepted swap
      const float dlt0 = dbeta * dE;
      const uint64_t spec = ballot(step && (dlt0 >= 0.0f || accept_u(dlt0, w)));
      okmask = 0ull;
      bool prev_ok = false;
      float prev = 0.0f;
      for (uint32_t k = 0; k < B1; ++k) {
        const float dEk = __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(dE), k));
        float eff = dEk;
        bool ok = ((spec >> k) & 1ull) != 0ull;
        if (prev_ok) {
          // after a swap at k-1, dE[k] += dE[k-1] (the value used at k-1)
          const float dbk = __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(dbeta), k));
          const uint32_t wk = __builtin_amdgcn_readlane(w, k);
          eff = dEk + prev;
          const float dlt = dbk * eff;
          ok = dlt >= 0.0f || accept_u(dlt, wk);
        }
        if (ok) okmask |= 1ull << k;
        prev_ok = ok;
        prev = eff;
      }
    }
    // slot j after the walk: j+1 if step j accepted, else the last restart point r(j)
    const uint64_t rejected = ~okmask & ((B1 >= 64u) ? ~0ull : ((1ull << B1) - 1ull));
    const uint64_t below = rejected & ((j >= 64u) ? ~0ull : ((1ull << j) - 1ull));
    const uint32_t r = below ? 64u - (uint32_t)__builtin_clzll(below) : 0u;
    const bool okj = step && ((okmask >> j) & 1ull);
    uint32_t np = 0u;
    uint8_t nt = 0;
    if (j < B) {
      const uint32_t src = okj ? j + 1u : r;
      np = perm[src];
      nt = tags[src];
    }
    if (j < B) {  // every read above is done before any write (one wave, LDS in order)
      perm[j] = np;
      tags[j] = nt;
      if (okj) c.nacc[ch * B + j] += 1u;
    }
    if (c.lane == 0) round_trips(c, ch);
  }
  __syncthreads();
  for (uint32_t k = c.tid; k < c.nch * B; k += c.nthreads) {
    const uint8_t tg = c.tags[k];
    if (tg == 1 || tg == 3) c.hist[2u * k] += 1u;
    else if (tg == 2 || tg == 4) c.hist[2u * k + 1u] += 1u;
    c.beta_of[c.perm[k]] = c4(X.betas)[k % B];
  }
  __syncthreads();
}

// local field of site i for the lane's replica (word g), summed in coupling_vecs order
template <int D>
__device__ __forceinline__ float site_field(const Ctx &c, uint32_t i, uint32_t g, uint32_t half, uint32_t sh) {
  const TAMC_AS4 PtGroup &X = *c.X;
  const uint32_t G = c.G;
  float d = 0.0f;
  d += c4(X.bias)[i];
  if constexpr (D > 0) {
    const TAMC_AS4 uint32_t *cp = c4(X.ell_col) + (size_t)i * D;
    const TAMC_AS4 float *jp = c4(X.ell_J) + (size_t)i * D;
    uint32_t w[D];
    float Jv[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      Jv[k] = jp[k];
      w[k] = c.sp[(cp[k] * G + g) * 2u + half];
    }
    // d += s_j J_j as an fma with s_j = +-1.0 built from the lane's spin bit: J_j s_j is exact, so
    // the fma rounds exactly like the reference's separate add (zeros' signs included)
    const uint32_t up = 31u - sh;
#pragma unroll
    for (int k = 0; k < D; ++k)
      d = __fmaf_rn(Jv[k], __uint_as_float(0x3f800000u | ((w[k] << up) & 0x80000000u)), d);
  } else {
    const uint32_t e0 = c4(X.row_ptr)[i], e1 = c4(X.row_ptr)[i + 1];
    for (uint32_t e = e0; e < e1; ++e) {
      const uint32_t j = c4(X.col)[e];
      const float Je = c4(X.J)[e];
      const uint32_t wj = c.sp[(j * G + g) * 2u + half];
      d += ((wj >> sh) & 1u) ? -Je : Je;
    }
  }
  return d;
}

// 3. Metropolis sweep, level by level.  Tasks are numbered level-major, word-minor: task k is word
// k % G of RNG group k / G.  Each task decides its four sites in two passes: a straight-line pass
// with the hardware exp2 (accept when u is below the estimate by more than the error band, reject
// when above), and, only when some lane of the wave fell inside the band, the exact test.

template <int D, bool kTeam>
__device__ __forceinline__ void sweep_phase_d(Ctx &c, uint32_t t) {
  const TAMC_AS4 PtGroup &X = *c.X;
  const uint32_t G = c.G;
  // accepted dE in fixed point, per (wave, plane word, lane) slot of the LDS partials: only this
  // lane adds into its slot, and no accumulator array is carried in registers through the loop
  unsigned long long *acc = reinterpret_cast<unsigned long long *>(c.partial) + c.wave * G * 64u + c.lane;
  for (uint32_t g = 0; g < G; ++g) acc[g * 64u] = 0ull;
  const uint32_t half = c.lane >> 5, sh = c.lane & 31u;
  const uint32_t nl = X.n_levels;
  const bool self = X.self_J != nullptr;
  uint32_t a = c4(X.lvl_off)[0];
  const bool blocks = (c.flags & kFlagTileBlocks) != 0u;
  const uint32_t K = kTeam ? c.members : 1u;
  // team meetings: one per level, numbered on from the member's count
  uint32_t e = kTeam ? c.misc[15] : 0u;
  uint32_t *abort_flag = kTeam ? X.team_ctr + 1 : nullptr;
  for (uint32_t l = 0; l < nl; ++l) {
    const uint32_t b = c4(X.lvl_off)[l + 1];
    if (K > 1u) ++e;
    // tiling of the level's tasks [a, b) over the waves: dealt in turn, or one contiguous run each.
    // In a team the member takes the RNG groups g == member (mod members) of the level, dealt in
    // turn: index j stands for word j % G of the level's (j / G)-th such group.
    const uint32_t ga = a / G, gb = b / G;
    const uint32_t gm = ga + (c.member + K - ga % K) % K;
    uint32_t k = a + c.wave, kend = b, kstep = c.W;
    if (K > 1u) {
      k = c.wave;
      kend = gm < gb ? (gb - gm + K - 1u) / K * G : 0u;
    } else if (blocks) {
      const uint32_t per = (b - a + c.W - 1u) / c.W;
      k = a + c.wave * per;
      kend = (k + per < b) ? k + per : b;
      kstep = 1u;
    }
    for (; k < kend; k += kstep) {
      const uint32_t kk = (K > 1u) ? (gm + (k / G) * K) * G + k % G : k;
      const uint32_t grp = (G == 1u) ? kk : kk / G, g = kk - grp * G;
      const uint32_t r = g * 64u + c.lane;
      const bool valid = r < c.R;
      const float beta = valid ? c.beta_of[r] : 0.0f;
      const uint32_t rho = X.chain0 * c.B + r;
      // re-materialise the key each task: otherwise the ten round keys are hoisted out of the
      // loop as twenty live scalars, which the register allocator spills to VGPR lanes
      uint32_t k0 = c.key0, k1 = c.key1;
      asm volatile("" : "+s"(k0), "+s"(k1));
      const U4 u = philox4x32_10(t, grp, rho, kTagSweep, k0, k1);
      uint32_t site[4];
      // the four sites of a group are mutually uncoupled: gather all fields first (independent
      // LDS reads in flight together), then decide and write the new spin words
      float d[4];
      uint32_t bi[4];
      if constexpr (D > 0) {
        // the group's record: sites, biases, cols, couplings, in one contiguous block
        const TAMC_AS4 uint32_t *rec = c4(X.records) + (size_t)grp * (8u + 8u * D);
        const uint32_t up = 31u - sh;
#pragma unroll
        for (uint32_t s = 0; s < 4u; ++s) site[s] = rec[s];
#pragma unroll
        for (uint32_t s = 0; s < 4u; ++s) {
          // padded sites have record zeros: a harmless field of site 0
          uint32_t w[D];
#pragma unroll
          for (int k = 0; k < D; ++k) w[k] = c.sp[(rec[8u + s * D + k] * G + g) * 2u + half];
          // the reference starts from 0.0 + bias; starting from the bias differs only in the sign
          // of a zero field, which changes no decision (dE = +-0 always accepts) and no energy
          float f = __uint_as_float(rec[4u + s]);
#pragma unroll
          for (int k = 0; k < D; ++k)
            f = __fmaf_rn(__uint_as_float(rec[8u + 4u * D + s * D + k]),
                          __uint_as_float(0x3f800000u | ((w[k] << up) & 0x80000000u)), f);
          d[s] = f;
          // padded sites read site 0's word: their decision is discarded
          const uint32_t si = (site[s] != kNone) ? site[s] : 0u;
          bi[s] = (c.sp[(si * G + g) * 2u + half] >> sh) & 1u;
        }
      } else {
#pragma unroll
        for (uint32_t s = 0; s < 4u; ++s) site[s] = c4(X.grp)[grp * 4u + s];
#pragma unroll
        for (uint32_t s = 0; s < 4u; ++s) {
          d[s] = 0.0f;
          bi[s] = 0u;
          if (site[s] != kNone) {
            d[s] = site_field<D>(c, site[s], g, half, sh);
            bi[s] = (c.sp[(site[s] * G + g) * 2u + half] >> sh) & 1u;
          }
        }
      }
      // decisions are kept as wave masks (one compare per ballot, combined on the scalar unit);
      // inverse_ballot turns a mask back into the lane's predicate at no vector cost
      const uint64_t vmask = ballot(valid);
      float dd[4], xs[4];
      uint32_t ks[4];
      uint64_t okm[4], band[4], any_band = 0ull;
#pragma unroll
      for (uint32_t s = 0; s < 4u; ++s) {
        dd[s] = d[s] * (bi[s] ? 2.0f : -2.0f);  // dE *= -2*s_i
        xs[s] = -beta * dd[s];
        ks[s] = pick(u, s) >> 8;
        // p = 2^-24 P; the thresholds scale by the exact constants 2^24 (1 -+ band), so the
        // products round exactly as accept_below(P) and reject_above(P) do
        const float p = __builtin_amdgcn_exp2f(xs[s] * 1.44269504088896341f);
        const float kf = (float)ks[s];
        const uint64_t acc_m = ballot(dd[s] < 0.0f) | ballot(kf < __fmaf_rn(p, 16777216.0f * (1.0f - kBand), -0x1p-100f));
        const uint64_t rej_m = ballot(kf > p * (16777216.0f * (1.0f + kBand)));
        okm[s] = acc_m;
        band[s] = (site[s] != kNone) ? ~(acc_m | rej_m) & vmask : 0ull;
        any_band |= band[s];
      }
      if (any_band != 0ull) {
#pragma unroll
        for (uint32_t s = 0; s < 4u; ++s) {
          if (band[s] == 0ull) continue;
          bool e = false;
          if (__builtin_amdgcn_inverse_ballot_w64(band[s])) e = accept_exact(xs[s], ks[s], exc_table());
          okm[s] |= ballot(e);
        }
      }
      float a_g = 0.f;
#pragma unroll
      for (uint32_t s = 0; s < 4u; ++s) {
        if (site[s] == kNone) break;
        const uint32_t i = site[s];
        const uint64_t accm = okm[s] & vmask;
        // the site's new plane word: the lanes' old bits with the accepted ones flipped
        const uint64_t nw = ballot(bi[s] != 0u) ^ accm;
        if (c.lane == 0u) *reinterpret_cast<uint64_t *>(c.sp + (i * G + g) * 2u) = nw;
        if (K > 1u && c.lane == 0u) tag_put(X.team_box + 2u * ((size_t)i * G + g), nw, e);
        const bool acc_ok = __builtin_amdgcn_inverse_ballot_w64(accm);
        if (acc_ok) a_g += dd[s];
        if (self) {
          const float sj = c4(X.self_J)[i];
          if (acc_ok && sj != 0.0f) atomicAdd((unsigned long long *)&c.drift[r], (unsigned long long)to_fix(-2.0f * sj));
        }
      }
      // the group's accepted dE, summed in group order, enter the fixed-point total
      atomicAdd(acc + g * 64u, (unsigned long long)to_fix(a_g));
    }
    if (K > 1u) {
      // after the last level, this member's share of the accepted energy goes out with its spin
      // words, so that one meeting serves both
      unsigned long long *sums = X.team_box + team_sums_slot(c.n, G);
      if (l + 1u == nl) {
        __syncthreads();
        for (uint32_t r = c.tid; r < c.R; r += c.nthreads) {
          const uint32_t g = r >> 6, ln = r & 63u;
          long long s = 0;
          for (uint32_t w = 0; w < c.W; ++w) s += c.partial[(w * G + g) * 64u + ln];
          tag_put(sums + 2u * ((size_t)c.member * c.R + r), (unsigned long long)s, e);
        }
      }
      // take the other members' new spin words of this level, four per thread at a time
      const uint32_t items = (gb - ga) * 4u * G;
      for (uint32_t it0 = c.tid; it0 < items; it0 += 4u * c.nthreads) {
        const unsigned long long *src[4];
        uint32_t dst[4], want = 0;
#pragma unroll
        for (uint32_t j = 0; j < 4u; ++j) {
          const uint32_t it = it0 + j * c.nthreads;
          src[j] = X.team_box;
          dst[j] = 0;
          if (it >= items) continue;
          const uint32_t q = ga + it / (4u * G);
          if (q % K == c.member) continue;
          const uint32_t site = c4(X.grp)[q * 4u + (it / G) % 4u], g = it % G;
          if (site == kNone) continue;
          src[j] = X.team_box + 2u * ((size_t)site * G + g);
          dst[j] = (site * G + g) * 2u;
          want |= 1u << j;
        }
        uint64_t v[4];
        tag_get4(src, want, e, abort_flag, v);
#pragma unroll
        for (uint32_t j = 0; j < 4u; ++j)
          if ((want >> j) & 1u) *reinterpret_cast<uint64_t *>(c.sp + dst[j]) = v[j];
      }
    }
    a = b;
    __syncthreads();
  }
  __syncthreads();
  if (K > 1u) {
    // every member adds every member's share, in member order: the same energies everywhere
    const unsigned long long *sums = X.team_box + team_sums_slot(c.n, G);
    for (uint32_t r = c.tid; r < c.R; r += c.nthreads) {
      const uint32_t g = r >> 6, ln = r & 63u;
      long long mine = 0;
      for (uint32_t w = 0; w < c.W; ++w) mine += c.partial[(w * G + g) * 64u + ln];
      c.energy[r] = add_fix(c.energy[r], tag_sum(sums + 2u * r, c.R, K, c.member, mine, e, abort_flag));
    }
    __syncthreads();
    if (c.tid == 0u) c.misc[15] = e;
    return;
  }
  for (uint32_t r = c.tid; r < c.R; r += c.nthreads) {
    const uint32_t g = r >> 6, l = r & 63u;
    long long s = 0;
    for (uint32_t w = 0; w < c.W; ++w) s += c.partial[(w * G + g) * 64u + l];
    c.energy[r] = add_fix(c.energy[r], s);
  }
  __syncthreads();
}

// the team variant is a separate instantiation, so that groups run by one work group keep the
// registers the team bookkeeping would take
__device__ __forceinline__ void sweep_phase(Ctx &c, uint32_t t) {
  if (c.members > 1u) {
    switch (c.X->ell_D) {
      case 3u: sweep_phase_d<3, true>(c, t); break;
      case 4u: sweep_phase_d<4, true>(c, t); break;
      case 6u: sweep_phase_d<6, true>(c, t); break;
      case 8u: sweep_phase_d<8, true>(c, t); break;
      default: sweep_phase_d<0, true>(c, t);
    }
    return;
  }
  switch (c.X->ell_D) {
    case 3u: sweep_phase_d<3, false>(c, t); break;
    case 4u: sweep_phase_d<4, false>(c, t); break;
    case 6u: sweep_phase_d<6, false>(c, t); break;
    case 8u: sweep_phase_d<8, false>(c, t); break;
    default: sweep_phase_d<0, false>(c, t);
  }
}

__device__ __forceinline__ uint32_t event_index(uint32_t stp, uint32_t S) {
  return (stp % S == 0u) ? stp / S : stp / S + 1u;
}

// 4. measurements (src/pt.rs:167-226) and ground-state tracking (src/pt.rs:559-581)
__device__ __forceinline__ void measure_phase(Ctx &c, uint32_t t) {
  const TAMC_AS4 PtGroup &X = *c.X;
  if (t < X.meas_init) return;
  const uint32_t stp = t - X.meas_init;
  const bool last = (t == X.T - 1u);
  const uint32_t B = c.B, G = c.G, n = c.n;
  if ((c.flags & kFlagSample) && (stp % X.sample == 0u || last)) {
    const uint32_t ev = event_index(stp, X.sample);
    for (uint32_t k = c.tid; k < c.nch * B; k += c.nthreads)
      X.out_e[(size_t)ev * c.nch * B + k] = c.energy[c.perm[k]];
    if (c.nch == 2) {
      // q[b] = n - 2 * #{i : replicas at slot b of the two chains differ}
      uint32_t *cnt = (uint32_t *)c.partial;
      for (uint32_t b = 0; b < B; ++b) {
        const uint32_t ra = c.perm[b], rb = c.perm[B + b];
        uint32_t tot = 0;
        for (uint32_t blk = c.wave; blk < c.nblk; blk += c.W) {
          const uint32_t i = blk * 64u + c.lane;
          const bool d = (i < n) && (sbit(c.sp, G, i, ra) != sbit(c.sp, G, i, rb));
          tot += popc64(ballot(d));
        }
        if (c.lane == 0) cnt[c.wave * B + b] = tot;
      }
      __syncthreads();
      for (uint32_t b = c.tid; b < B; b += c.nthreads) {
        uint32_t tot = 0;
        for (uint32_t w = 0; w < c.W; ++w) tot += cnt[w * B + b];
        X.out_q[(size_t)ev * B + b] = (int32_t)n - 2 * (int32_t)tot;
      }
      if (X.nchi) {
        // suscept: sum_i w_k[i] * q_i in f64, site order (src/ising.rs:427-433)
        for (uint32_t idx = c.tid; idx < B * X.nchi; idx += c.nthreads) {
          const uint32_t b = idx / X.nchi, k = idx % X.nchi;
          const uint32_t ra = c.perm[b], rb = c.perm[B + b];
          double x = 0.0;
          const TAMC_AS4 double *w = c4(X.chi) + (size_t)k * n;
          for (uint32_t i = 0; i < n; ++i) {
            const double qi = (sbit(c.sp, G, i, ra) != sbit(c.sp, G, i, rb)) ? -1.0 : 1.0;
            x = __dadd_rn(x, __dmul_rn(w[i], qi));
          }
          X.out_chi[((size_t)ev * B + b) * X.nchi + k] = (float)x;
        }
      }
      __syncthreads();
    }
  }
  if ((c.flags & kFlagSampleStates) && (stp % X.sample_states == 0u || last)) {
    const uint32_t sev = event_index(stp, X.sample_states);
    const uint32_t ns = X.n_samp_slots;
    for (uint32_t slot = 0; slot < ns; ++slot) {
      const uint32_t b = (X.sample_limiting == 0) ? slot : ((X.sample_limiting == 1) ? B - 1u - slot : B - 1u);
      for (uint32_t ch = 0; ch < c.nch; ++ch) {
        const uint32_t r = c.perm[ch * B + b];
        uint8_t *dst = X.out_samples + (((size_t)sev * ns + slot) * c.nch + ch) * X.nbytes;
        for (uint32_t by = c.tid; by < X.nbytes; by += c.nthreads) {
          uint32_t v = 0;
          for (uint32_t k = 0; k < 8u; ++k) {
            const uint32_t i = by * 8u + k;
            if (i < n) v |= sbit(c.sp, G, i, r) << k;
          }
          dst[by] = (uint8_t)v;
        }
      }
    }
  }
  // ground state: the first minimum over chains, then slots, of this group's energies (slot index
  // k = chain * B + slot); strict improvements are recorded.  Wave 0: each lane takes the first
  // minimum of its slots k = lane (mod 64), then the wave the smallest value and, among the lanes
  // holding it, the smallest index.  The record count and minimum live in misc[2], misc[3].
  if (c.wave == 0u) {
    const uint32_t NB = c.nch * B;
    float best = __uint_as_float(0x7F800000u);  // +inf
    uint32_t bk = kNone;
    for (uint32_t k = c.lane; k < NB; k += 64u) {
      const float e = c.energy[c.perm[k]];
      if (bk == kNone || e < best) {
        best = e;
        bk = k;
      }
    }
    const float m = wave_min_f(best);
    const uint32_t k = wave_min_u((bk != kNone && best == m) ? bk : kNone);
    if (c.lane == 0u) {
      const uint32_t cnt = c.misc[2];
      uint32_t rec = kNone;
      if ((cnt == 0u || m < __uint_as_float(c.misc[3])) && cnt < X.gs_cap) {
        rec = cnt;
        c.misc[2] = cnt + 1u;
        c.misc[3] = __float_as_uint(m);
        X.out_gs_steps[cnt] = t;
        X.out_gs_e[cnt] = m;
      }
      c.misc[0] = rec;
      c.misc[1] = c.perm[k];
    }
  }
  __syncthreads();
  const uint32_t rec = c.misc[0];
  if (rec != kNone) {
    const uint32_t r = c.misc[1];
    for (uint32_t w = c.wave; w < X.gs_words; w += c.W) {
      const uint32_t i = w * 64u + c.lane;
      const uint64_t m = ballot(i < n && sbit(c.sp, G, i, r));
      if (c.lane == 0) X.out_gs_states[(size_t)rec * X.gs_words + w] = m;
    }
  }
  __syncthreads();
}

// --------------------------------------------------------------------------------------------
// Main kernel: work group blockIdx.x runs group groups[blockIdx.x] for sweeps [t0, t1).
// One work group per PtGroup.  Sweeps [t0, t1) of every group, or, when `ranges` is given,
// sweeps [ranges[2b], ranges[2b+1]) of group b (time slices of a serving loop).  When
// `launch_ticks` is given, thread 0 stores the group's 100 MHz tick count for this launch there.
// Kernel context of the calling work group: sizes, key, LDS layout pointers.  Cheap to build (a
// few scalar loads and integer ops), so the sweep loop rebuilds it every sweep from the group
// pointer passed through an empty asm statement: that keeps the compiler from holding dozens of
// loop-invariant scalars live across all phases, which it would otherwise spill to VGPR lanes and
// read back with v_readlane inside the hot loops.
template <bool kAny>
__device__ __forceinline__ Ctx make_ctx(const PtGroup *group, uint32_t icm_lds, uint32_t lds_bytes) {
  extern __shared__ __align__(16) unsigned char smem[];
  const TAMC_AS4 PtGroup &X = *c4(group);
  Ctx c;
  c.X = &X;
  c.n = X.n;
  c.G = X.G;
  c.R = X.R;
  c.B = X.B;
  c.P = X.P;
  c.nch = X.nch;
  c.lo = X.lo;
  c.flags = X.flags;
  c.nblk = X.nblk;
  c.members = X.members;
  c.member = X.member;
  c.W = blockDim.x / 64u;
  c.tid = threadIdx.x;
  c.nthreads = blockDim.x;
  c.wave = uni(threadIdx.x / 64u);
  c.lane = lane_id();
  c.key0 = (uint32_t)X.seed;
  c.key1 = (uint32_t)(X.seed >> 32);
  const bool use_lds_icm = icm_lds != 0u && !(X.flags & kFlagIcmGlobal);
  c.icm_lds = use_lds_icm;
  // kAny: the spin planes may stay in global memory (kFlagSpinsGlobal), so c.sp is a generic
  // pointer; otherwise they are in LDS and every access is an LDS instruction
  const bool spins_global = kAny && (X.flags & kFlagSpinsGlobal) != 0u;
  const LdsLayout L = with_nbr(make_layout(c.n, c.G, c.R, c.B, c.P, c.W, use_lds_icm, !spins_global), c.n,
                               X.ell_col16 ? X.ell_D : 0u, lds_bytes);
  c.sp = spins_global ? X.spins : (uint32_t *)(smem + L.spins);
  c.energy = (float *)(smem + L.energy);
  c.beta_of = (float *)(smem + L.beta_of);
  c.perm = (uint32_t *)(smem + L.perm);
  c.tags = (uint8_t *)(smem + L.tags);
  c.nacc = (uint32_t *)(smem + L.nacc);
  c.hist = (uint32_t *)(smem + L.hist);
  c.rt = (uint32_t *)(smem + L.rt);
  c.pair_a = (uint32_t *)(smem + L.pair_a);
  c.pair_b = (uint32_t *)(smem + L.pair_b);
  c.pair_of = (uint32_t *)(smem + L.pair_of);
  c.seed_blk = (uint32_t *)(smem + L.seed_blk);
  c.misc = (uint32_t *)(smem + L.misc);
  c.partial = (long long *)(smem + L.partial);
  c.dEx = (float *)(smem + L.dE);
  c.drift = (long long *)(smem + L.drift);
  c.uw = (uint32_t *)(smem + L.uw);
  c.dEa = (long long *)(smem + L.dEa);
  if (use_lds_icm && c.P > 0) {
    c.qw = (uint32_t *)(smem + L.qw);
    c.cw = (uint32_t *)(smem + L.cw);
    c.blk = (uint32_t *)(smem + L.blk);
  } else {
    c.qw = X.g_qw;
    c.cw = X.g_cw;
    c.blk = X.g_blk;
  }
  c.nbr = nullptr;
  c.nbr_in_lds = false;
  if (c.P > 0 && X.ell_col16 != nullptr && X.ell_D != 0u) {
    if (L.nbr != kNone) {
      c.nbr = (const uint16_t *)(smem + L.nbr);
      c.nbr_in_lds = true;
    } else {
      c.nbr = X.ell_col16;
    }
  }
  c.moves = c.misc + 4;
  c.ticks = (unsigned long long *)(smem + L.ticks);
  c.fmask = (unsigned long long *)(smem + L.fmask);
  c.fscr = (uint32_t *)(smem + L.fscr);
  return c;
}

// One group (or one team member's copy of it): load the persistent state into LDS, run sweeps
// [t0, t1), store the state back.  kAny: the group's spin planes may stay in global memory.
template <bool kAny>
__device__ __forceinline__ void run_group(const PtGroup *group, uint32_t t0, uint32_t t1, uint32_t icm_lds,
                                          uint32_t lds_bytes, unsigned long long launch_start,
                                          unsigned long long *launch_ticks) {
  Ctx c = make_ctx<kAny>(group, icm_lds, lds_bytes);
  const TAMC_AS4 PtGroup &X = *c.X;
  const bool spins_lds = !(kAny && (c.flags & kFlagSpinsGlobal));
  const bool lead = c.member == 0u;  // the member that writes outputs and the final state
  if (c.P > 0 && c.nbr_in_lds) {
    uint16_t *nb = const_cast<uint16_t *>(c.nbr);
    for (uint32_t k = c.tid; k < c.n * nbr_stride(X.ell_D); k += c.nthreads) nb[k] = c4(X.ell_col16)[k];
  }
  if (c.tid == 0) {
    c.moves[0] = 0u;
    for (int k = 0; k <= (int)kPhaseSlots; ++k) c.ticks[k] = 0ull;
    c.misc[15] = c.members > 1u ? X.team_ctr[3u + c.member] : 0u;  // team meetings so far
    c.misc[2] = X.gs_count[0];  // the ground-state tracker, kept in LDS during the launch
    c.misc[3] = __float_as_uint(X.gs_min[0]);
  }
  // load persistent state
  const uint32_t nw = c.n * c.G * 2u;
  if (spins_lds)
    for (uint32_t k = c.tid; k < nw; k += c.nthreads) c.sp[k] = X.spins[k];
  for (uint32_t r = c.tid; r < c.R; r += c.nthreads) {
    c.energy[r] = X.energy[r];
    c.drift[r] = X.drift[r];
  }
  for (uint32_t k = c.tid; k < c.nch * c.B; k += c.nthreads) {
    c.perm[k] = X.perm[k];
    c.tags[k] = X.tags[k];
    c.nacc[k] = X.nacc[k];
    c.hist[2 * k] = X.hist[2 * k];
    c.hist[2 * k + 1] = X.hist[2 * k + 1];
  }
  if (c.tid < c.nch) c.rt[c.tid] = X.rt[c.tid];
  __syncthreads();
  for (uint32_t k = c.tid; k < c.nch * c.B; k += c.nthreads) c.beta_of[c.perm[k]] = c4(X.betas)[k % c.B];
  __syncthreads();
  const uint32_t tend = t1 < X.T ? t1 : X.T;
  if (c.tid == 0) c.ticks[kPhaseSlots] = wall_clock64();
  // one iteration per ICM chunk of at most kMaxPairs slots (a pair's moves depend on its own slot
  // only); the last chunk's iteration goes on with the rest of the sweep
  for (uint32_t t = t0, ck = 0; t < tend;) {
    const PtGroup *gp = group;
    uint32_t il = icm_lds, lb = lds_bytes;
    asm volatile("" : "+s"(gp), "+s"(il), "+s"(lb));
    Ctx cs = make_ctx<kAny>(gp, il, lb);
    if ((cs.flags & kFlagIcm) && cs.P > 0) {
      const uint32_t Pt = cs.P;
      cs.lo += ck * kMaxPairs;
      cs.P = icm_chunk(Pt - ck * kMaxPairs);
      if (cs.icm_lds) icm_phase<true>(cs, t);
      else if (kAny && !spins_lds) icm_phase<false, true>(cs, t);
      else icm_phase<false>(cs, t);
      if ((++ck) * kMaxPairs < Pt) continue;
    }
    ck = 0;
    lap(cs, kPhIcmFlip);
    if (cs.flags & kFlagIcmOnly) break;
    exchange_phase(cs, t);
    lap(cs, kPhExchange);
    sweep_phase(cs, t);
    lap(cs, kPhSweep);
    if (cs.member == 0u) measure_phase(cs, t);
    lap(cs, kPhMeasure);
    ++t;
  }
  __syncthreads();
  if (c.tid == 0 && c.members > 1u) X.team_ctr[3u + c.member] = c.misc[15];
  if (!lead) return;
  if (c.tid == 0 && X.phase_ticks)
    for (int k = 0; k < (int)kPhaseSlots; ++k) X.phase_ticks[k] += c.ticks[k];
  __syncthreads();
  if (c.tid == 0) {
    X.icm_moves[0] += c.moves[0];
    X.gs_count[0] = c.misc[2];
    X.gs_min[0] = __uint_as_float(c.misc[3]);
  }
  if (c.tid == 0 && launch_ticks) launch_ticks[blockIdx.x] = wall_clock64() - launch_start;
  // store persistent state
  if (spins_lds)
    for (uint32_t k = c.tid; k < nw; k += c.nthreads) X.spins[k] = c.sp[k];
  for (uint32_t r = c.tid; r < c.R; r += c.nthreads) {
    X.energy[r] = c.energy[r];
    X.drift[r] = c.drift[r];
  }
  for (uint32_t k = c.tid; k < c.nch * c.B; k += c.nthreads) {
    X.perm[k] = c.perm[k];
    X.tags[k] = c.tags[k];
    X.nacc[k] = c.nacc[k];
    X.hist[2 * k] = c.hist[2 * k];
    X.hist[2 * k + 1] = c.hist[2 * k + 1];
  }
  if (c.tid < c.nch) X.rt[c.tid] = c.rt[c.tid];
}

// Work group b runs the groups jobs[b] .. jobs[b + 1] - 1 one after another (without a job table,
// group b alone).  The members of a team are in different work groups of one launch, all resident
// at once, and come first in their work groups' job lists.
template <bool kAny>
__device__ __forceinline__ void sweep_kernel_body(const PtGroup *groups, uint32_t t0, uint32_t t1, uint32_t icm_lds,
                                                  uint32_t lds_bytes, const uint32_t *ranges,
                                                  unsigned long long *launch_ticks, const uint32_t *jobs) {
  if (ranges) {
    t0 = c4(ranges)[2u * blockIdx.x];
    t1 = c4(ranges)[2u * blockIdx.x + 1u];
  }
  const unsigned long long launch_start = wall_clock64();
  const uint32_t j0 = jobs ? c4(jobs)[blockIdx.x] : blockIdx.x;
  const uint32_t j1 = jobs ? c4(jobs)[blockIdx.x + 1u] : blockIdx.x + 1u;
  for (uint32_t j = j0; j < j1; ++j) {
    run_group<kAny>(groups + j, t0, t1, icm_lds, lds_bytes, launch_start, launch_ticks);
    __syncthreads();
  }
}

#ifndef TAMC_SWEEP_KERNEL_ANY
__global__ __launch_bounds__(1024) void pt_sweep_kernel(const PtGroup *groups, uint32_t t0, uint32_t t1,
                                                        uint32_t icm_lds, uint32_t lds_bytes, const uint32_t *ranges,
                                                        unsigned long long *launch_ticks, const uint32_t *jobs) {
  sweep_kernel_body<false>(groups, t0, t1, icm_lds, lds_bytes, ranges, launch_ticks, jobs);
}
#else
// The same for launches holding a group whose spin planes stay in global memory (a group too large
// for LDS): spin accesses go through generic pointers.  Built from pt_kernels_any.hip.
__global__ __launch_bounds__(1024) void pt_sweep_kernel_any(const PtGroup *groups, uint32_t t0, uint32_t t1,
                                                            uint32_t icm_lds, uint32_t lds_bytes, const uint32_t *ranges,
                                                            unsigned long long *launch_ticks, const uint32_t *jobs) {
  sweep_kernel_body<true>(groups, t0, t1, icm_lds, lds_bytes, ranges, launch_ticks, jobs);
}
#endif

}  // namespace tamc

#ifdef TAMC_SWEEP_KERNEL_ANY
extern "C" hipError_t tamc_launch_pt_sweeps_any(const tamc::PtGroup *groups, uint32_t ngroups, uint32_t t0, uint32_t t1,
                                                uint32_t waves, uint32_t lds_bytes, uint32_t icm_lds,
                                                const uint32_t *ranges, unsigned long long *launch_ticks, hipStream_t s,
                                                const uint32_t *jobs) {
  // dynamic LDS above 64 KiB must be requested explicitly; the attribute is raised to the largest
  // size any launch has used (callers may launch from several threads)
  static std::mutex mu;
  static uint32_t attr_bytes = 65536u;
  if (lds_bytes > 65536u) {
    std::lock_guard<std::mutex> lk(mu);
    if (lds_bytes > attr_bytes) {
      const hipError_t e = hipFuncSetAttribute((const void *)tamc::pt_sweep_kernel_any,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
      if (e != hipSuccess) return e;
      attr_bytes = lds_bytes;
    }
  }
  hipLaunchKernelGGL(tamc::pt_sweep_kernel_any, dim3(ngroups), dim3(waves * 64u), lds_bytes, s, groups, t0, t1,
                     icm_lds, lds_bytes, ranges, launch_ticks, jobs);
  return hipGetLastError();
}
#else
extern "C" hipError_t tamc_launch_pt_sweeps_any(const tamc::PtGroup *groups, uint32_t ngroups, uint32_t t0, uint32_t t1,
                                                uint32_t waves, uint32_t lds_bytes, uint32_t icm_lds,
                                                const uint32_t *ranges, unsigned long long *launch_ticks, hipStream_t s,
                                                const uint32_t *jobs);

// ------------------------------------------------------------------------ launch wrappers ----
// Work groups of pt_sweep_kernel that can be resident at once on the device (a launch with teams
// must not exceed it).
extern "C" hipError_t tamc_pt_sweep_capacity(uint32_t waves, uint32_t lds_bytes, uint32_t *capacity) {
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess && lds_bytes > 65536u)
    e = hipFuncSetAttribute((const void *)tamc::pt_sweep_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds_bytes);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)tamc::pt_sweep_kernel, (int)(waves * 64u),
                                                     (size_t)lds_bytes);
  *capacity = e == hipSuccess ? (uint32_t)(cus * per_cu) : 0u;
  return e;
}

extern "C" hipError_t tamc_launch_pt_init(const tamc::PtGroup *groups, uint32_t ngroups,
                                          const uint32_t *const *init_bits, hipStream_t s) {
  hipLaunchKernelGGL(tamc::pt_init_kernel, dim3(ngroups), dim3(256), 0, s, groups, init_bits);
  return hipGetLastError();
}

extern "C" hipError_t tamc_launch_pt_sweeps(const tamc::PtGroup *groups, uint32_t ngroups, uint32_t t0, uint32_t t1,
                                            uint32_t waves, uint32_t lds_bytes, uint32_t icm_lds, const uint32_t *ranges,
                                            unsigned long long *launch_ticks, hipStream_t s, const uint32_t *jobs,
                                            uint32_t spins_global) {
  if (spins_global)
    return tamc_launch_pt_sweeps_any(groups, ngroups, t0, t1, waves, lds_bytes, icm_lds, ranges, launch_ticks, s, jobs);
  // dynamic LDS above 64 KiB must be requested explicitly; the attribute is raised to the largest
  // size any launch has used (callers may launch from several threads)
  static std::mutex mu;
  static uint32_t attr_bytes = 65536u;
  if (lds_bytes > 65536u) {
    std::lock_guard<std::mutex> lk(mu);
    if (lds_bytes > attr_bytes) {
      const hipError_t e = hipFuncSetAttribute((const void *)tamc::pt_sweep_kernel,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
      if (e != hipSuccess) return e;
      attr_bytes = lds_bytes;
    }
  }
  // With teams (jobs set), members wait for each other, so every work group must be resident at
  // once: the planner keeps the grid within tamc_pt_sweep_capacity, which an otherwise idle device
  // dispatches in full, and the members' waits are bounded, so a shortfall ends the launch with an
  // error instead of a hang.  (A cooperative launch would guarantee residency, but the profiler
  // does not finish processes that use one.)
  hipLaunchKernelGGL(tamc::pt_sweep_kernel, dim3(ngroups), dim3(waves * 64u), lds_bytes, s, groups, t0, t1, icm_lds,
                     lds_bytes, ranges, launch_ticks, jobs);
  return hipGetLastError();
}
#endif  // TAMC_SWEEP_KERNEL_ANY
