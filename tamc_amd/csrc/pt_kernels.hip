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
// The sweep visits site quads level by level.  A quad's level is one more than the highest level
// of any lower-numbered quad it is coupled to, so each site sees exactly the neighbour spins it
// would see in the reference's sequential 0..N-1 order; within a quad one lane updates its replica's
// four sites in order.  Spin trajectories therefore equal the sequential algorithm's given the same
// uniforms.  The uniform for (sweep t, site i, replica rho) is word i&3 of
// Philox4x32-10(ctr = {t, i>>2, rho, 0}, key = seed): independent of tiling and of batch placement.
#include <hip/hip_runtime.h>

#include "pt_types.hpp"
#include "tamc_math.hpp"

#define TAMC_EXC_DECL static __device__ const
#include "generated/expf_exceptions.inc"

namespace tamc {

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

// u < expf(x) for u = (w >> 8) * 2^-24, decided with the hardware exp2 unless u falls within a
// relative 2^-12 band of the estimate; inside the band the exact double-precision path decides.
__device__ __forceinline__ bool accept_u(float x, uint32_t w) {
  const uint32_t k = w >> 8;
  const float p = __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
  const float P = p * 16777216.0f;
  const float kf = (float)k;
  if (kf < P * (1.0f - 0x1p-12f)) return true;
  if (kf > P * (1.0f + 0x1p-12f)) return false;
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

// --------------------------------------------------------------------------------------------
// Device-side initialisation of a group: bit planes from init states or Philox, energies by
// energy_ref in the reference's order (src/ising.rs:438-454), identity permutation, fresh tags.
// One work group per PtGroup; lane = physical replica.
// init_bits: nullable [n][G][2] bit planes prepared by the host from caller init states.
__global__ void pt_init_kernel(const PtGroup *groups, const uint32_t *const *init_bits) {
  const PtGroup &X = groups[blockIdx.x];
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
      h += X.bias[i];
      for (uint32_t e = X.row_ptr[i]; e < X.row_ptr[i + 1]; ++e) {
        const float sj = sbit(X.spins, G, X.col[e], r) ? -1.0f : 1.0f;
        h += __fmul_rn(X.J[e], sj) / 2.0f;
      }
      total = __fadd_rn(total, __fmul_rn(h, si));
    }
    X.energy[r] = total;
  }
  for (uint32_t k = threadIdx.x; k < X.nch * B; k += blockDim.x) {
    const uint32_t b = k % B;
    X.perm[k] = k;  // chain c slot b holds replica c*B+b
    X.tags[k] = (b == 0) ? 1u : ((b == B - 1) ? 2u : 0u);  // MinBeta / MaxBeta / None
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

// --------------------------------------------------------------------------------------------
struct Ctx {
  const PtGroup *X;
  uint32_t *sp;      // LDS bit planes
  float *energy, *beta_of, *dEx;
  long long *partial;
  uint32_t *perm, *pair_a, *pair_b, *pair_of, *seed_blk, *seed_rank, *misc;
  uint8_t *tags;
  uint32_t *qw, *cw, *blk;
  uint32_t n, G, R, B, P, nch, lo, W, wave, lane, tid, nthreads, key0, key1, flags, nblk;
};

// 1. Houdayer ICM for all pairs of the group at once (src/pt.rs:31-77).
__device__ __forceinline__ void icm_phase(Ctx &c, uint32_t t) {
  const PtGroup &X = *c.X;
  const uint32_t P = c.P, n = c.n, G = c.G;
  if (c.tid < P) {
    const uint32_t b = c.lo + c.tid;
    c.pair_a[c.tid] = c.perm[b];
    c.pair_b[c.tid] = c.perm[c.B + b];
  }
  for (uint32_t r = c.tid; r < c.R; r += c.nthreads) c.pair_of[r] = kNone;
  __syncthreads();
  if (c.tid < P) {
    c.pair_of[c.pair_a[c.tid]] = c.tid;
    c.pair_of[c.pair_b[c.tid]] = c.tid;
  }
  // (a) overlap words: bit p of qw[i] <=> replicas of pair p disagree at site i; block counts
  for (uint32_t blk = c.wave; blk < c.nblk; blk += c.W) {
    const uint32_t i = blk * 64u + c.lane;
    const bool valid = i < n;
    const uint32_t ii = valid ? i : 0u;
    uint32_t q = 0;
    for (uint32_t p = 0; p < P; ++p) {
      const uint32_t a = c.pair_a[p], b = c.pair_b[p];
      const uint32_t qb = (valid ? (sbit(c.sp, G, ii, a) ^ sbit(c.sp, G, ii, b)) : 0u);
      q |= qb << p;
      const uint32_t cnt = popc64(ballot(qb != 0u));
      if (c.lane == 0) c.blk[blk * P + p] = cnt;
    }
    if (valid) {
      c.qw[i] = q;
      c.cw[i] = 0u;
    }
  }
  __syncthreads();
  // (b) seed choice: k = floor(w * count / 2^32) into the ascending list of q<0 sites
  if (c.wave == 0 && c.lane < P) {
    const uint32_t p = c.lane;
    uint32_t total = 0;
    for (uint32_t blk = 0; blk < c.nblk; ++blk) total += c.blk[blk * P + p];
    uint32_t sb = kNone, sr = 0;
    if (total) {
      uint32_t k;
      if (X.icm_k) {
        k = X.icm_k[p];
        if (k >= total) k = total - 1u;
      } else {
        const U4 v = philox4x32_10(t, c.lo + p, X.pair_global, kTagIcm, c.key0, c.key1);
        k = (uint32_t)(((uint64_t)v.x * total) >> 32);
      }
      uint32_t cum = 0;
      for (uint32_t blk = 0; blk < c.nblk; ++blk) {
        const uint32_t cb = c.blk[blk * P + p];
        if (k < cum + cb) {
          sb = blk;
          sr = k - cum;
          break;
        }
        cum += cb;
      }
    }
    c.seed_blk[p] = sb;
    c.seed_rank[p] = sr;
  }
  __syncthreads();
  if (c.wave == 0) {
    uint32_t moves = 0;
    for (uint32_t p = 0; p < P; ++p) {
      const uint32_t sb = c.seed_blk[p];
      if (sb == kNone) continue;
      ++moves;
      const uint32_t i = sb * 64u + c.lane;
      const uint32_t qb = (i < n) ? ((c.qw[i] >> p) & 1u) : 0u;
      const uint64_t m = ballot(qb != 0u);
      if (qb && mbcnt64(m) == c.seed_rank[p]) c.cw[i] |= (1u << p);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    }
    if (c.lane == 0 && moves) atomicAdd(X.icm_moves, (unsigned long long)moves);
  }
  __syncthreads();
  // (c) cluster labels: grow every pair's cluster inside its q<0 sites until nothing changes
  for (;;) {
    int changed = 0;
    for (uint32_t i = c.tid; i < n; i += c.nthreads) {
      const uint32_t q = c.qw[i];
      if (!q) continue;
      const uint32_t cur = c.cw[i];
      uint32_t nb = 0;
      if (i < X.g_nodes) {
        const uint32_t e0 = X.g_row_ptr[i], e1 = X.g_row_ptr[i + 1];
        for (uint32_t e = e0; e < e1; ++e) nb |= c.cw[X.g_col[e]];
      }
      const uint32_t nc = cur | (q & nb);
      if (nc != cur) {
        c.cw[i] = nc;
        changed = 1;
      }
    }
    if (!__syncthreads_or(changed)) break;
  }
  // (d) swap the cluster spins of both replicas (= flip both) and update their energies
  //     incrementally: dE_r = sum_{i in cl} -2 s_i (h_i + sum_{j notin cl} J_ij s_j)
  long long acc[4] = {0, 0, 0, 0};
  for (uint32_t blk = c.wave; blk < c.nblk; blk += c.W) {
    const uint32_t i0 = blk * 64u;
    const uint32_t il = i0 + c.lane;
    uint32_t cwl = 0u;
    if (il < n) cwl = c.cw[il];
    uint64_t m = ballot(cwl != 0u);
    while (m) {
      const uint32_t off = (uint32_t)__builtin_ctzll(m);
      m &= m - 1;
      const uint32_t i = uni(i0 + off);
      const uint32_t cl = c.cw[i];
#pragma unroll
      for (uint32_t g = 0; g < 4u; ++g) {
        if (g >= G) break;
        const uint32_t r = g * 64u + c.lane;
        const uint32_t pr = (r < c.R) ? c.pair_of[r] : kNone;
        const bool in = (pr != kNone) && ((cl >> pr) & 1u);
        if (ballot(in) == 0ull) continue;
        const uint32_t w = c.sp[(i * G + g) * 2u + (c.lane >> 5)];
        const uint32_t bi = (w >> (c.lane & 31u)) & 1u;
        if (in) {
          float h = X.bias[i];
          const uint32_t e0 = X.row_ptr[i], e1 = X.row_ptr[i + 1];
          for (uint32_t e = e0; e < e1; ++e) {
            const uint32_t j = X.col[e];
            if (!((c.cw[j] >> pr) & 1u)) {
              const float Je = X.J[e];
              h += sbit(c.sp, G, j, r) ? -Je : Je;
            }
          }
          const float term = bi ? 2.0f * h : -2.0f * h;
          acc[g] += to_fix(term);
        }
        const uint64_t nw = ballot((bi ^ (in ? 1u : 0u)) != 0u);
        if ((c.lane & 31u) == 0u) c.sp[(i * G + g) * 2u + (c.lane >> 5)] = (c.lane >> 5) ? (uint32_t)(nw >> 32) : (uint32_t)nw;
      }
    }
  }
#pragma unroll
  for (uint32_t g = 0; g < 4u; ++g)
    if (g < G) c.partial[(c.wave * G + g) * 64u + c.lane] = acc[g];
  __syncthreads();
  for (uint32_t r = c.tid; r < c.R; r += c.nthreads) {
    if (c.pair_of[r] == kNone) continue;
    const uint32_t g = r >> 6, l = r & 63u;
    long long s = 0;
    for (uint32_t w = 0; w < c.W; ++w) s += c.partial[(w * G + g) * 64u + l];
    c.energy[r] = add_fix(c.energy[r], s);
  }
  __syncthreads();
}

// 2. replica exchange and round trips, one lane per chain of the group
__device__ __forceinline__ void exchange_phase(Ctx &c, uint32_t t) {
  const PtGroup &X = *c.X;
  const uint32_t B = c.B;
  if (c.wave == 0 && c.lane < c.nch) {
    const uint32_t ch = c.lane;
    const uint32_t cg = X.chain0 + ch;
    uint32_t *perm = c.perm + ch * B;
    uint8_t *tags = c.tags + ch * B;
    float *dE = c.dEx + ch * B;
    for (uint32_t j = 0; j + 1 < B; ++j) dE[j] = c.energy[perm[j + 1]] - c.energy[perm[j]];
    const bool threaded = (c.flags & kFlagThreaded) != 0u;
    U4 v = {0, 0, 0, 0};
    uint32_t vblk = kNone;
    for (uint32_t j = 0; j + 1 < B; ++j) {
      const float dbeta = X.betas[j + 1] - X.betas[j];
      const float dlt = dbeta * dE[j];
      bool ok = dlt >= 0.0f;
      if (!ok) {
        if ((j >> 2) != vblk) {
          vblk = j >> 2;
          v = philox4x32_10(t, vblk, cg, kTagExchange, c.key0, c.key1);
        }
        ok = accept_u(dlt, pick(v, j));
      }
      if (ok) {
        const uint32_t tp = perm[j];
        perm[j] = perm[j + 1];
        perm[j + 1] = tp;
        const uint8_t tt = tags[j];
        tags[j] = tags[j + 1];
        tags[j + 1] = tt;
        X.nacc[ch * B + j] += 1u;
        if (!threaded) {
          if (j + 2 < B) dE[j + 1] = dE[j + 1] + dE[j];
          dE[j] = -dE[j];
        }
      }
    }
    // update_round_trips (tamc-core/src/pt.rs:54-84); tags: 0 None 1 Min 2 Max 3 MinRefl 4 MaxRefl
    uint8_t lo = tags[0];
    if (lo == 0) lo = 1;
    else if (lo == 4) {
      X.rt[ch] += 1u;
      lo = 1;
    } else if (lo == 2) lo = 3;
    tags[0] = lo;
    uint8_t hi = tags[B - 1];
    if (hi == 0) hi = 2;
    else if (hi == 3) {
      X.rt[ch] += 1u;
      hi = 2;
    } else if (hi == 1) hi = 4;
    tags[B - 1] = hi;
    for (uint32_t b = 0; b < B; ++b) {
      const uint8_t tg = tags[b];
      if (tg == 1 || tg == 3) X.hist[(ch * B + b) * 2u] += 1u;
      else if (tg == 2 || tg == 4) X.hist[(ch * B + b) * 2u + 1u] += 1u;
    }
  }
  __syncthreads();
  for (uint32_t k = c.tid; k < c.nch * B; k += c.nthreads) c.beta_of[c.perm[k]] = X.betas[k % B];
  __syncthreads();
}

// 3. Metropolis sweep, quad levels
__device__ __forceinline__ void sweep_phase(Ctx &c, uint32_t t) {
  const PtGroup &X = *c.X;
  const uint32_t G = c.G, n = c.n;
  long long acc[4] = {0, 0, 0, 0};
  const uint32_t half = c.lane >> 5, sh = c.lane & 31u;
  for (uint32_t l = 0; l < X.n_levels; ++l) {
    const uint32_t a = X.lvl_off[l], b = X.lvl_off[l + 1];
    for (uint32_t k = a + c.wave; k < b; k += c.W) {
      const uint32_t task = uni(X.tasks[k]);
      const uint32_t quad = task / G, g = task - quad * G;
      const uint32_t r = g * 64u + c.lane;
      const bool valid = r < c.R;
      const float beta = valid ? c.beta_of[r] : 0.0f;
      const uint32_t rho = X.chain0 * c.B + r;
      const U4 u = philox4x32_10(t, quad, rho, kTagSweep, c.key0, c.key1);
      float a_g = 0.f;
      const uint32_t i0 = quad * 4u;
#pragma unroll
      for (uint32_t s = 0; s < 4u; ++s) {
        const uint32_t i = i0 + s;
        if (i >= n) break;
        float d = 0.0f;
        d += X.bias[i];
        const uint32_t e0 = X.row_ptr[i], e1 = X.row_ptr[i + 1];
        for (uint32_t e = e0; e < e1; ++e) {
          const uint32_t j = X.col[e];
          const float Je = X.J[e];
          const uint32_t wj = c.sp[(j * G + g) * 2u + half];
          d += ((wj >> sh) & 1u) ? -Je : Je;
        }
        const uint32_t wi = c.sp[(i * G + g) * 2u + half];
        const uint32_t bi = (wi >> sh) & 1u;
        d *= bi ? 2.0f : -2.0f;  // d *= -2*s_i
        bool acc_ok = d < 0.0f;
        if (!acc_ok) acc_ok = accept_u(-beta * d, pick(u, s));
        acc_ok = acc_ok && valid;
        const uint64_t nw = ballot((bi ^ (acc_ok ? 1u : 0u)) != 0u);
        if (sh == 0u) c.sp[(i * G + g) * 2u + half] = half ? (uint32_t)(nw >> 32) : (uint32_t)nw;
        if (acc_ok) a_g += d;
      }
      // the quad's accepted dE, summed in site order, enter the fixed-point total
      const long long fx = to_fix(a_g);
#pragma unroll
      for (uint32_t gg = 0; gg < 4u; ++gg)
        if (gg == g) acc[gg] += fx;
    }
    __syncthreads();
  }
#pragma unroll
  for (uint32_t g = 0; g < 4u; ++g)
    if (g < G) c.partial[(c.wave * G + g) * 64u + c.lane] = acc[g];
  __syncthreads();
  for (uint32_t r = c.tid; r < c.R; r += c.nthreads) {
    const uint32_t g = r >> 6, l = r & 63u;
    long long s = 0;
    for (uint32_t w = 0; w < c.W; ++w) s += c.partial[(w * G + g) * 64u + l];
    c.energy[r] = add_fix(c.energy[r], s);
  }
  __syncthreads();
}

__device__ __forceinline__ uint32_t event_index(uint32_t stp, uint32_t S) {
  return (stp % S == 0u) ? stp / S : stp / S + 1u;
}

// 4. measurements (src/pt.rs:167-226) and ground-state tracking (src/pt.rs:559-581)
__device__ __forceinline__ void measure_phase(Ctx &c, uint32_t t) {
  const PtGroup &X = *c.X;
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
          const double *w = X.chi + (size_t)k * n;
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
  // ground state: first minimum over chains then slots of this group; record strict improvements
  if (c.tid == 0) {
    float best = 0.f;
    uint32_t br = 0;
    bool first = true;
    for (uint32_t ch = 0; ch < c.nch; ++ch)
      for (uint32_t b = 0; b < B; ++b) {
        const uint32_t r = c.perm[ch * B + b];
        const float e = c.energy[r];
        if (first || e < best) {
          best = e;
          br = r;
          first = false;
        }
      }
    const uint32_t cnt = X.gs_count[0];
    uint32_t rec = kNone;
    if ((cnt == 0u || best < X.gs_min[0]) && cnt < X.gs_cap) {
      rec = cnt;
      X.gs_count[0] = cnt + 1u;
      X.gs_min[0] = best;
      X.out_gs_steps[cnt] = t;
      X.out_gs_e[cnt] = best;
    }
    c.misc[0] = rec;
    c.misc[1] = br;
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
__global__ __launch_bounds__(1024) void pt_sweep_kernel(const PtGroup *groups, uint32_t t0, uint32_t t1,
                                                        uint32_t icm_lds) {
  extern __shared__ __align__(16) unsigned char smem[];
  const PtGroup &X = groups[blockIdx.x];
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
  c.W = blockDim.x / 64u;
  c.tid = threadIdx.x;
  c.nthreads = blockDim.x;
  c.wave = uni(threadIdx.x / 64u);
  c.lane = lane_id();
  c.key0 = (uint32_t)X.seed;
  c.key1 = (uint32_t)(X.seed >> 32);
  const bool use_lds_icm = icm_lds != 0u && !(X.flags & kFlagIcmGlobal);
  const LdsLayout L = make_layout(c.n, c.G, c.R, c.B, c.P, c.W, use_lds_icm);
  c.sp = (uint32_t *)(smem + L.spins);
  c.energy = (float *)(smem + L.energy);
  c.beta_of = (float *)(smem + L.beta_of);
  c.perm = (uint32_t *)(smem + L.perm);
  c.tags = (uint8_t *)(smem + L.tags);
  c.pair_a = (uint32_t *)(smem + L.pair_a);
  c.pair_b = (uint32_t *)(smem + L.pair_b);
  c.pair_of = (uint32_t *)(smem + L.pair_of);
  c.seed_blk = (uint32_t *)(smem + L.seed_blk);
  c.seed_rank = (uint32_t *)(smem + L.seed_rank);
  c.misc = (uint32_t *)(smem + L.misc);
  c.partial = (long long *)(smem + L.partial);
  c.dEx = (float *)(smem + L.dE);
  if (use_lds_icm && c.P > 0) {
    c.qw = (uint32_t *)(smem + L.qw);
    c.cw = (uint32_t *)(smem + L.cw);
    c.blk = (uint32_t *)(smem + L.blk);
  } else {
    c.qw = X.g_qw;
    c.cw = X.g_cw;
    c.blk = X.g_blk;
  }
  // load persistent state
  const uint32_t nw = c.n * c.G * 2u;
  for (uint32_t k = c.tid; k < nw; k += c.nthreads) c.sp[k] = X.spins[k];
  for (uint32_t r = c.tid; r < c.R; r += c.nthreads) c.energy[r] = X.energy[r];
  for (uint32_t k = c.tid; k < c.nch * c.B; k += c.nthreads) {
    c.perm[k] = X.perm[k];
    c.tags[k] = X.tags[k];
  }
  __syncthreads();
  for (uint32_t k = c.tid; k < c.nch * c.B; k += c.nthreads) c.beta_of[c.perm[k]] = X.betas[k % c.B];
  __syncthreads();
  const uint32_t tend = t1 < X.T ? t1 : X.T;
  for (uint32_t t = t0; t < tend; ++t) {
    if ((c.flags & kFlagIcm) && c.P > 0) icm_phase(c, t);
    if (c.flags & kFlagIcmOnly) break;
    exchange_phase(c, t);
    sweep_phase(c, t);
    measure_phase(c, t);
  }
  // store persistent state
  for (uint32_t k = c.tid; k < nw; k += c.nthreads) X.spins[k] = c.sp[k];
  for (uint32_t r = c.tid; r < c.R; r += c.nthreads) X.energy[r] = c.energy[r];
  for (uint32_t k = c.tid; k < c.nch * c.B; k += c.nthreads) {
    X.perm[k] = c.perm[k];
    X.tags[k] = c.tags[k];
  }
}

}  // namespace tamc

// ------------------------------------------------------------------------ launch wrappers ----
extern "C" hipError_t tamc_launch_pt_init(const tamc::PtGroup *groups, uint32_t ngroups,
                                          const uint32_t *const *init_bits, hipStream_t s) {
  hipLaunchKernelGGL(tamc::pt_init_kernel, dim3(ngroups), dim3(256), 0, s, groups, init_bits);
  return hipGetLastError();
}

extern "C" hipError_t tamc_launch_pt_sweeps(const tamc::PtGroup *groups, uint32_t ngroups, uint32_t t0, uint32_t t1,
                                            uint32_t waves, uint32_t lds_bytes, uint32_t icm_lds, hipStream_t s) {
  static bool attr_set = false;
  if (!attr_set) {
    hipFuncSetAttribute((const void *)tamc::pt_sweep_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    attr_set = true;
  }
  hipLaunchKernelGGL(tamc::pt_sweep_kernel, dim3(ngroups), dim3(waves * 64u), lds_bytes, s, groups, t0, t1, icm_lds);
  return hipGetLastError();
}
