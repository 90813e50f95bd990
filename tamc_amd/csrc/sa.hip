// sa.hip -- simulated annealing on gfx950 (SaRunner::run / run_parallel, src/sa.rs:110-199 over
// tamc_core::sa::simulated_annealing, tamc-core/src/sa.rs:27-48).
//
// Restarts are independent, so one wavefront owns 64 of them and walks the sites in the reference's
// sequential order 0..N-1: the trajectory of every restart is exactly the sequential algorithm's.
// The wave's spins are bit planes in LDS (one 64-bit word per site, bit = restart), so a coupling
// row is fetched once per 64 restarts and each neighbour spin is one broadcast LDS read.  Uniforms
// come from Philox4x32-10(ctr = {sweep, i>>2, restart, 0}, key = seed), word i&3.
//
// A site's row is stored as blocks of eight entries (eight neighbour offsets and eight couplings,
// 64 bytes).  The next site's row is copied into the wave's LDS with one 16-byte vector load per
// lane while the current site is processed; the block loop then reads four blocks' entries and
// their 32 neighbour words together and sums them in row order, so memory latency never sits on
// the sum's dependency chain.  The last block of a row is padded with the site itself and a +0.0
// coupling, which adds a signed zero and changes no decision (dE = +-0 accepts either way).
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "runtime_common.hpp"
#include "sa_host.hpp"
#include "tamc_math.hpp"

#define TAMC_EXC_DECL static __device__ const
#include "generated/expf_exceptions.inc"

#define TAMC_AS3 __attribute__((address_space(3)))
#define TAMC_AS4 __attribute__((address_space(4)))

namespace tamc {

template <class T>
__device__ __forceinline__ const TAMC_AS4 T *sa_c4(const T *p) {
  return (const TAMC_AS4 T *)p;
}

__device__ __forceinline__ uint32_t sa_lane() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ bool sa_accept(float x, uint32_t w) {
  const uint32_t k = w >> 8;
  const float p = __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
  const float P = p * 16777216.0f;
  const float kf = (float)k;
  // the -2^-100 keeps a denormal or flushed estimate (possible only for k = 0) off the fast accept
  if (kf < __fmaf_rn(P, 1.0f - 0x1p-12f, -0x1p-100f)) return true;
  if (kf > P * (1.0f + 0x1p-12f)) return false;
  return accept_exact(x, k, ExpfExceptions{kExpfExcX, kExpfExcY, kExpfExcCount});
}

// +1.0 or -1.0 from the spin bit at position 31 - up of w (bit set <=> spin -1)
__device__ __forceinline__ float sa_sign(uint32_t w, uint32_t up) {
  return __uint_as_float(0x3f800000u | ((w << up) & 0x80000000u));
}

// Row staging: a site's row blocks are copied into a per-wave LDS buffer with one 16-byte vector
// load per lane (16 blocks = 128 entries per chunk); the copy for the next site is loaded while
// the current site is processed, so the row is read from LDS with no memory latency in the way.
// The buffer holds the chunk twice, once per half-wave, with each neighbour offset already turned
// into the LDS address of that half's word of the neighbour (planes in LDS): a neighbour read then
// needs no address arithmetic.
typedef uint32_t W4 __attribute__((ext_vector_type(4)));  // four words moved as one 16-byte access
constexpr uint32_t kRowChunkBlocks = 16;
constexpr uint32_t kRowBufWords = 2u * kRowChunkBlocks * 16u;
// LDS words of one wave: the plane words rounded up to 16 bytes, then the row buffer
__host__ __device__ __forceinline__ uint32_t sa_plane_words(uint32_t n) { return (2u * n + 3u) & ~3u; }
__host__ __device__ __forceinline__ uint32_t sa_wave_words(uint32_t n) { return sa_plane_words(n) + kRowBufWords; }

// One wave: 64 restarts of one solve, from their initial states to their final energies.
// lds: this wave's LDS, [n][2] plane words followed by the row buffer [kRowBufWords]; with kG the
// plane words stay in the task's global spins_out (an instance too large for LDS) and the LDS holds
// the row buffer only.
template <bool kG>
__device__ __forceinline__ void sa_wave(const SaTask &X, TAMC_AS3 uint32_t *lds) {
  using PW = std::conditional_t<kG, uint32_t, TAMC_AS3 uint32_t>;
  using PC = std::conditional_t<kG, const char, const TAMC_AS3 char>;
  const uint32_t lane = sa_lane(), half = lane >> 5, sh = lane & 31u, up = 31u - sh;
  const uint32_t n = X.n, nq = (n + 3u) / 4u;
  const uint32_t r = X.rbase + lane;
  const bool valid = r < X.R;
  const uint32_t key0 = (uint32_t)X.seed, key1 = (uint32_t)(X.seed >> 32);
  PW *sp = kG ? (PW *)X.spins_out : (PW *)lds;
  // this lane's half of every plane word: byte offsets are 8 * site + 4 * half
  PC *spl = (PC *)sp + half * 4u;
  // the row buffer: copy h (W4 64h .. 64h + 63) for half-wave h; with the planes in LDS its
  // offsets are absolute LDS byte addresses of that half's words
  TAMC_AS3 W4 *rb = (TAMC_AS3 W4 *)(lds + (kG ? 0u : sa_plane_words(n)));
  const TAMC_AS3 W4 *rbh = rb + half * 64u;
  const uint32_t base0 = kG ? 0u : (uint32_t)(uintptr_t)lds;
  // lane L holds W4 L of a chunk (block L / 4, part L % 4; parts 0 and 1 are neighbour offsets)
  auto stage = [&](W4 v) {
    W4 a = v, b = v;
    if (!kG && (lane & 2u) == 0u) {
      a += base0;
      b += base0 + 4u;
    }
    rb[lane] = a;
    rb[64u + lane] = b;
  };
  // the word of this lane's half at a neighbour offset from the row buffer
  auto word = [&](uint32_t off) -> uint32_t {
    if constexpr (kG) return *(const PW *)(spl + off);
    else return *(const TAMC_AS3 uint32_t *)(uintptr_t)off;
  };
  // the 32 entries of blocks b .. b + 3 of the row buffer
  auto load_batch = [&](uint32_t b, uint32_t *off, uint32_t *jb) {
#pragma unroll
    for (uint32_t u = 0; u < 4u; ++u) {
      const W4 o0 = rbh[4u * (b + u)], o1 = rbh[4u * (b + u) + 1u];
      const W4 j0 = rbh[4u * (b + u) + 2u], j1 = rbh[4u * (b + u) + 3u];
#pragma unroll
      for (uint32_t e = 0; e < 4u; ++e) {
        off[8u * u + e] = o0[e];
        off[8u * u + 4u + e] = o1[e];
        jb[8u * u + e] = j0[e];
        jb[8u * u + 4u + e] = j1[e];
      }
    }
  };
  const W4 *ginfo = (const W4 *)X.info;
  const W4 *gblk = (const W4 *)X.blocks;  // block b = W4 4b .. 4b+3
  auto put = [&](uint32_t i, uint64_t m) {
    if (sh == 0u) sp[i * 2u + half] = half ? (uint32_t)(m >> 32) : (uint32_t)m;
  };
  if (X.k0 > 0u) {
    // continue from the plane words the previous launch left (a wave's LDS writes and reads stay
    // in order, so no barrier is needed)
    if (!kG)
      for (uint32_t w = lane; w < n * 2u; w += 64u) sp[w] = X.spins_out[w];
  } else {
  // initial states: rand_ising_state with the production init stream (one Philox block per quad)
  for (uint32_t q = 0; q < nq; ++q) {
    const U4 o = philox4x32_10(0u, q, r, kTagInit, key0, key1);
    const uint32_t ws[4] = {o.x, o.y, o.z, o.w};
#pragma unroll
    for (uint32_t s = 0; s < 4u; ++s) {
      const uint32_t i = 4u * q + s;
      if (i >= n) break;
      const uint32_t bit = valid ? (((ws[s] >> 31) ^ 1u) & 1u) : 0u;
      put(i, __builtin_amdgcn_ballot_w64(bit != 0u));
    }
  }
  }
  // pipeline: site i is processed from rb while the next site's first chunk and the info of the
  // site after it are in flight
  W4 inf_c = ginfo[0], inf_n = ginfo[n > 1u ? 1u : 0u];
  stage(gblk[(size_t)inf_c.x * 4u + lane]);
  for (uint32_t k = X.k0; k < X.k1; ++k) {
    const float beta = sa_c4(X.betas)[k];
    for (uint32_t q = 0; q < nq; ++q) {
      const U4 u = philox4x32_10(k, q, r, kTagSweep, key0, key1);
      const uint32_t us[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
      for (uint32_t s = 0; s < 4u; ++s) {
        const uint32_t i = 4u * q + s;
        if (i >= n) break;
        const uint32_t i1 = (i + 1u == n) ? 0u : i + 1u, i2 = (i1 + 1u == n) ? 0u : i1 + 1u;
        const W4 chunk_n = gblk[(size_t)inf_n.x * 4u + lane];
        const W4 inf_nn = ginfo[i2];
        const uint32_t b0 = __builtin_amdgcn_readfirstlane(inf_c.x), nb = __builtin_amdgcn_readfirstlane(inf_c.y);
        // the reference starts from 0.0 and adds the bias (src/ising.rs:468-480)
        float d = __fadd_rn(0.0f, __uint_as_float(__builtin_amdgcn_readfirstlane(inf_c.z)));
        for (uint32_t c0 = 0; c0 < nb; c0 += kRowChunkBlocks) {
          if (c0 > 0u) stage(gblk[((size_t)b0 + c0) * 4u + lane]);  // rows beyond one chunk
          const uint32_t cb = (nb - c0 < kRowChunkBlocks) ? nb - c0 : kRowChunkBlocks;
          // four blocks at a time: their 32 neighbour words are read together and summed in row
          // order, while the entries of the next four blocks are read (buffers A and B in turn; a
          // read ahead past the chunk's last batch is clamped to blocks 12 .. 15 of the buffer and
          // not used); a remainder of one to three blocks goes singly
          uint32_t b = 0;
          if (cb >= 4u) {
            uint32_t oA[32], jA[32], oB[32], jB[32], w[32];
            load_batch(0u, oA, jA);
            for (; b + 8u <= cb; b += 8u) {
#pragma unroll
              for (int e = 0; e < 32; ++e) w[e] = word(oA[e]);
              load_batch(b + 4u, oB, jB);
              // d += J * s_j in row order; J * (+-1.0) is exact, so each fma rounds like the add
#pragma unroll
              for (int e = 0; e < 32; ++e) d = __fmaf_rn(__uint_as_float(jA[e]), sa_sign(w[e], up), d);
#pragma unroll
              for (int e = 0; e < 32; ++e) w[e] = word(oB[e]);
              load_batch(b + 8u < kRowChunkBlocks - 4u ? b + 8u : kRowChunkBlocks - 4u, oA, jA);
#pragma unroll
              for (int e = 0; e < 32; ++e) d = __fmaf_rn(__uint_as_float(jB[e]), sa_sign(w[e], up), d);
            }
            if (b + 4u <= cb) {
#pragma unroll
              for (int e = 0; e < 32; ++e) w[e] = word(oA[e]);
#pragma unroll
              for (int e = 0; e < 32; ++e) d = __fmaf_rn(__uint_as_float(jA[e]), sa_sign(w[e], up), d);
              b += 4u;
            }
          }
          for (; b < cb; ++b) {
            const W4 o0 = rbh[4u * b], o1 = rbh[4u * b + 1u], j0 = rbh[4u * b + 2u], j1 = rbh[4u * b + 3u];
            uint32_t w[8];
#pragma unroll
            for (uint32_t e = 0; e < 4u; ++e) {
              w[e] = word(o0[e]);
              w[4u + e] = word(o1[e]);
            }
#pragma unroll
            for (uint32_t e = 0; e < 4u; ++e) d = __fmaf_rn(__uint_as_float(j0[e]), sa_sign(w[e], up), d);
#pragma unroll
            for (uint32_t e = 0; e < 4u; ++e) d = __fmaf_rn(__uint_as_float(j1[e]), sa_sign(w[4u + e], up), d);
          }
        }
        const uint32_t bi = (sp[i * 2u + half] >> sh) & 1u;
        d *= bi ? 2.0f : -2.0f;
        bool acc = d < 0.0f;
        if (!acc) acc = sa_accept(-beta * d, us[s]);
        acc = acc && valid;
        put(i, __builtin_amdgcn_ballot_w64((bi ^ (acc ? 1u : 0u)) != 0u));
        stage(chunk_n);  // after this site's reads of the buffer: LDS operations stay in order
        inf_c = inf_n;
        inf_n = inf_nn;
      }
    }
  }
  if (X.k1 < X.K) {  // more sweeps follow in a later launch: keep the planes
    if (!kG)
      for (uint32_t w = lane; w < n * 2u; w += 64u) X.spins_out[w] = sp[w];
    return;
  }
  // AnnealMinResults::apply_measurements: energy_ref of every final state (src/sa.rs:55-64), the
  // real row entries in the reference's order
  const TAMC_AS4 uint32_t *info = sa_c4(X.info);
  const TAMC_AS4 uint32_t *blocks = sa_c4(X.blocks);
  float total = X.offset;
  for (uint32_t i = 0; i < n; ++i) {
    const TAMC_AS4 uint32_t *inf = info + 4u * i;
    const uint32_t b0 = inf[0], len = inf[3];
    const float si = sa_sign(*(const PW *)(spl + 8u * i), up);
    float h = 0.0f;
    h += __uint_as_float(inf[2]);
    const TAMC_AS4 uint32_t *blk = blocks + (size_t)b0 * 16u;
    for (uint32_t e = 0; e < len; ++e) {
      const uint32_t bl = e >> 3, k = e & 7u;
      const float sj = sa_sign(*(const PW *)(spl + blk[bl * 16u + k]), up);
      h += __fmul_rn(__uint_as_float(blk[bl * 16u + 8u + k]), sj) / 2.0f;
    }
    total = __fadd_rn(total, __fmul_rn(h, si));
  }
  if (X.states_out) {
    if (valid) {
      X.energy_out[lane] = total;
      int8_t *row = X.states_out + (size_t)lane * n;
      for (uint32_t i = 0; i < n; ++i) row[i] = ((sp[i * 2u + half] >> sh) & 1u) ? (int8_t)-1 : (int8_t)1;
    }
    return;
  }
  X.energy_out[lane] = total;
  if (!kG)
    for (uint32_t w = lane; w < n * 2u; w += 64u) X.spins_out[w] = sp[w];
}

// Work groups are spread over the XCDs round robin (block b runs on XCD b mod 8); the logical
// group index groups the blocks of one XCD together, so the waves of one solve (consecutive tasks)
// share that XCD's L2 for their row records.
__device__ __forceinline__ uint32_t sa_logical_group() {
  const uint32_t per = (gridDim.x + 7u) / 8u;
  return (blockIdx.x % 8u) * per + blockIdx.x / 8u;
}

// kG: the instances' plane words stay in global memory (some instance of the launch is too large
// for LDS), the LDS holds the row buffers only
template <bool kG>
__global__ __launch_bounds__(256) void sa_kernel(const SaTask *tasks, uint32_t ntasks, uint32_t words_per_wave,
                                                 uint32_t ngroups) {
  extern __shared__ __align__(16) uint32_t sa_smem[];
  const uint32_t g = sa_logical_group();
  if (g >= ngroups) return;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64u);
  const uint32_t tix = g * (blockDim.x / 64u) + wave;
  if (tix >= ntasks) return;
  sa_wave<kG>(tasks[tix], (TAMC_AS3 uint32_t *)sa_smem + (size_t)wave * words_per_wave);
}

// One solve whose waves are described by one task passed by value: wave w takes restarts
// 64w .. 64w+63 and writes int8 states and energies straight into the caller's [R] arrays.
__global__ __launch_bounds__(256) void sa_kernel_direct(SaTask base, uint32_t ntasks, uint32_t words_per_wave,
                                                        uint32_t ngroups) {
  extern __shared__ __align__(16) uint32_t sa_smem[];
  const uint32_t g = sa_logical_group();
  if (g >= ngroups) return;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64u);
  const uint32_t tix = g * (blockDim.x / 64u) + wave;
  if (tix >= ntasks) return;
  SaTask X = base;
  X.rbase = tix * 64u;
  X.energy_out = base.energy_out + X.rbase;
  X.states_out = base.states_out + (size_t)X.rbase * base.n;
  sa_wave<false>(X, (TAMC_AS3 uint32_t *)sa_smem + (size_t)wave * words_per_wave);
}

// Scratch of the allocation-free entry point: the schedule's betas are staged here.  Host code
// uses the address hipGetSymbolAddress returns for the current device, never the array's name.
constexpr size_t kNoAllocScratchBytes = 512u * 1024u;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// grid of the launch: 8 * ceil(groups / 8) blocks, so every logical group has a block
inline uint32_t sa_grid(uint32_t ngroups) { return 8u * ((ngroups + 7u) / 8u); }

tamc_status sa_validate(const tamc_instance *inst, const tamc_sa_params &P) {
  if (!inst || P.num_betas < 1 || !P.betas) return fail(TAMC_E_INVALID_ARG, "invalid SA parameters");
  if (P.num_replicas > 1 && P.threads == 0) return fail(TAMC_E_INVALID_ARG, "threads must be >= 1");
  for (uint32_t k = 1; k < P.num_betas; ++k)
    if (!(P.betas[k] - P.betas[k - 1] >= 0.0f)) return fail(TAMC_E_INVALID_ARG, "beta array must be non-decreasing");
  return TAMC_OK;
}

uint32_t sa_wave_lds_words(uint32_t n) {
  const uint32_t w = sa_wave_words(n);
  return w * 4u <= 160u * 1024u ? w : 0u;
}

// waves per work group for instances of up to maxn sites: four when their LDS fits
static uint32_t sa_waves_per_group(uint32_t wpw) {
  uint32_t wpg = 4;
  while (wpg > 1 && wpg * wpw * 4u > 160u * 1024u) wpg >>= 1;
  return wpg;
}

template <class K>
static tamc_status sa_lds_attribute(K kernel, uint32_t lds) {
  if (lds > 65536u) TAMC_HIP(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  return TAMC_OK;
}

// LDS words per wave of a launch whose instances have at most maxn sites: the planes and the row
// buffer, or the row buffer alone when the planes stay in global memory
static uint32_t sa_launch_wave_words(uint32_t maxn) {
  return sa_wave_lds_words(maxn) ? sa_wave_words(maxn) : kRowBufWords;
}

tamc_status sa_launch(const SaTask *d_tasks, uint32_t ntasks, uint32_t maxn, hipStream_t s) {
  if (!ntasks) return TAMC_OK;
  const bool global = sa_wave_lds_words(maxn) == 0u;
  const uint32_t wpw = sa_launch_wave_words(maxn), wpg = sa_waves_per_group(wpw), lds = wpg * wpw * 4u;
  const uint32_t ngroups = (ntasks + wpg - 1) / wpg;
  if (global) {
    if (tamc_status st = sa_lds_attribute(sa_kernel<true>, lds)) return st;
    hipLaunchKernelGGL(sa_kernel<true>, dim3(sa_grid(ngroups)), dim3(64 * wpg), lds, s, d_tasks, ntasks, wpw, ngroups);
  } else {
    if (tamc_status st = sa_lds_attribute(sa_kernel<false>, lds)) return st;
    hipLaunchKernelGGL(sa_kernel<false>, dim3(sa_grid(ngroups)), dim3(64 * wpg), lds, s, d_tasks, ntasks, wpw, ngroups);
  }
  TAMC_HIP(hipGetLastError());
  return TAMC_OK;
}

tamc_sa_result *sa_assemble(uint32_t n, uint32_t R, const uint32_t *sp, const float *en) {
  tamc_sa_result *res = (tamc_sa_result *)calloc(1, sizeof(tamc_sa_result));
  res->n = n;
  res->R = R;
  res->energies = (float *)malloc(std::max<size_t>(1, (size_t)R * 4));
  res->final_states = (int8_t *)malloc(std::max<size_t>(1, (size_t)R * n));
  for (uint32_t r = 0; r < R; ++r) {
    const uint32_t w = r / 64, l = r % 64;
    res->energies[r] = en[(size_t)w * 64 + l];
    for (uint32_t i = 0; i < n; ++i) {
      const uint32_t bit = (sp[((size_t)w * n + i) * 2 + (l >> 5)] >> (l & 31)) & 1u;
      res->final_states[(size_t)r * n + i] = bit ? -1 : 1;
    }
    if (r == 0 || res->energies[r] < res->min_energy) res->min_energy = res->energies[r];
  }
  return res;
}

}  // namespace tamc

using namespace tamc;

namespace {

struct SaSolve {
  const tamc_instance *inst;
  tamc_sa_params P;
  uint32_t R, waves;
  float *d_betas;
  uint32_t *d_spins;
  float *d_energy;
};

tamc_status run_sa(std::vector<SaSolve> &solves, tamc_sa_result **results, tamc_batch_stats *stats) {
  Arena A;
  uint32_t maxn = 0, ntasks = 0;
  for (auto &S : solves) {
    if (tamc_status st = sa_validate(S.inst, S.P)) return st;
    S.R = sa_kept_replicas(S.P);
    S.waves = (S.R + 63) / 64;
    maxn = std::max(maxn, S.inst->n);
    ntasks += S.waves;
    A.want(&S.d_betas, S.P.num_betas);
    A.want(&S.d_spins, (size_t)S.waves * S.inst->n * 2);
    A.want(&S.d_energy, (size_t)S.waves * 64);
  }
  SaTask *d_tasks;
  A.want(&d_tasks, ntasks);
  TAMC_HIP(A.commit());
  std::vector<SaTask> tasks;
  for (auto &S : solves) {
    tamc_status st;
    const DeviceInstance *D = device_sa_rows(S.inst, &st);
    if (!D) return st;
    TAMC_HIP(hipMemcpy(S.d_betas, S.P.betas, S.P.num_betas * 4, hipMemcpyHostToDevice));
    for (uint32_t w = 0; w < S.waves; ++w) {
      SaTask t{};
      t.info = (const uint32_t *)D->sa_info.p;
      t.blocks = (const uint32_t *)D->sa_blocks.p;
      t.betas = S.d_betas;
      t.offset = S.inst->offset;
      t.n = S.inst->n;
      t.K = S.P.num_betas;
      t.R = S.R;
      t.rbase = w * 64;
      t.k0 = 0;
      t.k1 = S.P.num_betas;
      t.seed = S.P.seed;
      t.spins_out = S.d_spins + (size_t)w * S.inst->n * 2;
      t.energy_out = S.d_energy + (size_t)w * 64;
      tasks.push_back(t);
    }
  }
  if (!tasks.empty()) TAMC_HIP(hipMemcpy(d_tasks, tasks.data(), tasks.size() * sizeof(SaTask), hipMemcpyHostToDevice));
  hipEvent_t e0, e1;
  TAMC_HIP(hipEventCreate(&e0));
  TAMC_HIP(hipEventCreate(&e1));
  const auto w0 = std::chrono::steady_clock::now();
  TAMC_HIP(hipEventRecord(e0, 0));
  if (tamc_status st = sa_launch(d_tasks, ntasks, maxn, 0)) return st;
  TAMC_HIP(hipEventRecord(e1, 0));
  TAMC_HIP(hipEventSynchronize(e1));
  const auto w1 = std::chrono::steady_clock::now();
  float ms = 0;
  TAMC_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  uint64_t attempts = 0;
  for (size_t s = 0; s < solves.size(); ++s) {
    SaSolve &S = solves[s];
    const uint32_t n = S.inst->n;
    std::vector<uint32_t> sp((size_t)S.waves * n * 2);
    std::vector<float> en((size_t)S.waves * 64);
    TAMC_HIP(hipMemcpy(sp.data(), S.d_spins, sp.size() * 4, hipMemcpyDeviceToHost));
    TAMC_HIP(hipMemcpy(en.data(), S.d_energy, en.size() * 4, hipMemcpyDeviceToHost));
    tamc_sa_result *R = sa_assemble(n, S.R, sp.data(), en.data());
    R->attempts = (uint64_t)S.P.num_betas * n * S.R;
    R->device_ms = ms;
    R->timing_us = std::chrono::duration<double, std::micro>(w1 - w0).count();
    attempts += R->attempts;
    results[s] = R;
  }
  if (stats) {
    stats->device_ms = ms;
    stats->wall_ms = std::chrono::duration<double, std::milli>(w1 - w0).count();
    stats->attempts = attempts;
    stats->waves_per_group = sa_waves_per_group(sa_launch_wave_words(maxn));
    stats->launches = 1;
    stats->groups = ntasks;
  }
  return TAMC_OK;
}

}  // namespace

extern "C" {

tamc_status tamc_sa_run_noalloc(const tamc_instance *inst, const tamc_sa_params *params, float *energies,
                                int8_t *states, void *stream) {
  clear_error();
  if (!inst || !params || !params->betas || !energies || !states) return fail(TAMC_E_INVALID_ARG, "null argument");
  const tamc_sa_params &P = *params;
  if (tamc_status st = sa_validate(inst, P)) return st;
  if (!sa_wave_lds_words(inst->n))
    return fail(TAMC_E_UNSUPPORTED, "the allocation-free SA keeps its planes in LDS (n <= 20224)");
  if ((size_t)P.num_betas * 4u > kNoAllocScratchBytes)
    return fail(TAMC_E_INVALID_ARG, "schedule longer than the allocation-free scratch holds (131072 betas)");
  const uint32_t n = inst->n;
  const uint32_t R = sa_kept_replicas(P);
  tamc_status st = device_init();
  if (st) return st;
  const DeviceInstance *D = resident_instance(inst);
  if (!D) return fail(TAMC_E_INVALID_ARG, "instance is not resident on the device: run SA on it once through tamc_sa_run");
  // the scratch address of the current device, looked up once per device
  static std::mutex mu;
  static void *dev_scratch[64] = {};
  int dev = 0;
  TAMC_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(TAMC_E_DEVICE, "device index out of range");
  void *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(mu);
    if (!dev_scratch[dev]) {
      void *p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) return fail(TAMC_E_DEVICE, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
      dev_scratch[dev] = p;
    }
    base = dev_scratch[dev];
  }
  if (R == 0) return TAMC_OK;
  hipStream_t s = (hipStream_t)stream;
  float *d_betas = (float *)base;
  TAMC_HIP(hipMemcpyAsync(d_betas, P.betas, (size_t)P.num_betas * 4u, hipMemcpyHostToDevice, s));
  SaTask t{};
  t.info = (const uint32_t *)D->sa_info.p;
  t.blocks = (const uint32_t *)D->sa_blocks.p;
  t.betas = d_betas;
  t.offset = inst->offset;
  t.n = n;
  t.K = P.num_betas;
  t.R = R;
  t.seed = P.seed;
  t.energy_out = energies;
  t.states_out = states;
  t.k0 = 0;
  t.k1 = P.num_betas;
  const uint32_t ntasks = (R + 63) / 64, wpw = sa_wave_words(n), wpg = sa_waves_per_group(wpw);
  const uint32_t lds = wpg * wpw * 4u;
  if (tamc_status e = sa_lds_attribute(sa_kernel_direct, lds)) return e;
  const uint32_t ngroups = (ntasks + wpg - 1) / wpg;
  hipLaunchKernelGGL(sa_kernel_direct, dim3(sa_grid(ngroups)), dim3(64 * wpg), lds, s, t, ntasks, wpw, ngroups);
  TAMC_HIP(hipGetLastError());
  return TAMC_OK;
}

tamc_status tamc_sa_run(const tamc_instance *inst, const tamc_sa_params *params, tamc_sa_result **out) {
  return tamc_sa_run_batch(1, &inst, params, out, nullptr);
}

tamc_status tamc_sa_run_batch(uint32_t count, const tamc_instance *const *insts, const tamc_sa_params *params,
                              tamc_sa_result **results, tamc_batch_stats *stats) {
  clear_error();
  if (!insts || !params || !results) return fail(TAMC_E_INVALID_ARG, "null argument");
  tamc_status st = device_init();
  if (st) return st;
  std::vector<SaSolve> solves(count);
  for (uint32_t k = 0; k < count; ++k) {
    solves[k].inst = insts[k];
    solves[k].P = params[k];
  }
  return run_sa(solves, results, stats);
}

void tamc_sa_result_free(tamc_sa_result *r) {
  if (!r) return;
  free(r->energies);
  free(r->final_states);
  free(r);
}

}  // extern "C"
