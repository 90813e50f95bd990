// lockstep.hip -- lock-step correctness mode of the PT-ICM path.
//
// Every random number comes from caller-supplied u32 word streams, consumed in exactly the order
// the reference consumes its generators, with the rand 0.8 conversions done on the device:
//   f32 uniform  = (w >> 8) * 2^-24                       (Standard, tamc-core/src/metropolis.rs:50)
//   ICM choice   = gen_range(0..len) widening multiply     (SliceRandom::choose, src/pt.rs:53)
// threaded (params.threads > 1, parallel_pt_loop src/pt.rs:414-457): stream c*B+b serves slot b of
// chain c; the exchange of chain c draws from stream c*B before that slot's sweep; ICM draws from
// stream 0 first.  serial (pt_loop src/pt.rs:459-504): one stream for everything, in program order.
// Spins are int8 per replica, energies are cached and recomputed by energy_ref after ICM exactly as
// src/ising.rs:455-463 does, so trajectories and measurements are bit-identical to the restatement.
// One work group runs the whole solve; thread = (chain, slot).
#include <hip/hip_runtime.h>

#include "pt_types.hpp"
#include "tamc_math.hpp"

#define TAMC_EXC_DECL static __device__ const
#include "generated/expf_exceptions.inc"

namespace tamc {

struct LockParams {
  const uint32_t *row_ptr, *col, *g_row_ptr, *g_col;
  const float *J, *bias, *betas;
  const double *chi;
  float offset;
  uint32_t n, g_nodes, nchi, B, C, lo, icm, threaded, T, meas_init;
  uint32_t sample_en, sample, sstates_en, sample_states, sample_limiting, n_slots, n_meas, n_sev, nbytes, gs_words,
      gs_cap;
  // streams
  const uint32_t *words;
  const uint64_t *off, *len;
  uint64_t *cur;
  uint32_t nstreams;
  // state
  int8_t *spins;  // [C*B][n] physical replicas
  float *energy;
  uint8_t *valid;
  uint32_t *perm, *nacc, *hist, *rt;
  uint8_t *tags;
  uint32_t *queue;
  uint8_t *mark;
  // outputs (final layout)
  float *e;
  int32_t *q;
  float *chiout;
  uint8_t *samples;
  uint32_t *gs_steps;
  float *gs_e;
  uint64_t *gs_states;
  uint32_t *gs_count;
  unsigned long long *icm_moves;
  int32_t *error;
};

__device__ __forceinline__ ExpfExceptions ls_table() { return ExpfExceptions{kExpfExcX, kExpfExcY, kExpfExcCount}; }

struct Stream {
  const LockParams *P;
  uint32_t id;
  __device__ bool next(uint32_t &w) const {
    const uint64_t c = P->cur[id];
    if (c >= P->len[id]) {
      atomicExch(P->error, 1);
      return false;
    }
    w = P->words[P->off[id] + c];
    P->cur[id] = c + 1;
    return true;
  }
  __device__ float f32(bool &ok) const {
    uint32_t w = 0;
    ok = next(w);
    return (float)(w >> 8) * (1.0f / 16777216.0f);
  }
  __device__ bool gen_index(uint32_t range, uint32_t &out) const {
    const uint32_t zone = (range << __builtin_clz(range)) - 1u;
    for (;;) {
      uint32_t v;
      if (!next(v)) return false;
      const uint64_t m = (uint64_t)v * range;
      if ((uint32_t)m <= zone) {
        out = (uint32_t)(m >> 32);
        return true;
      }
    }
  }
};

__device__ float ls_energy_ref(const LockParams &P, const int8_t *s) {
  float total = P.offset;
  for (uint32_t i = 0; i < P.n; ++i) {
    float h = 0.0f;
    const float si = (float)s[i];
    h += P.bias[i];
    for (uint32_t e = P.row_ptr[i]; e < P.row_ptr[i + 1]; ++e) h += __fmul_rn(P.J[e], (float)s[P.col[e]]) / 2.0f;
    total = __fadd_rn(total, __fmul_rn(h, si));
  }
  return total;
}

// one Metropolis sweep in site order; returns false if the stream ran dry
__device__ bool ls_sweep(const LockParams &P, int8_t *s, float beta, const Stream &st, float &de) {
  de = 0.0f;
  for (uint32_t i = 0; i < P.n; ++i) {
    float d = 0.0f;
    d += P.bias[i];
    for (uint32_t e = P.row_ptr[i]; e < P.row_ptr[i + 1]; ++e) d = __fadd_rn(d, __fmul_rn(P.J[e], (float)s[P.col[e]]));
    d = __fmul_rn(d, -2.0f * (float)s[i]);
    bool acc = d < 0.0f;
    if (!acc) {
      uint32_t w;
      if (!st.next(w)) return false;
      acc = accept_exact(__fmul_rn(-beta, d), w >> 8, ls_table());
    }
    if (acc) {
      s[i] = (int8_t)-s[i];
      de = __fadd_rn(de, d);
    }
  }
  return true;
}

__global__ void lockstep_kernel(LockParams Pv) {
  const LockParams &P = Pv;
  const uint32_t tid = threadIdx.x, B = P.B, C = P.C, n = P.n;
  const uint32_t R = B * C;
  __shared__ int s_stop;
  __shared__ float s_best;
  __shared__ uint32_t s_rec, s_rec_r;
  if (tid == 0) s_stop = 0;
  __syncthreads();
  for (uint32_t t = 0; t < P.T; ++t) {
    // ---- ICM (src/pt.rs:521-540), sequential: pairs in order, slots lo..B-1
    if (tid == 0 && P.icm && !s_stop) {
      const Stream st{&P, 0u};
      for (uint32_t p = 0; p < C / 2 && !s_stop; ++p)
        for (uint32_t b = P.lo; b < B; ++b) {
          const uint32_t r1 = P.perm[(2 * p) * B + b], r2 = P.perm[(2 * p + 1) * B + b];
          int8_t *s1 = P.spins + (size_t)r1 * n, *s2 = P.spins + (size_t)r2 * n;
          uint32_t cnt = 0;
          for (uint32_t i = 0; i < n; ++i) cnt += ((int)s1[i] * (int)s2[i] < 0);
          if (!cnt) continue;
          uint32_t k;
          if (!st.gen_index(cnt, k)) {
            s_stop = 1;
            break;
          }
          uint32_t seed = 0;
          for (uint32_t i = 0, c = 0; i < n; ++i)
            if ((int)s1[i] * (int)s2[i] < 0) {
              if (c == k) {
                seed = i;
                break;
              }
              ++c;
            }
          // BFS over the q<0-filtered coupling graph
          for (uint32_t i = 0; i < n; ++i) P.mark[i] = 0;
          uint32_t head = 0, tail = 0;
          P.mark[seed] = 1;
          P.queue[tail++] = seed;
          while (head < tail) {
            const uint32_t v = P.queue[head++];
            if (v >= P.g_nodes) continue;
            for (uint32_t e = P.g_row_ptr[v]; e < P.g_row_ptr[v + 1]; ++e) {
              const uint32_t w = P.g_col[e];
              if ((int)s1[w] * (int)s2[w] < 0 && !P.mark[w]) {
                P.mark[w] = 1;
                P.queue[tail++] = w;
              }
            }
          }
          for (uint32_t k2 = 0; k2 < tail; ++k2) {
            const uint32_t i = P.queue[k2];
            const int8_t a = s1[i];
            s1[i] = s2[i];
            s2[i] = a;
          }
          P.valid[r1] = 0;
          P.valid[r2] = 0;
          atomicAdd(P.icm_moves, 1ull);
        }
    }
    __syncthreads();
    if (s_stop) break;
    // ---- chains (tamc-core/src/pt.rs:120-150 or tamc-core/src/parallel/pt.rs:41-69)
    for (uint32_t c = 0; c < C; ++c) {
      // energies: recompute invalid ones
      if (tid < B) {
        const uint32_t r = P.perm[c * B + tid];
        if (!P.valid[r]) {
          P.energy[r] = ls_energy_ref(P, P.spins + (size_t)r * n);
          P.valid[r] = 1;
        }
      }
      __syncthreads();
      if (tid == 0 && !s_stop) {
        const Stream st{&P, P.threaded ? c * B : 0u};
        uint32_t *perm = P.perm + c * B;
        uint8_t *tags = P.tags + c * B;
        float dE[257];
        const uint32_t Bl = B < 257 ? B : 257;
        for (uint32_t j = 0; j + 1 < Bl; ++j) dE[j] = P.energy[perm[j + 1]] - P.energy[perm[j]];
        for (uint32_t j = 0; j + 1 < Bl; ++j) {
          const float dbeta = P.betas[j + 1] - P.betas[j];
          const float dlt = __fmul_rn(dbeta, dE[j]);
          bool ok = dlt >= 0.0f;
          if (!ok) {
            uint32_t w;
            if (!st.next(w)) {
              s_stop = 1;
              break;
            }
            ok = accept_exact(dlt, w >> 8, ls_table());
          }
          if (ok) {
            const uint32_t a = perm[j];
            perm[j] = perm[j + 1];
            perm[j + 1] = a;
            const uint8_t tg = tags[j];
            tags[j] = tags[j + 1];
            tags[j + 1] = tg;
            P.nacc[c * B + j] += 1u;
            if (!P.threaded) {
              if (j + 2 < B) dE[j + 1] = dE[j + 1] + dE[j];
              dE[j] = -dE[j];
            }
          }
        }
        uint8_t lo = tags[0];
        if (lo == 0) lo = 1;
        else if (lo == 4) {
          P.rt[c] += 1u;
          lo = 1;
        } else if (lo == 2) lo = 3;
        tags[0] = lo;
        uint8_t hi = tags[B - 1];
        if (hi == 0) hi = 2;
        else if (hi == 3) {
          P.rt[c] += 1u;
          hi = 2;
        } else if (hi == 1) hi = 4;
        tags[B - 1] = hi;
        for (uint32_t b = 0; b < B; ++b) {
          if (tags[b] == 1 || tags[b] == 3) P.hist[(c * B + b) * 2] += 1u;
          else if (tags[b] == 2 || tags[b] == 4) P.hist[(c * B + b) * 2 + 1] += 1u;
        }
      }
      __syncthreads();
      if (s_stop) break;
      // sweeps
      if (P.threaded) {
        if (tid < B) {
          const uint32_t r = P.perm[c * B + tid];
          float de;
          if (!ls_sweep(P, P.spins + (size_t)r * n, P.betas[tid], Stream{&P, c * B + tid}, de)) s_stop = 1;
          else P.energy[r] = __fadd_rn(P.energy[r], de);
        }
      } else if (tid == 0) {
        for (uint32_t b = 0; b < B && !s_stop; ++b) {
          const uint32_t r = P.perm[c * B + b];
          float de;
          if (!ls_sweep(P, P.spins + (size_t)r * n, P.betas[b], Stream{&P, 0u}, de)) s_stop = 1;
          else P.energy[r] = __fadd_rn(P.energy[r], de);
        }
      }
      __syncthreads();
      if (s_stop) break;
    }
    if (s_stop) break;
    // ---- measurements (src/pt.rs:542-583)
    if (t >= P.meas_init) {
      const uint32_t stp = t - P.meas_init;
      const bool last = t == P.T - 1;
      const uint32_t np = C / 2;
      if (P.sample_en && (stp % P.sample == 0 || last)) {
        const uint32_t ev = (stp % P.sample == 0) ? stp / P.sample : stp / P.sample + 1;
        if (tid < R) {
          const uint32_t c = tid / B, b = tid % B;
          P.e[(size_t)b * P.n_meas * C + (size_t)ev * C + c] = P.energy[P.perm[c * B + b]];
        }
        if (tid < np * B) {
          const uint32_t p = tid / B, b = tid % B;
          const int8_t *s1 = P.spins + (size_t)P.perm[(2 * p) * B + b] * n;
          const int8_t *s2 = P.spins + (size_t)P.perm[(2 * p + 1) * B + b] * n;
          int64_t q = 0;
          for (uint32_t i = 0; i < n; ++i) q += (int64_t)(s1[i] * s2[i]);
          P.q[(size_t)b * P.n_meas * np + (size_t)ev * np + p] = (int32_t)q;
          for (uint32_t k = 0; k < P.nchi; ++k) {
            double x = 0.0;
            for (uint32_t i = 0; i < n; ++i) x = __dadd_rn(x, __dmul_rn(P.chi[(size_t)k * n + i], (double)(s1[i] * s2[i])));
            P.chiout[((size_t)b * P.nchi + k) * P.n_meas * np + (size_t)ev * np + p] = (float)x;
          }
        }
      }
      if (P.sstates_en && (stp % P.sample_states == 0 || last)) {
        const uint32_t sev = (stp % P.sample_states == 0) ? stp / P.sample_states : stp / P.sample_states + 1;
        if (tid < P.n_slots * C) {
          const uint32_t slot = tid / C, c = tid % C;
          const uint32_t b = (P.sample_limiting == 0) ? slot : ((P.sample_limiting == 1) ? B - 1 - slot : B - 1);
          const int8_t *s = P.spins + (size_t)P.perm[c * B + b] * n;
          uint8_t *dst = P.samples + ((size_t)slot * P.n_sev * C + (size_t)sev * C + c) * P.nbytes;
          for (uint32_t by = 0; by < P.nbytes; ++by) {
            uint32_t v = 0;
            for (uint32_t k = 0; k < 8; ++k) {
              const uint32_t i = by * 8 + k;
              if (i < n && !(s[i] > 0)) v |= 1u << k;
            }
            dst[by] = (uint8_t)v;
          }
        }
      }
      if (tid == 0) {
        float best = 0.f;
        uint32_t br = 0;
        for (uint32_t c = 0; c < C; ++c)
          for (uint32_t b = 0; b < B; ++b) {
            const uint32_t r = P.perm[c * B + b];
            if ((c == 0 && b == 0) || P.energy[r] < best) {
              best = P.energy[r];
              br = r;
            }
          }
        s_rec = 0xFFFFFFFFu;
        const uint32_t cnt = P.gs_count[0];
        if ((cnt == 0 || best < s_best) && cnt < P.gs_cap) {
          s_best = best;
          s_rec = cnt;
          s_rec_r = br;
          P.gs_count[0] = cnt + 1;
          P.gs_steps[cnt] = t;
          P.gs_e[cnt] = best;
        }
      }
      __syncthreads();
      if (s_rec != 0xFFFFFFFFu) {
        const int8_t *s = P.spins + (size_t)s_rec_r * n;
        for (uint32_t w = tid; w < P.gs_words; w += blockDim.x) {
          uint64_t v = 0;
          for (uint32_t k = 0; k < 64; ++k) {
            const uint32_t i = w * 64 + k;
            if (i < n && !(s[i] > 0)) v |= 1ull << k;
          }
          P.gs_states[(size_t)s_rec * P.gs_words + w] = v;
        }
      }
      __syncthreads();
    }
  }
}

}  // namespace tamc

extern "C" hipError_t tamc_launch_lockstep(const tamc::LockParams *P, uint32_t threads, hipStream_t s) {
  hipLaunchKernelGGL(tamc::lockstep_kernel, dim3(1), dim3(threads), 0, s, *P);
  return hipGetLastError();
}

// ------------------------------------------------------------------------------ host side ----
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <vector>

#include "runtime_common.hpp"

using namespace tamc;

template <class T>
static T *ls_dup(const std::vector<T> &v) {
  T *p = (T *)malloc(v.size() ? v.size() * sizeof(T) : 1);
  if (!v.empty()) memcpy(p, v.data(), v.size() * sizeof(T));
  return p;
}

static uint32_t ls_events(uint32_t T, uint32_t mi, uint32_t S) {
  if (T <= mi || S == 0) return 0;
  const uint32_t last = T - 1 - mi;
  return last / S + 1 + (last % S ? 1 : 0);
}

extern "C" tamc_status tamc_pt_run_lockstep(const tamc_instance *inst, const tamc_pt_params *P0,
                                            const int8_t *init_states, uint32_t num_streams,
                                            const uint32_t *const *stream_words, const uint64_t *stream_lens,
                                            tamc_pt_result **out) {
  clear_error();
  if (!inst || !P0 || !out || !init_states) return fail(TAMC_E_INVALID_ARG, "null argument (init_states required)");
  const tamc_pt_params &Pp = *P0;
  const uint32_t B = Pp.num_betas, C = Pp.num_replica_chains, n = inst->n, T = Pp.num_sweeps;
  if (B < 1 || B > 256 || C < 1 || B * C > 1024) return fail(TAMC_E_UNSUPPORTED, "lock-step needs B<=256, B*C<=1024");
  for (uint32_t b = 1; b < B; ++b)
    if (!(Pp.betas[b] - Pp.betas[b - 1] >= 0.0f)) return fail(TAMC_E_INVALID_ARG, "beta array must be non-decreasing");
  if ((Pp.sample_enabled && !Pp.sample) || (Pp.sample_states_enabled && !Pp.sample_states))
    return fail(TAMC_E_INVALID_ARG, "sample intervals must be positive");
  const bool threaded = Pp.threads > 1;
  const uint32_t need = threaded ? C * B : 1;
  if (num_streams < need || !stream_words || !stream_lens) return fail(TAMC_E_INVALID_ARG, "not enough streams");
  tamc_status st = device_init();
  if (st) return st;
  DeviceInstance *D = device_instance(inst, 1, &st);
  if (!D) return st;
  const double mi = Pp.warmup_fraction * (double)T;
  const uint32_t meas_init = (mi != mi || mi <= 0.0) ? 0u : (mi >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)mi);
  const uint32_t nm = Pp.sample_enabled ? ls_events(T, meas_init, Pp.sample) : 0;
  const uint32_t nsev = Pp.sample_states_enabled ? ls_events(T, meas_init, Pp.sample_states) : 0;
  const uint32_t nslots = Pp.sample_limiting == 0 ? B : (Pp.sample_limiting == 1 ? B / 2 : 1);
  const uint32_t nbytes = (n + 7) / 8, gw = (n + 63) / 64, gcap = T > meas_init ? T - meas_init : 0;
  const uint32_t np = C / 2, nchi = (uint32_t)inst->chi.size(), R = B * C;
  for (auto &v : inst->chi)
    if (v.size() != n) return fail(TAMC_E_INVALID_ARG, "suscept vector length differs from the instance size");
  // streams
  std::vector<uint64_t> off(need), len(need);
  uint64_t total = 0;
  for (uint32_t k = 0; k < need; ++k) {
    off[k] = total;
    len[k] = stream_lens[k];
    total += len[k];
  }
  std::vector<uint32_t> words(total ? total : 1);
  for (uint32_t k = 0; k < need; ++k)
    if (len[k]) memcpy(&words[off[k]], stream_words[k], len[k] * 4);
  LockParams L{};
  Arena A;
  float *betas;
  uint32_t *dwords;
  uint64_t *doff, *dlen;
  A.want(&betas, B);
  A.want(&dwords, words.size());
  A.want(&doff, need);
  A.want(&dlen, need);
  A.want(&L.cur, need);
  A.want(&L.spins, (size_t)R * n);
  A.want(&L.energy, R);
  A.want(&L.valid, R);
  A.want(&L.perm, R);
  A.want(&L.nacc, R);
  A.want(&L.hist, 2 * R);
  A.want(&L.rt, C);
  A.want(&L.tags, R);
  A.want(&L.queue, n);
  A.want(&L.mark, n);
  A.want(&L.e, (size_t)B * nm * C);
  A.want(&L.q, (size_t)B * nm * np);
  A.want(&L.chiout, (size_t)B * nchi * nm * np);
  A.want(&L.samples, (size_t)nslots * nsev * C * nbytes);
  A.want(&L.gs_steps, gcap);
  A.want(&L.gs_e, gcap);
  A.want(&L.gs_states, (size_t)gcap * gw);
  A.want(&L.gs_count, 1);
  A.want(&L.icm_moves, 1);
  A.want(&L.error, 1);
  TAMC_HIP(A.commit());
  TAMC_HIP(hipMemset(A.buf.p, 0, A.used));
  TAMC_HIP(hipMemcpy(betas, Pp.betas, B * 4, hipMemcpyHostToDevice));
  TAMC_HIP(hipMemcpy(dwords, words.data(), words.size() * 4, hipMemcpyHostToDevice));
  TAMC_HIP(hipMemcpy(doff, off.data(), need * 8, hipMemcpyHostToDevice));
  TAMC_HIP(hipMemcpy(dlen, len.data(), need * 8, hipMemcpyHostToDevice));
  TAMC_HIP(hipMemcpy(L.spins, init_states, (size_t)R * n, hipMemcpyHostToDevice));
  std::vector<uint32_t> perm(R);
  std::vector<uint8_t> tags(R, 0);
  for (uint32_t k = 0; k < R; ++k) {
    perm[k] = k;
    const uint32_t b = k % B;
    tags[k] = b == B - 1 ? 2 : (b == 0 ? 1 : 0);  // PTState::new: the last slot's MaxBeta wins
  }
  TAMC_HIP(hipMemcpy(L.perm, perm.data(), R * 4, hipMemcpyHostToDevice));
  TAMC_HIP(hipMemcpy(L.tags, tags.data(), R, hipMemcpyHostToDevice));
  L.row_ptr = (const uint32_t *)D->row_ptr.p;
  L.col = (const uint32_t *)D->col.p;
  L.J = (const float *)D->J.p;
  L.bias = (const float *)D->bias.p;
  L.g_row_ptr = (const uint32_t *)D->g_row_ptr.p;
  L.g_col = (const uint32_t *)D->g_col.p;
  L.chi = (const double *)D->chi.p;
  L.betas = betas;
  L.offset = inst->offset;
  L.n = n;
  L.g_nodes = inst->g_nodes;
  L.nchi = nchi;
  L.B = B;
  L.C = C;
  L.lo = Pp.lo_beta_idx;
  L.icm = Pp.icm ? 1 : 0;
  L.threaded = threaded ? 1 : 0;
  L.T = T;
  L.meas_init = meas_init;
  L.sample_en = Pp.sample_enabled ? 1 : 0;
  L.sample = Pp.sample_enabled ? Pp.sample : 1;
  L.sstates_en = Pp.sample_states_enabled ? 1 : 0;
  L.sample_states = Pp.sample_states_enabled ? Pp.sample_states : 1;
  L.sample_limiting = Pp.sample_limiting;
  L.n_slots = nslots;
  L.n_meas = nm;
  L.n_sev = nsev;
  L.nbytes = nbytes;
  L.gs_words = gw;
  L.gs_cap = gcap;
  L.words = dwords;
  L.off = doff;
  L.len = dlen;
  L.nstreams = need;
  const uint32_t threads = std::min(1024u, std::max(64u, ((R + 63) / 64) * 64));
  const auto w0 = std::chrono::steady_clock::now();
  TAMC_HIP(tamc_launch_lockstep(&L, threads, 0));
  TAMC_HIP(hipDeviceSynchronize());
  const auto w1 = std::chrono::steady_clock::now();
  int32_t err = 0;
  TAMC_HIP(hipMemcpy(&err, L.error, 4, hipMemcpyDeviceToHost));
  if (err) return fail(TAMC_E_STREAM, "lock-step uniform stream exhausted");
  auto get = [&](auto &v, const auto *d, size_t count) -> tamc_status {
    v.resize(count);
    if (count) TAMC_HIP(hipMemcpy(v.data(), d, count * sizeof(v[0]), hipMemcpyDeviceToHost));
    return TAMC_OK;
  };
  std::vector<float> e, chi, energy, gse;
  std::vector<int32_t> q;
  std::vector<uint8_t> smp, otags;
  std::vector<int8_t> spins;
  std::vector<uint32_t> operm, nacc, hist, rt, gss, gsc;
  std::vector<uint64_t> gsst;
  std::vector<unsigned long long> mv;
  if ((st = get(e, L.e, (size_t)B * nm * C)) || (st = get(q, L.q, (size_t)B * nm * np)) ||
      (st = get(chi, L.chiout, (size_t)B * nchi * nm * np)) || (st = get(smp, L.samples, (size_t)nslots * nsev * C * nbytes)) ||
      (st = get(spins, L.spins, (size_t)R * n)) || (st = get(energy, L.energy, R)) || (st = get(operm, L.perm, R)) ||
      (st = get(otags, L.tags, R)) || (st = get(nacc, L.nacc, R)) || (st = get(hist, L.hist, 2 * R)) ||
      (st = get(rt, L.rt, C)) || (st = get(gsc, L.gs_count, 1)) || (st = get(mv, L.icm_moves, 1)))
    return st;
  const uint32_t ngs = std::min(gsc[0], gcap);
  if ((st = get(gss, L.gs_steps, ngs)) || (st = get(gse, L.gs_e, ngs)) || (st = get(gsst, L.gs_states, (size_t)ngs * gw)))
    return st;
  tamc_pt_result *Rr = (tamc_pt_result *)calloc(1, sizeof(tamc_pt_result));
  Rr->n = n;
  Rr->B = B;
  Rr->C = C;
  Rr->T = T;
  Rr->meas_init = meas_init;
  Rr->nchi = nchi;
  Rr->npairs = np;
  Rr->n_meas = nm;
  Rr->n_samp_slots = nslots;
  Rr->n_samp_events = nsev;
  Rr->nbytes = nbytes;
  Rr->gs_words = gw;
  Rr->n_gs = ngs;
  std::vector<int8_t> fin((size_t)R * n);
  std::vector<float> fe(R);
  std::vector<uint32_t> accs(B, 0);
  for (uint32_t k = 0; k < R; ++k) {
    const uint32_t r = operm[k];
    memcpy(&fin[(size_t)k * n], &spins[(size_t)r * n], n);
    fe[k] = energy[r];
    accs[k % B] += nacc[k];
  }
  Rr->e = ls_dup(e);
  Rr->q = ls_dup(q);
  Rr->suscept = ls_dup(chi);
  Rr->samples = ls_dup(smp);
  Rr->gs_steps = ls_dup(gss);
  Rr->gs_energies = ls_dup(gse);
  Rr->gs_states = ls_dup(gsst);
  Rr->acceptance_counts = ls_dup(accs);
  Rr->final_states = ls_dup(fin);
  Rr->final_energies = ls_dup(fe);
  Rr->final_ids = ls_dup(operm);
  Rr->num_acceptances = ls_dup(nacc);
  Rr->round_trips = ls_dup(rt);
  Rr->tags = ls_dup(otags);
  Rr->diffusion_hist = ls_dup(hist);
  Rr->icm_moves = mv[0];
  Rr->attempts = (uint64_t)T * n * B * C;
  Rr->timing_us = std::chrono::duration<double, std::micro>(w1 - w0).count();
  *out = Rr;
  return TAMC_OK;
}
