// pt_runtime.hip -- host side of the production PT-ICM path: PtIcmRunner::new resolution
// (src/pt.rs:307-351), group planning, device buffers, kernel launches and the assembly of
// PtIcmMinResults / PtIcmThermalSamples / final PTState (src/pt.rs:96-227, 542-591).
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <functional>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <thread>
#include <vector>

#include "pt_host.hpp"

namespace tamc {

static uint32_t count_events(uint32_t T, uint32_t meas_init, uint32_t S) {
  if (T <= meas_init || S == 0) return 0;
  const uint32_t last = T - 1 - meas_init;
  return last / S + 1 + (last % S ? 1 : 0);
}

tamc_status plan_solve(SolvePlan &S, const tamc_instance *inst, const tamc_pt_params *P,
                              const int8_t *init_states, uint32_t waves) {
  if (!inst || !P) return fail(TAMC_E_INVALID_ARG, "null argument");
  if (P->num_betas < 1 || !P->betas) return fail(TAMC_E_INVALID_ARG, "num_betas must be >= 1");
  if (P->num_replica_chains < 1) return fail(TAMC_E_INVALID_ARG, "num_replica_chains must be >= 1");
  for (uint32_t b = 1; b < P->num_betas; ++b)
    if (!(P->betas[b] - P->betas[b - 1] >= 0.0f)) return fail(TAMC_E_INVALID_ARG, "beta array must be non-decreasing");
  if (P->lo_beta_idx > P->num_betas) return fail(TAMC_E_INVALID_ARG, "lo_beta_idx out of range");
  if (P->sweep_order > TAMC_ORDER_SEQUENTIAL) return fail(TAMC_E_INVALID_ARG, "unknown sweep_order");
  if ((P->sample_enabled && P->sample == 0) || (P->sample_states_enabled && P->sample_states == 0))
    return fail(TAMC_E_INVALID_ARG, "sample intervals must be positive");
  if (inst->n == 0) return fail(TAMC_E_INVALID_ARG, "empty instance");
  for (auto &v : inst->chi)
    if (v.size() != inst->n) return fail(TAMC_E_INVALID_ARG, "suscept vector length differs from the instance size");
  S.inst = inst;
  S.P = *P;
  S.betas.assign(P->betas, P->betas + P->num_betas);
  S.P.betas = S.betas.data();
  S.init_states = init_states;
  S.n = inst->n;
  S.B = P->num_betas;
  S.C = P->num_replica_chains;
  S.T = P->num_sweeps;
  const double mi = P->warmup_fraction * (double)P->num_sweeps;  // `as u32` saturates
  S.meas_init = (mi != mi || mi <= 0.0) ? 0u : (mi >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)mi);
  S.n_meas = P->sample_enabled ? count_events(S.T, S.meas_init, P->sample) : 0;
  S.n_samp_events = P->sample_states_enabled ? count_events(S.T, S.meas_init, P->sample_states) : 0;
  S.n_samp_slots = P->sample_limiting == 0 ? S.B : (P->sample_limiting == 1 ? S.B / 2 : 1);
  S.nbytes = (S.n + 7) / 8;
  S.gs_words = (S.n + 63) / 64;
  S.gs_cap = S.T > S.meas_init ? S.T - S.meas_init : 0;
  S.npairs = S.C / 2;
  S.nchi = (uint32_t)inst->chi.size();
  for (uint32_t c0 = 0; c0 < S.C; c0 += 2) {
    GroupPlan g;
    g.chain0 = c0;
    g.nch = std::min(2u, S.C - c0);
    g.R = g.nch * S.B;
    if (g.R > kMaxGroupReplicas)
      return fail(TAMC_E_UNSUPPORTED, "a chain pair may hold at most 1024 replicas (num_betas <= 512)");
    g.G = (g.R + 63) / 64;
    g.P = (P->icm && g.nch == 2 && P->lo_beta_idx < S.B) ? S.B - P->lo_beta_idx : 0;
    // the group's state in LDS when it fits; else the ICM scratch, then also the spin planes, in
    // global memory
    const LdsLayout L = make_layout(S.n, g.G, g.R, S.B, g.P, waves, true);
    if (L.total > kLdsLimit) {
      g.icm_global = true;
      if (make_layout(S.n, g.G, g.R, S.B, g.P, waves, false).total > kLdsLimit) {
        g.spins_global = true;
        if (make_layout(S.n, g.G, g.R, S.B, g.P, waves, false, false).total > kLdsLimit)
          return fail(TAMC_E_UNSUPPORTED, "a chain pair's ladder state does not fit in LDS");
      }
    }
    if (init_states) {
      // bit planes [n][G][2]: bit l of half h of word g <=> replica 64g+32h+l is -1
      g.init_bits.assign((size_t)S.n * g.G * 2, 0u);
      for (uint32_t r = 0; r < g.R; ++r) {
        const uint32_t c = g.chain0 + r / S.B, b = r % S.B;
        const int8_t *s = init_states + ((size_t)c * S.B + b) * S.n;
        for (uint32_t i = 0; i < S.n; ++i)
          if (!(s[i] > 0)) g.init_bits[((size_t)i * g.G + (r >> 6)) * 2 + ((r >> 5) & 1)] |= 1u << (r & 31);
      }
    }
    S.groups.push_back(std::move(g));
  }
  return TAMC_OK;
}

tamc_status upload_resume(const SolvePlan &S) {
  const tamc_pt_state *st = S.resume;
  if (!st) return TAMC_OK;
  const uint32_t B = S.B;
  for (const auto &g : S.groups) {
    const PtGroup &X = g.g;
    const size_t c0 = g.chain0, k0 = c0 * B, nk = (size_t)g.nch * B;
    if (st->tags) TAMC_HIP(hipMemcpy(X.tags, st->tags + k0, nk, hipMemcpyHostToDevice));
    if (st->num_acceptances) TAMC_HIP(hipMemcpy(X.nacc, st->num_acceptances + k0, nk * 4, hipMemcpyHostToDevice));
    if (st->diffusion_hist) TAMC_HIP(hipMemcpy(X.hist, st->diffusion_hist + 2 * k0, nk * 8, hipMemcpyHostToDevice));
    if (st->round_trips) TAMC_HIP(hipMemcpy(X.rt, st->round_trips + c0, (size_t)g.nch * 4, hipMemcpyHostToDevice));
    // the identity permutation holds after init: physical replica r = local chain * B + slot.
    // The init kernel left energy_ref of each state in X.energy; a carried energy that differs from
    // it (QUBO self entries shift the tracked energy, src/ising.rs:308-310) starts as drift, which
    // the next ICM move on the replica clears as the reference's recomputation does.
    if (st->energies) {
      std::vector<float> eref(nk);
      std::vector<long long> drift(nk);
      TAMC_HIP(hipMemcpy(eref.data(), X.energy, nk * 4, hipMemcpyDeviceToHost));
      for (size_t k = 0; k < nk; ++k)
        drift[k] = (long long)std::rint(((double)st->energies[k0 + k] - (double)eref[k]) * 4294967296.0);
      TAMC_HIP(hipMemcpy(X.energy, st->energies + k0, nk * 4, hipMemcpyHostToDevice));
      TAMC_HIP(hipMemcpy(X.drift, drift.data(), nk * 8, hipMemcpyHostToDevice));
    }
  }
  return TAMC_OK;
}

uint32_t lds_bytes_for(const SolvePlan &S, const GroupPlan &g, uint32_t waves) {
  return with_nbr(make_layout(S.n, g.G, g.R, S.B, g.P, waves, !g.icm_global, !g.spins_global), S.n,
                  g.g.ell_col16 ? g.g.ell_D : 0u, kLdsLimit)
      .total;
}

void want_buffers(SolvePlan &S, Arena &A, Arena *Ags, Arena *Ab) {
  (Ab ? *Ab : A).want(&S.d_betas, S.B);
  for (auto &g : S.groups) {
    PtGroup &X = g.g;
    const uint32_t nb = (S.n + 63) / 64;
    A.want(&X.spins, (size_t)S.n * g.G * 2);
    A.want(&X.energy, g.R);
    A.want(&X.drift, g.R);
    A.want(&X.phase_ticks, kPhaseSlots);
    A.want(&X.perm, (size_t)g.nch * S.B);
    A.want(&X.tags, (size_t)g.nch * S.B);
    A.want(&X.nacc, (size_t)g.nch * S.B);
    A.want(&X.hist, (size_t)g.nch * S.B * 2);
    A.want(&X.rt, g.nch);
    A.want(&X.gs_min, 1);
    A.want(&X.gs_count, 1);
    A.want(&X.icm_moves, 1);
    A.want(&X.out_e, (size_t)S.n_meas * g.nch * S.B);
    A.want(&X.out_q, (size_t)S.n_meas * S.B);
    A.want(&X.out_chi, (size_t)S.n_meas * S.B * S.nchi);
    A.want(&X.out_samples, (size_t)S.n_samp_events * S.n_samp_slots * g.nch * S.nbytes);
    A.want(&X.out_gs_steps, S.gs_cap);
    A.want(&X.out_gs_e, S.gs_cap);
    (Ags ? *Ags : A).want(&X.out_gs_states, (size_t)S.gs_cap * S.gs_words);
    if (g.icm_global) {
      A.want(&X.g_qw, S.n);
      A.want(&X.g_cw, S.n);
      A.want(&X.g_blk, (size_t)nb * icm_chunk(g.P));
    }
  }
}

// Default ICM growth order, from timings of both orders on one work group of 16 waves: compact
// tiles for rows wider than four couplings (3D lattices: growth 25-35 % shorter); for rows of four
// (2D lattices), tiles up to 1024 sites and above 4096, strided sites between (four sites per
// thread, whose rows, q and cluster words stay in registers: the 64 x 64 lattice's sweep takes
// 127 us against 191 us with tiles, while from 96 x 96 on, with rows in HBM, tiles take 714 us
// against 1138 us); strided sites on random 3-regular graphs (1.3-2x faster there).
static bool use_icm_tiles(bool have_tiles, uint32_t D, uint32_t n) {
  return have_tiles && (D > 4u || (D == 4u && (n <= 1024u || n > 4096u)));
}

tamc_status fill_groups(SolvePlan &S, hipStream_t s, bool upload_betas) {
  if (upload_betas && s) TAMC_HIP(hipMemcpyAsync(S.d_betas, S.betas.data(), S.B * 4, hipMemcpyHostToDevice, s));
  else if (upload_betas) TAMC_HIP(hipMemcpy(S.d_betas, S.betas.data(), S.B * 4, hipMemcpyHostToDevice));
  for (auto &g : S.groups) {
    tamc_status st;
    DeviceInstance *D = device_instance(S.inst, g.G, &st, S.P.sweep_order);
    if (!D) return st;
    const DeviceSchedule *Sc = device_schedule(S.inst, g.G, &st, S.P.sweep_order);
    if (!Sc) return st;
    PtGroup &X = g.g;
    X.row_ptr = (const uint32_t *)D->row_ptr.p;
    X.col = (const uint32_t *)D->col.p;
    X.J = (const float *)D->J.p;
    X.bias = (const float *)D->bias.p;
    X.g_row_ptr = (const uint32_t *)D->g_row_ptr.p;
    X.g_col = (const uint32_t *)D->g_col.p;
    X.chi = (const double *)D->chi.p;
    X.ell_col = (const uint32_t *)D->ell_col.p;
    X.ell_col16 = S.n < 65536u ? (const uint16_t *)D->ell_col16.p : nullptr;
    X.ell_J = (const float *)D->ell_J.p;
    X.ell_D = D->ell_D;
    X.icm_tiles = X.ell_col16 ? (const uint32_t *)D->icm_tiles.p : nullptr;
    X.icm_ntiles = X.icm_tiles ? D->icm_ntiles : 0u;
    X.lat_L = X.ell_col16 ? D->lat_L : 0u;
    X.lat_dim = D->lat_dim;
    X.icm_trows = X.icm_tiles ? (const uint16_t *)D->icm_trows.p : nullptr;
    X.self_J = D->has_self ? (const float *)D->self_J.p : nullptr;
    X.offset = S.inst->offset;
    X.n = S.n;
    X.g_nodes = S.inst->g_nodes;
    X.nchi = S.nchi;
    X.nblk = (S.n + 63) / 64;
    X.lvl_off = (const uint32_t *)Sc->lvl_off.p;
    X.tasks = (const uint32_t *)Sc->tasks.p;
    X.n_levels = Sc->n_levels;
    X.grp = (const uint32_t *)Sc->groups.p;
    X.records = (const uint32_t *)Sc->records.p;
    X.betas = S.d_betas;
    X.B = S.B;
    X.nch = g.nch;
    X.chain0 = g.chain0;
    X.lo = S.P.lo_beta_idx;
    X.G = g.G;
    X.R = g.R;
    X.P = g.P;
    X.flags = (S.P.icm ? kFlagIcm : 0u) | (S.P.threads > 1 ? kFlagThreaded : 0u) |
              (S.P.sample_enabled ? kFlagSample : 0u) | (S.P.sample_states_enabled ? kFlagSampleStates : 0u) |
              (g.icm_global ? kFlagIcmGlobal : 0u) | (g.spins_global ? kFlagSpinsGlobal : 0u) | (use_icm_tiles(X.icm_tiles != nullptr, X.ell_D, S.n) ? kFlagIcmTiles : 0u);
    X.T = S.T;
    X.meas_init = S.meas_init;
    X.sample = S.P.sample_enabled ? S.P.sample : 1u;
    X.sample_states = S.P.sample_states_enabled ? S.P.sample_states : 1u;
    X.sample_limiting = S.P.sample_limiting;
    X.n_samp_slots = S.n_samp_slots;
    X.gs_cap = S.gs_cap;
    X.nbytes = S.nbytes;
    X.gs_words = S.gs_words;
    X.pair_global = g.chain0 / 2;
    X.seed = S.P.seed;
    X.members = 1;
    X.member = 0;
    X.team_ctr = nullptr;
    X.team_box = nullptr;
  }
  return TAMC_OK;
}

// Allocate device buffers for all groups of all solves and fill their PtGroup descriptors.
static tamc_status allocate(std::vector<SolvePlan> &solves, Arena &A, Arena &Ags, Arena &Ab) {
  for (auto &S : solves) want_buffers(S, A, &Ags, &Ab);
  TAMC_HIP(A.commit());
  TAMC_HIP(Ags.commit());
  TAMC_HIP(Ab.commit());
  // every solve's betas in one upload
  std::vector<char> img(Ab.used, 0);
  for (auto &S : solves) memcpy(img.data() + ((const char *)S.d_betas - (const char *)Ab.buf.p), S.betas.data(), S.B * 4);
  if (Ab.used) TAMC_HIP(hipMemcpy(Ab.buf.p, img.data(), Ab.used, hipMemcpyHostToDevice));
  for (auto &S : solves) {
    const tamc_status st = fill_groups(S, nullptr, false);
    if (st) return st;
  }
  return TAMC_OK;
}

struct Launch {
  std::vector<PtGroup> order;        // groups in launch order (largest first)
  std::vector<const uint32_t *> init_ptrs;
  DevBuf d_groups, d_init_ptrs;
  std::vector<DevBuf *> init_bufs;
  uint32_t lds = 0, icm_lds = 1, max_T = 0;
  uint32_t spins_global = 0;         // some group keeps its spin planes in global memory
  // with teams: the sweep launch's descriptors (one per team member or group), grouped by work
  // group through `jobs` (work group b runs desc[jobs[b]] .. desc[jobs[b + 1] - 1])
  std::vector<PtGroup> desc;
  std::vector<uint32_t> jobs;
  DevBuf d_desc, d_jobs, d_team;
  std::vector<size_t> team_ctr;      // word offset of each team's counters in d_team
  uint32_t teams = 0;
  ~Launch() {
    for (auto *b : init_bufs) delete b;
  }
};

// Teams.  Per-sweep cost of a group in microseconds, from the measured phase split of the
// batched kernel at 16 waves (random 3-regular graph, N = 3976, one plane word: Metropolis
// 27.7 ns per site and word, ICM 12.3 ns per site and chunk of 32 pairs, 8 us of exchange,
// measurement and barriers), scaled for the row width; a team of K pays a meeting (arrival, wait,
// import of the other members' spin words) per colour class and per ICM flip, charged at 4 us: the
// measured cost is 5-6 us, but 4 us plans faster launches of the batched workload (2.25e11
// against 2.21e11 attempts/s at 5.5 us), as the model's other terms overstate small groups.
static double sweep_cost_us(const SolvePlan &S, const GroupPlan &g, uint32_t K) {
  const double n = (double)S.n, D = (double)std::max<uint32_t>(3u, (uint32_t)((S.inst->row_ptr[S.n] + S.n - 1) / S.n));
  const double metro = 0.0277 * n * g.G * (1.0 + 0.15 * (D - 3.0));
  // ICM: overlap, seeds and growth (57 %) run in every member, the flip (43 %) is shared with one
  // more meeting per chunk
  const double chunks = g.P ? (double)((g.P + kMaxPairs - 1) / kMaxPairs) : 0.0;
  const double icm = 0.0123 * n * chunks * (0.57 + 0.43 / K);
  const double levels = (double)S.inst->sched[S.P.sweep_order].n_classes;
  return metro / K + icm + 8.0 + (K > 1 ? 4.0 * (levels + chunks) : 0.0);
}

// Teams need a group whose spin planes are in LDS (every member holds the whole state; a group
// whose ICM scratch is in global memory gives each member a private copy of it), no QUBO self
// entries (their drift is per flip) and at least two colour classes: a member rewrites a class's
// exchange words only after the next meeting, which every member reaches after reading them.
static bool team_eligible(const SolvePlan &S, const GroupPlan &g) {
  return !g.spins_global && !(g.g.flags & kFlagIcmOnly) && g.g.self_J == nullptr && S.n >= 64u &&
         S.inst->sched[S.P.sweep_order].n_classes >= 2u;
}

// 32-bit words of a team member's private ICM scratch: q words, cluster words, per-block counts,
// then the parking area of the spin planes
static size_t icm_scratch_words(const SolvePlan &S, const GroupPlan &g) {
  return 2u * (size_t)S.n + (size_t)((S.n + 63u) / 64u) * icm_chunk(g.P) + (size_t)S.n * g.G * 2u;
}

struct TeamJob {
  uint32_t item, member, members;
};

// Assign groups to work groups: each group runs alone, or as a team of K members in K work groups,
// and groups that run alone may follow one another in a work group.  For a target makespan M,
// every group takes the smallest K whose cost fits M; team members open work groups, and the
// single groups go, longest first, to the least loaded work group.  The smallest feasible M is
// found by bisection.  force_k > 1 makes every eligible group a team of force_k.  Returns the
// predicted makespan (us), or 0 when no plan fits the capacity.
static double plan_teams(const std::vector<double> &cost1, const std::vector<std::vector<double>> &costk,
                         const std::vector<uint32_t> &kvals, const std::vector<bool> &eligible, uint32_t capacity,
                         uint32_t force_k, std::vector<std::vector<TeamJob>> &out) {
  const size_t ni = cost1.size();
  auto build = [&](double M, std::vector<std::vector<TeamJob>> *wg) -> double {
    std::vector<uint32_t> K(ni, 1u);
    uint32_t used = 0;
    for (size_t i = 0; i < ni; ++i) {
      if (!eligible[i]) continue;
      if (force_k > 1u) {
        K[i] = force_k;
      } else if (cost1[i] > M) {
        for (size_t q = 0; q < kvals.size(); ++q)
          if (costk[i][q] <= M) {
            K[i] = kvals[q];
            break;
          }
      }
      if (K[i] > 1u) used += K[i];
    }
    for (size_t i = 0; i < ni; ++i)
      if (K[i] == 1u && cost1[i] > M && force_k <= 1u) return 0.0;
    if (used > capacity) return 0.0;
    std::vector<double> load;
    std::vector<std::vector<TeamJob>> bins;
    for (size_t i = 0; i < ni; ++i)
      if (K[i] > 1u) {
        const size_t q = std::find(kvals.begin(), kvals.end(), K[i]) - kvals.begin();
        const double c = q < kvals.size() ? costk[i][q] : cost1[i];
        for (uint32_t m = 0; m < K[i]; ++m) {
          load.push_back(c);
          bins.push_back({TeamJob{(uint32_t)i, m, K[i]}});
        }
      }
    while (bins.size() < capacity) {
      load.push_back(0.0);
      bins.emplace_back();
    }
    std::vector<size_t> singles;
    for (size_t i = 0; i < ni; ++i)
      if (K[i] == 1u) singles.push_back(i);
    std::sort(singles.begin(), singles.end(), [&](size_t a, size_t b) { return cost1[a] > cost1[b]; });
    for (size_t i : singles) {
      const size_t b = std::min_element(load.begin(), load.end()) - load.begin();
      load[b] += cost1[i];
      bins[b].push_back(TeamJob{(uint32_t)i, 0u, 1u});
    }
    double mk = 0.0;
    for (double l : load) mk = std::max(mk, l);
    if (wg) {
      wg->clear();
      for (auto &b : bins)
        if (!b.empty()) wg->push_back(b);
    }
    return mk;
  };
  if (force_k > 1u) return build(0.0, &out);
  double hi = 0.0, lo = 0.0;
  for (size_t i = 0; i < ni; ++i) {
    hi = std::max(hi, cost1[i]);
    double best = cost1[i];
    if (eligible[i])
      for (double c : costk[i]) best = std::min(best, c);
    lo = std::max(lo, best);
  }
  double best_mk = 0.0;
  for (int it = 0; it < 40 && hi - lo > 0.01; ++it) {
    const double M = 0.5 * (lo + hi);
    const double mk = build(M, nullptr);
    if (mk > 0.0 && mk <= M * 1.0001) {
      hi = M;
    } else {
      lo = M;
    }
  }
  best_mk = build(hi, &out);
  return best_mk;
}


// teams: 0 = none, 1 = where the cost model predicts a makespan at least 10 % shorter, K >= 2 =
// every eligible group as a team of K work groups
static tamc_status prepare_launch(std::vector<SolvePlan> &solves, uint32_t waves, uint32_t tiling, uint32_t growth,
                                  uint32_t teams, Launch &Lc) {
  struct Item {
    SolvePlan *S;
    GroupPlan *g;
    double cost;
  };
  std::vector<Item> items;
  bool any_init = false;
  for (auto &S : solves)
    for (auto &g : S.groups) {
      items.push_back({&S, &g, (double)S.n * g.R * S.T});
      any_init |= !g.init_bits.empty();
      Lc.lds = std::max(Lc.lds, lds_bytes_for(S, g, waves));
      Lc.spins_global |= g.spins_global ? 1u : 0u;
      Lc.max_T = std::max(Lc.max_T, S.T);
    }
  std::stable_sort(items.begin(), items.end(), [](const Item &a, const Item &b) { return a.cost > b.cost; });
  for (auto &it : items) {
    Lc.order.push_back(it.g->g);
    if (tiling == 2u) Lc.order.back().flags |= kFlagTileBlocks;
    if (growth == 1u) Lc.order.back().flags &= ~kFlagIcmTiles;
    if (growth == 2u && Lc.order.back().icm_tiles) Lc.order.back().flags |= kFlagIcmTiles;
    if (any_init) {
      if (it.g->init_bits.empty()) {
        Lc.init_ptrs.push_back(nullptr);
      } else {
        DevBuf *b = new DevBuf();
        Lc.init_bufs.push_back(b);
        TAMC_HIP(b->alloc(it.g->init_bits.size() * 4));
        TAMC_HIP(hipMemcpy(b->p, it.g->init_bits.data(), it.g->init_bits.size() * 4, hipMemcpyHostToDevice));
        Lc.init_ptrs.push_back((const uint32_t *)b->p);
      }
    }
  }
  TAMC_HIP(Lc.d_groups.alloc(Lc.order.size() * sizeof(PtGroup)));
  TAMC_HIP(hipMemcpy(Lc.d_groups.p, Lc.order.data(), Lc.order.size() * sizeof(PtGroup), hipMemcpyHostToDevice));
  if (any_init) {
    TAMC_HIP(Lc.d_init_ptrs.alloc(Lc.init_ptrs.size() * sizeof(void *)));
    TAMC_HIP(hipMemcpy(Lc.d_init_ptrs.p, Lc.init_ptrs.data(), Lc.init_ptrs.size() * sizeof(void *),
                       hipMemcpyHostToDevice));
  }
  if (!teams) return TAMC_OK;
  uint32_t capacity = 0;
  if (tamc_pt_sweep_capacity(waves, Lc.lds, &capacity) != hipSuccess || capacity < 2u || items.size() > capacity) {
    (void)hipGetLastError();
    return TAMC_OK;  // no room for teams: every group runs alone
  }
  std::vector<uint32_t> kvals;
  for (uint32_t k : {2u, 3u, 4u, 5u, 6u, 8u, 10u, 12u, 16u})
    if (k <= kMaxMembers && k <= capacity) kvals.push_back(k);
  const size_t ni = items.size();
  std::vector<double> cost1(ni);
  std::vector<std::vector<double>> costk(ni);
  std::vector<bool> eligible(ni);
  bool any = false;
  for (size_t i = 0; i < ni; ++i) {
    const SolvePlan &S = *items[i].S;
    const GroupPlan &g = *items[i].g;
    cost1[i] = sweep_cost_us(S, g, 1u) * S.T;
    for (uint32_t k : kvals) costk[i].push_back(sweep_cost_us(S, g, k) * S.T);
    eligible[i] = team_eligible(S, g);
    any |= eligible[i];
  }
  if (!any) return TAMC_OK;
  std::vector<std::vector<TeamJob>> wgs;
  const uint32_t force = teams >= 2u ? std::min(teams, std::min(kMaxMembers, capacity)) : 0u;
  const double mk = plan_teams(cost1, costk, kvals, eligible, capacity, force, wgs);
  double base = 0.0;
  for (double c : cost1) base = std::max(base, c);
  bool used = false;
  for (auto &w : wgs)
    for (auto &j : w) used |= j.members > 1u;
  if (!used || mk <= 0.0 || (!force && mk > 0.9 * base) || wgs.size() > capacity) return TAMC_OK;
  // team scratch: counters (64-byte aligned) and the spin-word / energy-sum box of each team
  std::vector<size_t> ctr_off(ni, 0), box_off(ni, 0), scr_off(ni, 0);
  std::vector<uint32_t> members(ni, 1u);
  for (auto &w : wgs)
    for (auto &j : w) members[j.item] = j.members;
  size_t words = 0;  // in 8-byte units
  for (size_t i = 0; i < ni; ++i) {
    if (members[i] < 2u) continue;
    ctr_off[i] = words;
    words += (kTeamCtrWords * 4u + 63u) / 64u * 8u;
    box_off[i] = words;
    // level words, energy sums, then two flip buffers of block words and per-pair sums
    words += team_box_slots(items[i].S->n, items[i].g->G, items[i].g->R, members[i]);
    words = (words + 7u) & ~(size_t)7u;
    if (items[i].g->icm_global || Lc.order[i].lat_L != 0u) {
      // members 1 .. K-1 grow their clusters in private copies of the ICM scratch (member 0 uses the
      // group's own): members run the unshared ICM steps at different times, and one member's
      // clearing of the cluster words for the next sweep must not meet another's seeds; and they
      // park their spin planes (HBM scratch, or lattice growth) in private areas
      scr_off[i] = words;
      words += ((members[i] - 1u) * icm_scratch_words(*items[i].S, *items[i].g) + 15u) / 16u * 8u;
    }
    ++Lc.teams;
  }
  TAMC_HIP(Lc.d_team.alloc(words * 8u));
  TAMC_HIP(hipMemset(Lc.d_team.p, 0, words * 8u));
  unsigned long long *tb = (unsigned long long *)Lc.d_team.p;
  Lc.jobs.push_back(0u);
  for (auto &w : wgs) {
    for (auto &j : w) {
      PtGroup d = Lc.order[j.item];
      if (j.members > 1u) {
        d.members = j.members;
        d.member = j.member;
        d.team_ctr = (uint32_t *)(tb + ctr_off[j.item]);
        d.team_box = tb + box_off[j.item];
        if (j.member > 0u && (items[j.item].g->icm_global || Lc.order[j.item].lat_L != 0u)) {
          uint32_t *scr = (uint32_t *)(tb + scr_off[j.item]) +
                          (size_t)(j.member - 1u) * icm_scratch_words(*items[j.item].S, *items[j.item].g);
          d.g_qw = scr;
          d.g_cw = scr + items[j.item].S->n;
          d.g_blk = scr + 2u * (size_t)items[j.item].S->n;
          d.g_park = d.g_blk + (size_t)((items[j.item].S->n + 63u) / 64u) * icm_chunk(items[j.item].g->P);
        }
        if (j.member == 0u) Lc.team_ctr.push_back(ctr_off[j.item] * 2u);
      }
      Lc.desc.push_back(d);
    }
    Lc.jobs.push_back((uint32_t)Lc.desc.size());
  }
  TAMC_HIP(Lc.d_desc.alloc(Lc.desc.size() * sizeof(PtGroup)));
  TAMC_HIP(hipMemcpy(Lc.d_desc.p, Lc.desc.data(), Lc.desc.size() * sizeof(PtGroup), hipMemcpyHostToDevice));
  TAMC_HIP(Lc.d_jobs.alloc(Lc.jobs.size() * 4u));
  TAMC_HIP(hipMemcpy(Lc.d_jobs.p, Lc.jobs.data(), Lc.jobs.size() * 4u, hipMemcpyHostToDevice));
  return TAMC_OK;
}

// Device -> host for result arrays: from the host mirror when the range lies inside it (one bulk
// copy of the arena instead of one small copy per array), else a copy of its own.
template <class T>
static tamc_status d2h(std::vector<T> &v, const T *d, size_t count, const HostMirror *m) {
  v.resize(count);
  if (!count) return TAMC_OK;
  const char *dp = (const char *)d;
  if (m && dp >= m->dev && dp + count * sizeof(T) <= m->dev + m->bytes) {
    memcpy(v.data(), m->host + (dp - m->dev), count * sizeof(T));
    return TAMC_OK;
  }
  TAMC_HIP(hipMemcpy(v.data(), d, count * sizeof(T), hipMemcpyDeviceToHost));
  return TAMC_OK;
}

tamc_status assemble(const SolvePlan &S, tamc_pt_result **out, const HostMirror *m) {
  const uint32_t n = S.n, B = S.B, C = S.C, nm = S.n_meas, np = S.npairs;
  tamc_pt_result *R = (tamc_pt_result *)calloc(1, sizeof(tamc_pt_result));
  R->n = n;
  R->B = B;
  R->C = C;
  R->T = S.T;
  R->meas_init = S.meas_init;
  R->nchi = S.nchi;
  R->npairs = np;
  R->n_meas = nm;
  R->n_samp_slots = S.n_samp_slots;
  R->n_samp_events = S.n_samp_events;
  R->nbytes = S.nbytes;
  R->gs_words = S.gs_words;
  std::vector<float> e((size_t)B * nm * C), chi((size_t)B * S.nchi * nm * np);
  std::vector<int32_t> q((size_t)B * nm * np);
  std::vector<uint8_t> smp((size_t)S.n_samp_slots * S.n_samp_events * C * S.nbytes);
  std::vector<int8_t> fin((size_t)C * B * n);
  std::vector<float> fe((size_t)C * B);
  std::vector<uint32_t> ids((size_t)C * B), nacc((size_t)C * B), hist((size_t)C * B * 2), rts(C), accs(B, 0);
  std::vector<uint8_t> tags((size_t)C * B);
  struct GsRec {
    uint32_t step;
    float e;
    uint32_t group;
    std::vector<uint64_t> st;
  };
  std::vector<std::vector<GsRec>> gsr(S.groups.size());
  uint64_t moves = 0;
  tamc_status st;
  for (size_t gi = 0; gi < S.groups.size(); ++gi) {
    const GroupPlan &g = S.groups[gi];
    const PtGroup &X = g.g;
    std::vector<float> oe, ochi, energy, gsmin, gse;
    std::vector<int32_t> oq;
    std::vector<uint8_t> os, otags;
    std::vector<uint32_t> spins, perm, onacc, ohist, ort, gsc, gss;
    std::vector<uint64_t> gsst;
    std::vector<unsigned long long> mv;
    if ((st = d2h(oe, X.out_e, (size_t)nm * g.nch * B, m)) || (st = d2h(oq, X.out_q, (size_t)nm * B, m)) ||
        (st = d2h(ochi, X.out_chi, (size_t)nm * B * S.nchi, m)) ||
        (st = d2h(os, X.out_samples, (size_t)S.n_samp_events * S.n_samp_slots * g.nch * S.nbytes, m)) ||
        (st = d2h(spins, X.spins, (size_t)n * g.G * 2, m)) || (st = d2h(energy, X.energy, g.R, m)) ||
        (st = d2h(perm, X.perm, (size_t)g.nch * B, m)) || (st = d2h(otags, X.tags, (size_t)g.nch * B, m)) ||
        (st = d2h(onacc, X.nacc, (size_t)g.nch * B, m)) || (st = d2h(ohist, X.hist, (size_t)g.nch * B * 2, m)) ||
        (st = d2h(ort, X.rt, g.nch, m)) || (st = d2h(gsc, X.gs_count, 1, m)) || (st = d2h(mv, X.icm_moves, 1, m))) {
      free(R);
      return st;
    }
    moves += mv[0];
    const uint32_t ngs = std::min(gsc[0], S.gs_cap);
    if ((st = d2h(gss, X.out_gs_steps, ngs, m)) || (st = d2h(gse, X.out_gs_e, ngs, m)) ||
        (st = d2h(gsst, X.out_gs_states, (size_t)ngs * S.gs_words, m))) {
      free(R);
      return st;
    }
    for (uint32_t k = 0; k < ngs; ++k)
      gsr[gi].push_back(GsRec{gss[k], gse[k], (uint32_t)gi,
                              std::vector<uint64_t>(gsst.begin() + (size_t)k * S.gs_words,
                                                    gsst.begin() + (size_t)(k + 1) * S.gs_words)});
    for (uint32_t ev = 0; ev < nm; ++ev)
      for (uint32_t cl = 0; cl < g.nch; ++cl)
        for (uint32_t b = 0; b < B; ++b)
          e[(size_t)b * nm * C + (size_t)ev * C + g.chain0 + cl] = oe[((size_t)ev * g.nch + cl) * B + b];
    if (g.nch == 2) {
      const uint32_t p = g.chain0 / 2;
      for (uint32_t ev = 0; ev < nm; ++ev)
        for (uint32_t b = 0; b < B; ++b) {
          q[(size_t)b * nm * np + (size_t)ev * np + p] = oq[(size_t)ev * B + b];
          for (uint32_t k = 0; k < S.nchi; ++k)
            chi[((size_t)b * S.nchi + k) * nm * np + (size_t)ev * np + p] = ochi[((size_t)ev * B + b) * S.nchi + k];
        }
    }
    for (uint32_t sev = 0; sev < S.n_samp_events; ++sev)
      for (uint32_t slot = 0; slot < S.n_samp_slots; ++slot)
        for (uint32_t cl = 0; cl < g.nch; ++cl)
          memcpy(&smp[(((size_t)slot * S.n_samp_events * C) + (size_t)sev * C + g.chain0 + cl) * S.nbytes],
                 &os[(((size_t)sev * S.n_samp_slots + slot) * g.nch + cl) * S.nbytes], S.nbytes);
    for (uint32_t cl = 0; cl < g.nch; ++cl) {
      const uint32_t c = g.chain0 + cl;
      for (uint32_t b = 0; b < B; ++b) {
        const uint32_t r = perm[cl * B + b];
        int8_t *out_s = &fin[((size_t)c * B + b) * n];
        if (g.G == 1u) {
          uint64_t w;
          for (uint32_t i = 0; i < n; ++i) {
            memcpy(&w, &spins[(size_t)i * 2], 8);
            out_s[i] = (int8_t)(1 - 2 * (int)((w >> r) & 1u));
          }
        } else {
          for (uint32_t i = 0; i < n; ++i) {
            const uint32_t bit = (spins[((size_t)i * g.G + (r >> 6)) * 2 + ((r >> 5) & 1)] >> (r & 31)) & 1u;
            out_s[i] = bit ? -1 : 1;
          }
        }
        fe[(size_t)c * B + b] = energy[r];
        ids[(size_t)c * B + b] = g.chain0 * B + r;
        nacc[(size_t)c * B + b] = onacc[cl * B + b];
        accs[b] += onacc[cl * B + b];
        tags[(size_t)c * B + b] = otags[cl * B + b];
        hist[((size_t)c * B + b) * 2] = ohist[(cl * B + b) * 2];
        hist[((size_t)c * B + b) * 2 + 1] = ohist[(cl * B + b) * 2 + 1];
      }
      rts[c] = ort[cl];
    }
  }
  // Merge per-group ground-state records into the global trace (src/pt.rs:559-581): a global
  // improvement is always an improvement of the group holding the minimum, and ties go to the
  // lowest chain, i.e. the lowest group.
  std::vector<uint32_t> gsteps;
  std::vector<float> genergies;
  std::vector<uint64_t> gstates;
  {
    std::vector<size_t> pos(gsr.size(), 0);
    bool have = false;
    float cur = 0.f;
    for (;;) {
      uint32_t tmin = 0xFFFFFFFFu;
      for (size_t gi = 0; gi < gsr.size(); ++gi)
        if (pos[gi] < gsr[gi].size()) tmin = std::min(tmin, gsr[gi][pos[gi]].step);
      if (tmin == 0xFFFFFFFFu) break;
      const GsRec *best = nullptr;
      for (size_t gi = 0; gi < gsr.size(); ++gi)
        if (pos[gi] < gsr[gi].size() && gsr[gi][pos[gi]].step == tmin) {
          const GsRec &r = gsr[gi][pos[gi]];
          if (!best || r.e < best->e) best = &r;
          ++pos[gi];
        }
      if (!have || best->e < cur) {
        have = true;
        cur = best->e;
        gsteps.push_back(best->step);
        genergies.push_back(best->e);
        gstates.insert(gstates.end(), best->st.begin(), best->st.end());
      }
    }
  }
  R->n_gs = (uint32_t)gsteps.size();
  R->e = host_dup(e);
  R->q = host_dup(q);
  R->suscept = host_dup(chi);
  R->samples = host_dup(smp);
  R->gs_steps = host_dup(gsteps);
  R->gs_energies = host_dup(genergies);
  R->gs_states = host_dup(gstates);
  R->acceptance_counts = host_dup(accs);
  R->final_states = host_dup(fin);
  R->final_energies = host_dup(fe);
  R->final_ids = host_dup(ids);
  R->num_acceptances = host_dup(nacc);
  R->round_trips = host_dup(rts);
  R->tags = host_dup(tags);
  R->diffusion_hist = host_dup(hist);
  R->icm_moves = moves;
  R->attempts = (uint64_t)S.T * n * B * C;
  *out = R;
  return TAMC_OK;
}

struct RunOptions {
  uint32_t waves = 8;
  uint32_t sweeps_per_launch = 0;  // 0 = all
  uint32_t tiling = 1;             // 1 tasks dealt to the waves in turn, 2 contiguous runs
  uint32_t growth = 0;             // ICM cluster growth: 1 strided sites per thread, 2 compact tiles,
                                   // 0 each group's default (tiles for rows wider than 4 entries)
  uint32_t teams = 1;              // 0 none, 1 where the cost model gains, K >= 2 forced teams of K
  bool use_graph = false;
  std::chrono::steady_clock::time_point t_entry = std::chrono::steady_clock::now();  // for prep_ms
};

// Plan, allocate, launch and assemble a set of solves that run concurrently in one kernel.
static tamc_status run_solves(std::vector<SolvePlan> &solves, const RunOptions &opt, tamc_pt_result **results,
                              tamc_batch_stats *stats, const std::function<void(std::vector<SolvePlan> &)> &patch = {}) {
  Arena A, Ags, Ab;  // Ags: ground-state trace buffers, large and mostly unused, downloaded by prefix
  tamc_status st = allocate(solves, A, Ags, Ab);
  if (st) return st;
  if (patch) patch(solves);
  Launch Lc;
  if ((st = prepare_launch(solves, opt.waves, opt.tiling, opt.growth, opt.teams, Lc))) return st;
  const uint32_t ng = (uint32_t)Lc.order.size();
  hipStream_t s;
  TAMC_HIP(hipStreamCreate(&s));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } sg{s};
  const PtGroup *dg = (const PtGroup *)Lc.d_groups.p;
  TAMC_HIP(tamc_launch_pt_init(dg, ng, Lc.init_ptrs.empty() ? nullptr : (const uint32_t *const *)Lc.d_init_ptrs.p, s));
  TAMC_HIP(hipStreamSynchronize(s));
  for (const auto &S : solves)
    if ((st = upload_resume(S))) return st;
  hipEvent_t e0, e1;
  TAMC_HIP(hipEventCreate(&e0));
  TAMC_HIP(hipEventCreate(&e1));
  const uint32_t spl = opt.sweeps_per_launch ? opt.sweeps_per_launch : std::max(1u, Lc.max_T);
  uint32_t launches = 0;
  const auto w0 = std::chrono::steady_clock::now();
  hipGraphExec_t gexec = nullptr;
  hipGraph_t hgraph = nullptr;
  TAMC_HIP(hipEventRecord(e0, s));
  const bool graph = opt.use_graph;
  // with teams, the work groups run the descriptors their job lists name
  const PtGroup *lg = Lc.teams ? (const PtGroup *)Lc.d_desc.p : dg;
  const uint32_t lng = Lc.teams ? (uint32_t)Lc.jobs.size() - 1u : ng;
  const uint32_t *ljobs = Lc.teams ? (const uint32_t *)Lc.d_jobs.p : nullptr;
  if (graph) {
    TAMC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (uint32_t t0 = 0; t0 < Lc.max_T; t0 += spl, ++launches)
      (void)tamc_launch_pt_sweeps(lg, lng, t0, t0 + spl, opt.waves, Lc.lds, Lc.icm_lds, nullptr, nullptr, s, ljobs, Lc.spins_global);
    TAMC_HIP(hipStreamEndCapture(s, &hgraph));
    TAMC_HIP(hipGraphInstantiate(&gexec, hgraph, nullptr, nullptr, 0));
    TAMC_HIP(hipGraphLaunch(gexec, s));
  } else {
    for (uint32_t t0 = 0; t0 < Lc.max_T; t0 += spl, ++launches)
      TAMC_HIP(tamc_launch_pt_sweeps(lg, lng, t0, t0 + spl, opt.waves, Lc.lds, Lc.icm_lds, nullptr, nullptr, s, ljobs, Lc.spins_global));
  }
  TAMC_HIP(hipEventRecord(e1, s));
  TAMC_HIP(hipEventSynchronize(e1));
  const auto w1 = std::chrono::steady_clock::now();
  float ms = 0.f;
  TAMC_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (hgraph) (void)hipGraphDestroy(hgraph);
  for (size_t off : Lc.team_ctr) {
    uint32_t ab = 0;
    TAMC_HIP(hipMemcpy(&ab, (const uint32_t *)Lc.d_team.p + off + 1u, 4, hipMemcpyDeviceToHost));
    if (ab) return fail(TAMC_E_DEVICE, "a team of work groups gave up waiting for its members");
  }
  const double wall_us = std::chrono::duration<double, std::micro>(w1 - w0).count();
  uint64_t attempts = 0;
  std::vector<char> mirror_buf(A.used);
  if (A.used) TAMC_HIP(hipMemcpy(mirror_buf.data(), A.buf.p, A.used, hipMemcpyDeviceToHost));
  const HostMirror mirror{mirror_buf.data(), (const char *)A.buf.p, A.used};
  // results are independent: assemble them on several host threads (errors are thread-local, so
  // each worker keeps its message for this thread to report)
  const size_t ns = solves.size();
  std::vector<tamc_status> sts(ns, TAMC_OK);
  std::vector<std::string> msgs(ns);
  for (size_t k = 0; k < ns; ++k) results[k] = nullptr;
  std::atomic<size_t> next{0};
  auto work = [&]() {
    for (size_t k; (k = next.fetch_add(1)) < ns;) {
      sts[k] = assemble(solves[k], &results[k], &mirror);
      if (sts[k]) msgs[k] = tamc_last_error();
    }
  };
  const size_t nt = std::min<size_t>(ns, std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
  std::vector<std::thread> pool;
  for (size_t t = 1; t < nt; ++t) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
  for (size_t k = 0; k < ns; ++k)
    if (sts[k]) {
      for (size_t j = 0; j < ns; ++j) {
        tamc_pt_result_free(results[j]);
        results[j] = nullptr;
      }
      return fail(sts[k], msgs[k]);
    }
  for (size_t k = 0; k < ns; ++k) {
    results[k]->timing_us = wall_us;
    results[k]->device_ms = ms;
    attempts += results[k]->attempts;
  }
  if (stats) {
    // phase split of the group that took longest
    unsigned long long best[kPhaseSlots] = {0, 0, 0, 0, 0, 0, 0, 0}, best_tot = 0, all_tot = 0;
    for (auto &S : solves)
      for (auto &g : S.groups) {
        unsigned long long tk[kPhaseSlots];
        const char *src = (const char *)g.g.phase_ticks;
        if (src < mirror.dev || src + sizeof tk > mirror.dev + mirror.bytes) continue;
        memcpy(tk, mirror.host + (src - mirror.dev), sizeof tk);
        unsigned long long tot = 0;
        for (uint32_t k = 0; k < kPhIcmPasses; ++k) tot += tk[k];
        all_tot += tot;
        if (tot > best_tot) {
          best_tot = tot;
          memcpy(best, tk, sizeof tk);
        }
      }
    for (uint32_t k = 0; k < kPhIcmPasses; ++k) stats->phase_ms[k] = (double)best[k] * 1e-5;  // 100 MHz ticks
    stats->icm_passes = (double)best[kPhIcmPasses];
    stats->groups_busy_ms = (double)all_tot * 1e-5;
    stats->device_ms = ms;
    stats->wall_ms = wall_us / 1000.0;
    stats->attempts = attempts;
    stats->waves_per_group = opt.waves;
    stats->sweeps_per_launch = spl;
    stats->launches = launches;
    stats->groups = ng;
    stats->task_tiling = opt.tiling;
    stats->icm_growth = opt.growth;
    stats->graph = graph ? 1 : 0;
    stats->teams = Lc.teams;
    stats->work_groups = Lc.teams ? (uint32_t)Lc.jobs.size() - 1u : ng;
    stats->prep_ms = std::chrono::duration<double, std::milli>(w0 - opt.t_entry).count();
    stats->post_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w1).count();
  }
  return TAMC_OK;
}

uint32_t default_waves(const tamc_instance *inst) {
  // enough waves to cover the typical level width, within one work group
  const SiteSchedule &H = inst->sched[kOrderColour];
  const uint32_t nq = H.n_groups;
  const uint32_t width = H.n_classes ? (nq + H.n_classes - 1) / H.n_classes : 1;
  uint32_t w = 4;
  while (w < width && w < 16) w *= 2;
  return w;
}

}  // namespace tamc

using namespace tamc;

extern "C" {

tamc_status tamc_pt_resolve(int32_t geometric, float beta_min, float beta_max, uint32_t num_beta, const float *arr,
                            uint32_t narr, float lo_beta, int64_t lo_num_beta, int32_t icm, float *betas_out,
                            uint32_t *num_betas_out, uint32_t *lo_beta_idx_out) {
  clear_error();
  if (!betas_out || !num_betas_out || !lo_beta_idx_out) return fail(TAMC_E_INVALID_ARG, "null argument");
  uint32_t nb;
  if (geometric) {
    // geometric_beta_schedule (tamc-core/src/sa.rs:13-22) in f64, then `as f32`
    nb = num_beta;
    const double l0 = std::log((double)beta_min), lf = std::log((double)beta_max);
    const double d = (lf - l0) / (double)(nb - 1);
    for (uint32_t i = 0; i < nb; ++i) betas_out[i] = (float)std::exp(l0 + d * (double)i);
  } else {
    nb = narr;
    if (narr && !arr) return fail(TAMC_E_INVALID_ARG, "null beta array");
    memcpy(betas_out, arr, narr * sizeof(float));
  }
  if (nb == 0) return fail(TAMC_E_INVALID_ARG, "empty beta array");
  for (uint32_t b = 1; b < nb; ++b)
    if (!(betas_out[b] - betas_out[b - 1] >= 0.0f)) return fail(TAMC_E_INVALID_ARG, "beta array must be non-decreasing");
  uint32_t lo;
  if (lo_num_beta >= 0) {
    const uint64_t ln = std::min<uint64_t>((uint64_t)lo_num_beta, nb);
    lo = nb - (uint32_t)ln;
  } else if (lo_beta == lo_beta) {
    lo = nb - 1;  // out of bounds: the largest beta is assigned (src/pt.rs:328-331)
    for (uint32_t b = 0; b < nb; ++b)
      if (betas_out[b] >= lo_beta) {
        lo = b;
        break;
      }
  } else {
    lo = 0;
    if (icm) return fail(TAMC_E_INVALID_ARG, "Must specify either lo_beta or lo_num_beta for ICM");
  }
  *num_betas_out = nb;
  *lo_beta_idx_out = lo;
  return TAMC_OK;
}

tamc_status tamc_pt_run(const tamc_instance *inst, const tamc_pt_params *params, const int8_t *init_states,
                        tamc_pt_result **out) {
  clear_error();
  if (!out) return fail(TAMC_E_INVALID_ARG, "null result pointer");
  tamc_status st = device_init();
  if (st) return st;
  std::vector<SolvePlan> solves(1);
  RunOptions opt;
  opt.waves = inst ? default_waves(inst) : 8;
  if ((st = plan_solve(solves[0], inst, params, init_states, opt.waves))) return st;
  return run_solves(solves, opt, out, nullptr);
}

tamc_status tamc_pt_run_resume(const tamc_instance *inst, const tamc_pt_params *params, const tamc_pt_state *state,
                               tamc_pt_result **out) {
  clear_error();
  if (!out || !state || !state->states) return fail(TAMC_E_INVALID_ARG, "null argument");
  if (params && state->tags)
    for (size_t k = 0; k < (size_t)params->num_replica_chains * params->num_betas; ++k)
      if (state->tags[k] > 4u) return fail(TAMC_E_INVALID_ARG, "round-trip tag ordinal out of range (0..4)");
  tamc_status st = device_init();
  if (st) return st;
  std::vector<SolvePlan> solves(1);
  RunOptions opt;
  opt.waves = inst ? default_waves(inst) : 8;
  if ((st = plan_solve(solves[0], inst, params, state->states, opt.waves))) return st;
  solves[0].resume = state;
  return run_solves(solves, opt, out, nullptr);
}

// Autotuner for the batched launch.  Waves per work group and the task tiling of every colour
// level are chosen together by timing each candidate on a short prefix of the same batch
// (measurements off, results discarded); then, on the winner, one launch for the whole prefix is
// timed against four launches of a quarter each captured in a hipGraph, and the faster choice is
// scaled to the full run (all sweeps in one launch, or a quarter per launch).  Choices are cached
// per batch shape for the life of the process.
namespace {
struct TuneKey {
  uint32_t count, max_n, B, C, order, fixed_teams;
  uint64_t sum_n;
  bool operator<(const TuneKey &o) const {
    return std::tie(count, max_n, B, C, order, fixed_teams, sum_n) <
           std::tie(o.count, o.max_n, o.B, o.C, o.order, o.fixed_teams, o.sum_n);
  }
};
struct TuneChoice {
  uint32_t waves = 16, tiling = 1, growth = 0, teams = 0;
  bool quarters = false;  // a quarter of the sweeps per launch, launches captured in a hipGraph
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneChoice> g_tune_cache;
constexpr uint32_t kTunePrefix = 24u;
}  // namespace

// fixed_teams: the RunOptions value the caller fixed (0 none, K >= 2 forced), or kTeamsTuned
constexpr uint32_t kTeamsTuned = 0xFFFFFFFFu;

static tamc_status autotune_batch(uint32_t count, const tamc_instance *const *insts, const tamc_pt_params *params,
                                  uint32_t fixed_waves, uint32_t fixed_tiling, uint32_t fixed_growth,
                                  uint32_t fixed_teams, TuneChoice *out) {
  TuneKey key{count, 0, 0, 0, 0, fixed_teams, 0};
  for (uint32_t k = 0; k < count; ++k) {
    key.order |= 1u << params[k].sweep_order;
    key.max_n = std::max(key.max_n, insts[k]->n);
    key.sum_n += insts[k]->n;
    key.B = std::max(key.B, params[k].num_betas);
    key.C = std::max(key.C, params[k].num_replica_chains);
  }
  key.order |= (fixed_waves << 8) | (fixed_tiling << 16) | (fixed_growth << 24);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      *out = it->second;
      return TAMC_OK;
    }
  }
  std::vector<tamc_pt_params> probe(params, params + count);
  for (auto &p : probe) {
    p.num_sweeps = std::min(p.num_sweeps, kTunePrefix);
    p.warmup_fraction = 1.0;
    p.sample_enabled = 0;
    p.sample_states_enabled = 0;
  }
  std::vector<tamc_pt_result *> tmp(count, nullptr);
  auto time_one = [&](const RunOptions &opt, double *ms) -> tamc_status {
    std::vector<SolvePlan> solves(count);
    for (uint32_t k = 0; k < count; ++k) {
      const tamc_status st = plan_solve(solves[k], insts[k], &probe[k], nullptr, opt.waves);
      if (st) {
        *ms = -1.0;  // this shape does not fit (LDS): not a candidate
        return TAMC_OK;
      }
    }
    tamc_batch_stats bs{};
    const tamc_status st = run_solves(solves, opt, tmp.data(), &bs);
    if (st) return st;
    for (auto *r : tmp) tamc_pt_result_free(r);
    *ms = bs.device_ms;
    return TAMC_OK;
  };
  TuneChoice best;
  const uint32_t base_teams = fixed_teams == kTeamsTuned ? 0u : fixed_teams;
  double best_ms = 1e300;
  bool any = false;
  for (uint32_t w : {4u, 8u, 16u}) {
    if (fixed_waves && w != fixed_waves) continue;
    for (uint32_t tl : {1u, 2u}) {
      if (fixed_tiling && tl != fixed_tiling) continue;
      RunOptions opt;
      opt.waves = w;
      opt.tiling = tl;
      opt.growth = fixed_growth;
      opt.teams = base_teams;
      double ms;
      const tamc_status st = time_one(opt, &ms);
      if (st) return st;
      if (ms >= 0.0 && ms < best_ms) {
        best_ms = ms;
        best.waves = w;
        best.tiling = tl;
        any = true;
      }
    }
  }
  if (!any && fixed_waves) {  // a caller's wave count that the tuner's candidates do not hold
    best.waves = fixed_waves;
    best.tiling = fixed_tiling ? fixed_tiling : 1u;
  }
  best.growth = fixed_growth;
  for (uint32_t gr : {1u, 2u}) {
    if (!any || fixed_growth) break;
    RunOptions opt;
    opt.waves = best.waves;
    opt.tiling = best.tiling;
    opt.growth = gr;
    opt.teams = base_teams;
    double ms;
    const tamc_status st = time_one(opt, &ms);
    if (st) return st;
    if (ms >= 0.0 && ms < best_ms) {
      best_ms = ms;
      best.growth = gr;
    }
  }
  best.teams = base_teams;
  if (any && fixed_teams == kTeamsTuned) {
    // teams where the cost model places them, kept only when clearly faster (timed twice: the
    // first team launch of a process pays a one-time set-up)
    RunOptions opt;
    opt.waves = best.waves;
    opt.tiling = best.tiling;
    opt.growth = best.growth;
    opt.teams = 1u;
    double ms, ms2;
    tamc_status st = time_one(opt, &ms);
    if (st) return st;
    if ((st = time_one(opt, &ms2))) return st;
    ms = std::min(ms, ms2);
    if (ms >= 0.0 && ms < 0.97 * best_ms) {
      best_ms = ms;
      best.teams = 1u;
    }
  }
  if (any) {
    RunOptions opt;
    opt.waves = best.waves;
    opt.tiling = best.tiling;
    opt.growth = best.growth;
    opt.teams = best.teams;
    opt.sweeps_per_launch = (kTunePrefix + 3u) / 4u;
    opt.use_graph = true;
    double ms;
    const tamc_status st = time_one(opt, &ms);
    if (st) return st;
    best.quarters = ms >= 0.0 && ms < best_ms;
  }
  clear_error();
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_cache[key] = best;
  *out = best;
  return TAMC_OK;
}

tamc_status tamc_pt_run_batch(uint32_t count, const tamc_instance *const *insts, const tamc_pt_params *params,
                              const tamc_batch_options *opts, tamc_pt_result **results, tamc_batch_stats *stats) {
  return tamc_pt_run_batch_init(count, insts, params, nullptr, opts, results, stats);
}

tamc_status tamc_pt_run_batch_init(uint32_t count, const tamc_instance *const *insts, const tamc_pt_params *params,
                                   const int8_t *const *init_states, const tamc_batch_options *opts,
                                   tamc_pt_result **results, tamc_batch_stats *stats) {
  clear_error();
  const auto t_entry = std::chrono::steady_clock::now();
  if (!insts || !params || !results) return fail(TAMC_E_INVALID_ARG, "null argument");
  for (uint32_t k = 0; k < count; ++k)
    if (!insts[k]) return fail(TAMC_E_INVALID_ARG, "null instance");
  tamc_status st = device_init();
  if (st) return st;
  RunOptions opt;
  uint32_t w = 4;
  for (uint32_t k = 0; k < count; ++k) w = std::max(w, default_waves(insts[k]));
  if (opts && opts->waves_per_group) w = opts->waves_per_group;
  if (w < 1 || w > kMaxWaves) return fail(TAMC_E_INVALID_ARG, "waves_per_group must be 1..16");
  if (opts && opts->task_tiling > 2u) return fail(TAMC_E_INVALID_ARG, "task_tiling must be 0, 1 or 2");
  if (opts && opts->icm_growth > 2u) return fail(TAMC_E_INVALID_ARG, "icm_growth must be 0, 1 or 2");
  if (opts && opts->teams > kMaxMembers) return fail(TAMC_E_INVALID_ARG, "teams must be 0..16");
  opt.waves = w;
  opt.tiling = (opts && opts->task_tiling) ? opts->task_tiling : 1u;
  opt.growth = opts ? opts->icm_growth : 0u;
  // options value -> run value: 0 the cost model, 1 none, K forced teams of K
  const uint32_t teams_opt = opts ? opts->teams : 0u;
  opt.teams = teams_opt == 0u ? 1u : (teams_opt == 1u ? 0u : teams_opt);
  opt.sweeps_per_launch = opts ? opts->sweeps_per_launch : 0;
  opt.use_graph = opts && opts->use_graph;
  if (opts && opts->autotune && count &&
      (!opts->waves_per_group || !opts->task_tiling || !opts->icm_growth || !opts->teams ||
       !opts->sweeps_per_launch)) {
    TuneChoice tc;
    if ((st = autotune_batch(count, insts, params, opts->waves_per_group, opts->task_tiling, opts->icm_growth,
                             teams_opt == 0u ? kTeamsTuned : opt.teams, &tc)))
      return st;
    opt.waves = tc.waves;
    opt.tiling = tc.tiling;
    opt.growth = tc.growth;
    opt.teams = tc.teams;
    if (!opts->sweeps_per_launch && tc.quarters) {
      uint32_t maxT = 0;
      for (uint32_t k = 0; k < count; ++k) maxT = std::max(maxT, params[k].num_sweeps);
      opt.sweeps_per_launch = std::max(1u, (maxT + 3u) / 4u);
      opt.use_graph = true;
    }
  }
  opt.t_entry = t_entry;  // the autotuner's probe launches count as preparation
  std::vector<SolvePlan> solves(count);
  for (uint32_t k = 0; k < count; ++k)
    if ((st = plan_solve(solves[k], insts[k], &params[k], init_states ? init_states[k] : nullptr, opt.waves)))
      return st;
  return run_solves(solves, opt, results, stats);
}

tamc_status tamc_pt_run_plain(uint32_t count, const tamc_instance *const *insts, const tamc_pt_params *params,
                              tamc_pt_result **results, tamc_batch_stats *stats) {
  clear_error();
  if (!insts || !params || !results) return fail(TAMC_E_INVALID_ARG, "null argument");
  tamc_status st = device_init();
  if (st) return st;
  tamc_batch_stats acc{};
  for (uint32_t k = 0; k < count; ++k) {
    std::vector<SolvePlan> solves(1);
    RunOptions opt;
    opt.waves = insts[k] ? default_waves(insts[k]) : 8;
    tamc_batch_stats one{};
    if ((st = plan_solve(solves[0], insts[k], &params[k], nullptr, opt.waves)) ||
        (st = run_solves(solves, opt, &results[k], &one))) {
      for (uint32_t j = 0; j < k; ++j) tamc_pt_result_free(results[j]);
      return st;
    }
    acc.device_ms += one.device_ms;
    acc.wall_ms += one.wall_ms;
    acc.attempts += one.attempts;
    acc.launches += one.launches;
    acc.groups += one.groups;
    acc.groups_busy_ms += one.groups_busy_ms;
    acc.teams += one.teams;
    acc.work_groups = std::max(acc.work_groups, one.work_groups);
  }
  if (stats) *stats = acc;
  return TAMC_OK;
}

tamc_status tamc_houdayer(const tamc_instance *inst, uint32_t npairs, int8_t *s1, int8_t *s2, const uint32_t *k,
                          int64_t *cluster_size, uint8_t *clusters) {
  clear_error();
  if (!inst || (npairs && (!s1 || !s2 || !k || !cluster_size))) return fail(TAMC_E_INVALID_ARG, "null argument");
  tamc_status st = device_init();
  if (st) return st;
  const uint32_t n = inst->n;
  for (uint32_t p = 0; p < npairs; ++p) {
    uint32_t cnt = 0;
    for (uint32_t i = 0; i < n; ++i) cnt += ((int)s1[(size_t)p * n + i] * (int)s2[(size_t)p * n + i] < 0);
    if (cnt && k[p] >= cnt) return fail(TAMC_E_INVALID_ARG, "seed index beyond the number of q<0 sites");
  }
  for (uint32_t p0 = 0; p0 < npairs; p0 += kMaxPairs) {
    const uint32_t np = std::min(kMaxPairs, npairs - p0);
    std::vector<float> betas(np, 0.0f);
    std::vector<int8_t> init((size_t)2 * np * n);
    memcpy(init.data(), s1 + (size_t)p0 * n, (size_t)np * n);
    memcpy(init.data() + (size_t)np * n, s2 + (size_t)p0 * n, (size_t)np * n);
    tamc_pt_params P{};
    P.num_sweeps = 1;
    P.warmup_fraction = 1.0;
    P.num_betas = np;
    P.betas = betas.data();
    P.lo_beta_idx = 0;
    P.icm = 1;
    P.num_replica_chains = 2;
    P.threads = 1;
    std::vector<SolvePlan> solves(1);
    RunOptions opt;
    opt.waves = default_waves(inst);
    if ((st = plan_solve(solves[0], inst, &P, init.data(), opt.waves))) return st;
    DevBuf dk;
    TAMC_HIP(dk.alloc(np * 4));
    TAMC_HIP(hipMemcpy(dk.p, k + p0, np * 4, hipMemcpyHostToDevice));
    tamc_pt_result *R = nullptr;
    st = run_solves(solves, opt, &R, nullptr, [&](std::vector<SolvePlan> &ss) {
      ss[0].groups[0].g.flags |= kFlagIcmOnly;
      ss[0].groups[0].g.icm_k = (const uint32_t *)dk.p;
    });
    if (st) return st;
    for (uint32_t p = 0; p < np; ++p) {
      int8_t *a = s1 + (size_t)(p0 + p) * n, *b = s2 + (size_t)(p0 + p) * n;
      const int8_t *na = R->final_states + (size_t)p * n, *nb = R->final_states + ((size_t)np + p) * n;
      int64_t sz = 0;
      for (uint32_t i = 0; i < n; ++i) {
        const bool ch = a[i] != na[i];
        sz += ch;
        if (clusters) clusters[(size_t)(p0 + p) * n + i] = ch ? 1 : 0;
        a[i] = na[i];
        b[i] = nb[i];
      }
      cluster_size[p0 + p] = sz;
    }
    tamc_pt_result_free(R);
  }
  return TAMC_OK;
}

void tamc_pt_result_free(tamc_pt_result *R) {
  if (!R) return;
  void *ptrs[] = {R->e, R->q, R->suscept, R->samples, R->gs_steps, R->gs_energies, R->gs_states,
                  R->acceptance_counts, R->final_states, R->final_energies, R->final_ids, R->num_acceptances,
                  R->round_trips, R->tags, R->diffusion_hist};
  for (void *p : ptrs) free(p);
  free(R);
}

}  // extern "C"
