// pt_types.hpp -- layout of one PT-ICM work group ("group" = one instance x one ICM chain pair)
// as seen by the gfx950 kernels, plus the LDS arena layout.  Shared by kernels.hip and runtime.cpp.
//
// Spin storage ("bit planes"): for every site i the group keeps G = ceil(R/64) 64-bit words, one bit
// per replica (bit set <=> spin -1, the reference's packing convention).  A wavefront of 64 lanes
// handles the 64 replicas of one word, so coupling rows are read once per 64 replicas and every
// neighbour-spin read is a single broadcast LDS load.
#pragma once
#include <stdint.h>

namespace tamc {

constexpr uint32_t kMaxWaves = 16;       // waves per work group (1024 threads)
constexpr uint32_t kMaxGroupReplicas = 1024;  // 16 plane words per site
constexpr uint32_t kMaxPairs = 32;       // ICM pairs per group held in one 32-bit cluster word
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kMaxTableWords = 4;   // ICM overlap by byte tables for 2 to 4 plane words per site

// Per-group phase timers (100 MHz wall-clock ticks of the group's thread 0); the last slot counts
// ICM cluster-growth passes.
enum : uint32_t {
  kPhIcmOverlap = 0, kPhIcmSeed, kPhIcmGrow, kPhIcmFlip, kPhExchange, kPhSweep, kPhMeasure, kPhIcmPasses,
  kPhaseSlots
};

enum : uint32_t {
  kFlagIcm = 1u,
  kFlagThreaded = 2u,        // stale exchange dE (tamc-core/src/parallel/pt.rs:51-60)
  kFlagSample = 4u,
  kFlagSampleStates = 8u,
  kFlagIcmGlobal = 16u,      // ICM scratch lives in global memory (group too large for LDS)
  kFlagIcmOnly = 32u,        // run only the ICM step with seeds given in icm_k (tamc_houdayer)
  kFlagTileBlocks = 64u,     // sweep tiling: each wave takes one contiguous run of a level's tasks
                             // (else the tasks are dealt to the waves in turn)
  kFlagIcmTiles = 128u,      // ICM cluster growth over compact tiles of the coupling graph
  kFlagSpinsGlobal = 256u,   // the spin planes stay in the group's global state array (group too
                             // large for LDS; implies kFlagIcmGlobal)
};

struct PtGroup {
  // ---- instance (device pointers, read only) ----
  const uint32_t *row_ptr;   // [n+1] coupling_vecs
  const uint32_t *col;       // [nnz]
  const float *J;            // [nnz]
  const float *bias;         // [n]
  const uint32_t *g_row_ptr; // [g_nodes+1] coupling CsMat pattern (ICM graph)
  const uint32_t *g_col;
  const double *chi;         // [nchi][n]
  float offset;
  uint32_t n, g_nodes, nchi, nblk;
  // ---- sequential-order schedule: levels of mutually uncoupled sites, in groups of four ----
  const uint32_t *lvl_off;   // [n_levels+1] offsets into tasks
  const uint32_t *tasks;     // group*G + g
  const uint32_t *grp;       // [n_groups][4] sites of each group (kNone pads)
  const uint32_t *records;   // [n_groups][8 + 8 * ell_D]: sites, bias, cols, J of each group
  uint32_t n_levels;
  // ---- ladder ----
  const float *betas;        // [B]
  uint32_t B, nch, chain0, lo; // chains in group (1 or 2), first global chain, first ICM slot
  uint32_t G, R, P, flags;     // words per site, replicas, ICM pairs, kFlag*
  uint32_t T, meas_init, sample, sample_states;
  uint32_t sample_limiting, n_samp_slots, gs_cap, nbytes;
  uint32_t gs_words, pair_global, pad0, pad1;
  uint64_t seed;
  // ---- persistent state (global memory, carried between launches) ----
  uint32_t *spins;           // [n][G][2] 32-bit halves of the bit-plane words
  float *energy;             // [R] tracked energies per physical replica
  uint32_t *perm;            // [nch][B] slot -> physical replica index in the group
  uint8_t *tags;             // [nch][B] round-trip tags in slot order
  uint32_t *nacc;            // [nch][B]
  uint32_t *hist;            // [nch][B][2]
  uint32_t *rt;              // [nch]
  float *gs_min;             // [1]
  uint32_t *gs_count;        // [1]
  unsigned long long *icm_moves; // [1]
  // ---- outputs ----
  float *out_e;              // [n_meas][nch][B]
  int32_t *out_q;            // [n_meas][B]
  float *out_chi;            // [n_meas][B][nchi]
  uint8_t *out_samples;      // [n_samp_events][n_samp_slots][nch][nbytes]
  uint32_t *out_gs_steps;    // [gs_cap]
  float *out_gs_e;           // [gs_cap]
  uint64_t *out_gs_states;   // [gs_cap][gs_words]
  // ---- ICM scratch when it does not fit in LDS ----
  uint32_t *g_qw, *g_cw, *g_blk;
  // where the spin planes wait while ICM clusters grow in the LDS they leave: null for the group's
  // spin array; a private area for team members other than the lead (the lead alone writes the
  // group's array: plain stores of other members would sit in their XCDs' L2 caches and could be
  // written back after the lead's final state)
  uint32_t *g_park;
  // ---- kFlagIcmOnly: seed index per pair (into the ascending q<0 list) ----
  const uint32_t *icm_k;
  // ---- fixed-stride copy of coupling_vecs for rows of at most ell_D entries (0 = use CSR) ----
  const uint32_t *ell_col;   // [n][ell_D], padding entries point at the row's own site
  const uint16_t *ell_col16; // 16-bit neighbour lists [n][nbr_stride(ell_D)] (n < 65536), for ICM
  const float *ell_J;        // [n][ell_D], padding entries are +0.0
  uint32_t ell_D, icm_ntiles;
  uint32_t lat_L, lat_dim;   // a periodic lattice with row-major sites: L x L (lat_dim 2) or L x L x L
                             // (lat_dim 3); lat_L = 0 for other graphs
  const uint32_t *icm_tiles; // [icm_ntiles][64] sites of each compact tile (kNone pads), for ICM
  const uint16_t *icm_trows; // [icm_ntiles * 64][nbr_stride(ell_D)] their neighbour rows, tile order
  // ---- QUBO self entries (src/ising.rs:308-310): J_ii per site, null when the instance has none.
  // Their flips shift the reference's tracked energy by -2 J_ii each until energy_ref is recomputed
  // after an ICM move; `drift` carries that shift (2^-32 fixed point) per replica.
  const float *self_J;
  long long *drift;          // [R]
  // ---- per-phase time of the group (100 MHz wall clock ticks): ICM overlap+seeds, ICM cluster growth,
  //      ICM flip, exchange, sweep, measurement ----
  unsigned long long *phase_ticks;  // [kPhaseSlots]
  // ---- teams: a group whose Metropolis sweep is shared by `members` work groups of one launch,
  //      all resident at once.  Every member holds the whole state in its LDS and runs ICM and
  //      the exchange itself (the same inputs give the same results); member m updates the RNG
  //      groups g with g % members == m and publishes the new spin words (and, after the last
  //      colour class, its share of the accepted energy) in team_box, tagged with the number of the
  //      meeting; the others read them when the tags match.  In ICM every member flips every
  //      cluster and sums the terms of its own blocks; the members exchange the sums.  Member 0
  //      alone writes measurements and the final state. ----
  uint32_t members, member;  // 1, 0 for a group run by one work group
  uint32_t *team_ctr;        // [kTeamCtrWords]: -, abort flag, -, meeting counts of member 0, 1, ...
  unsigned long long *team_box;  // tagged values (team_box_slots)
};

constexpr uint32_t kMaxMembers = 16;
constexpr uint32_t kTeamCtrWords = 3 + kMaxMembers;

// Team box, in 64-bit slots; every value takes two slots (32 bits of it beside the epoch of the
// meeting that published it): the spin words [n * G], the members' accepted-energy sums
// [members][R], then two buffers of the members' ICM per-pair sums [members][kMaxPairs].
__host__ __device__ inline size_t team_sums_slot(uint32_t n, uint32_t G) { return 2u * (size_t)n * G; }
__host__ __device__ inline size_t team_pair_slot(uint32_t n, uint32_t G, uint32_t R, uint32_t K, uint32_t par) {
  return team_sums_slot(n, G) + 2u * (size_t)K * R + par * 2u * (size_t)K * kMaxPairs;
}
__host__ __device__ inline size_t team_box_slots(uint32_t n, uint32_t G, uint32_t R, uint32_t K) {
  return team_pair_slot(n, G, R, K, 2u);
}

struct LdsLayout {
  uint32_t spins, energy, beta_of, perm, tags, nacc, hist, rt, pair_a, pair_b, pair_of, seed_blk, misc, ticks, fmask, fscr, partial, dE, drift,
      uw, dEa, qw, cw, blk, nbr, total;
};

__host__ __device__ inline uint32_t lds_align(uint32_t x) { return (x + 15u) & ~15u; }

// ICM runs the pairs of a group in chunks of at most kMaxPairs (one bit per pair in the 32-bit
// overlap and cluster words); the number of pairs one chunk holds
__host__ __device__ inline uint32_t icm_chunk(uint32_t P) { return P < kMaxPairs ? P : kMaxPairs; }

// icm_lds: ICM arrays (q words, cluster words, per-block counts) are placed in LDS; spins_lds: so
// are the spin planes.
__host__ __device__ inline LdsLayout make_layout(uint32_t n, uint32_t G, uint32_t R, uint32_t B, uint32_t P,
                                                 uint32_t W, bool icm_lds, bool spins_lds = true) {
  LdsLayout L;
  uint32_t o = 0;
  const uint32_t nblk = (n + 63u) / 64u;
  L.spins = spins_lds ? o : kNone;
  if (spins_lds) o += lds_align(n * G * 8u);
  L.energy = o;
  o += lds_align(R * 4u);
  L.beta_of = o;
  o += lds_align(R * 4u);
  L.perm = o;
  o += lds_align(2u * B * 4u);
  L.tags = o;
  o += lds_align(2u * B);
  L.nacc = o;
  o += lds_align(2u * B * 4u);
  L.hist = o;
  o += lds_align(4u * B * 4u);
  L.rt = o;
  o += lds_align(2u * 4u);
  L.pair_a = o;
  o += lds_align(kMaxPairs * 4u);
  L.pair_b = o;
  o += lds_align(kMaxPairs * 4u);
  L.pair_of = o;
  o += lds_align(R * 4u);
  L.seed_blk = o;
  o += lds_align(kMaxPairs * 4u);
  L.misc = o;  // [0..1] measurement, [2..3] ground-state record count and minimum, [4] ICM moves,
                // [8..10] growth pass flags, [12..14] round flags of block_any, [15] team meetings
  o += lds_align(16u * 4u);
  L.ticks = o;  // phase timers [kPhaseSlots] + last timestamp
  o += lds_align((kPhaseSlots + 1u) * 8u);
  L.fmask = o;  // ICM flip masks [G][kMaxPairs / 4][16]
  o += lds_align(P > 0 ? G * (kMaxPairs / 4u * 16u * 8u) : 8u);
  L.fscr = o;  // ICM cluster labelling: a site of each straggling pair's partial cluster
  o += lds_align(P > 0 ? kMaxPairs * 4u : 4u);
  L.partial = o;  // per-wave partial sums: W*G*64 int64 (also W*B u32 counters since B <= 64*G);
                  // during ICM the overlap tables, 8 KB per plane word (two to four words)
  const uint32_t tabs = (P > 0 && G > 1u && G <= kMaxTableWords) ? G * 8192u : 0u;
  o += lds_align(W * G * 64u * 8u > tabs ? W * G * 64u * 8u : tabs);
  L.dE = o;  // exchange energy differences [2][B]
  o += lds_align(2u * B * 4u);
  L.drift = o;
  o += lds_align(R * 8u);
  L.uw = o;  // exchange uniforms [2][B]
  o += lds_align(2u * B * 4u);
  L.dEa = o;  // ICM energy change of the first replica of each pair (fixed point)
  o += lds_align(kMaxPairs * 8u);
  if (icm_lds && P > 0) {
    L.qw = o;
    o += lds_align(n * 4u);
    L.cw = o;
    o += lds_align(n * 4u);
    L.blk = o;  // per-pair counts of q<0 sites in each 64-site block, [P][nblk] (one chunk of pairs)
    o += lds_align(nblk * icm_chunk(P) * 4u);
  } else {
    L.qw = L.cw = L.blk = kNone;
  }
  L.nbr = kNone;
  L.total = o;
  return L;
}

// Adds the LDS copy of the 16-bit neighbour lists used by ICM when it fits under `limit`.
// Row stride of the 16-bit neighbour lists: D rounded up to 4, so that a row is one 8- or 16-byte
// load (padding entries repeat the row's own site).
__host__ __device__ inline uint32_t nbr_stride(uint32_t D) { return (D + 3u) & ~3u; }

__host__ __device__ inline LdsLayout with_nbr(LdsLayout L, uint32_t n, uint32_t D, uint32_t limit) {
  if (D == 0u || L.qw == kNone) return L;
  const uint32_t need = lds_align(n * nbr_stride(D) * 2u);
  if (L.total + need > limit) return L;
  L.nbr = L.total;
  L.total += need;
  return L;
}

}  // namespace tamc
