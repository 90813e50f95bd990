/*
 * tamc_oracle.h -- CPU restatement of the tamc sampling path (test infrastructure).
 *
 * THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, the smoke check and
 * bench.py's cpu_baseline leg may load it.  It restates, sequentially and in the
 * reference's floating-point operation order, the functions of USCqserver/tamc that
 * the MI355X path replaces, so the HIP kernels can be checked against it:
 *
 *   parser            src/util.rs:43-109, instance build src/ising.rs:286-324
 *   energy_ref        src/ising.rs:438-454      delta_energy  src/ising.rs:468-480
 *   Metropolis sweep  tamc-core/src/metropolis.rs:43-57, src/ising.rs:225-233
 *   PT exchange       tamc-core/src/pt.rs:120-150 (serial), tamc-core/src/parallel/pt.rs:41-69 (threaded)
 *   round trips       tamc-core/src/pt.rs:54-84
 *   Houdayer ICM      src/pt.rs:31-77, apply_icm src/pt.rs:521-540
 *   measurements      src/pt.rs:167-226, 542-583
 *   runner loops      src/pt.rs:414-504, SA src/sa.rs:167-210, tamc-core/src/sa.rs:13-48
 *
 * Random inputs come from one of three sources (the uniform-feed contract):
 *   TAMCO_RNG_XOSHIRO  Xoshiro256++ with the rand-0.8 conversions (reference RNG)
 *   TAMCO_RNG_FEED     caller-supplied u32 word streams (lock-step mode)
 *   TAMCO_RNG_PHILOX   counter-based Philox4x32-10 keyed per (sweep, site-quad, replica),
 *                      the production scheme of the GPU path.
 *
 * Parity status of each function is listed in the header comment of tamc_oracle.cpp.
 */
#ifndef TAMC_ORACLE_H
#define TAMC_ORACLE_H
#include <stdint.h>
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { TAMCO_RNG_XOSHIRO = 0, TAMCO_RNG_FEED = 1, TAMCO_RNG_PHILOX = 2 };

/* ---- RNG primitives ---- */
void tamco_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void tamco_xoshiro_seed(uint64_t seed, uint64_t st[4]);
uint64_t tamco_xoshiro_next(uint64_t st[4]);
void tamco_xoshiro_jump(uint64_t st[4]);
/* rand 0.8 gen_range(0..len as u32) over a word array; returns index, advances *cursor */
uint32_t tamco_gen_index(uint32_t len, const uint32_t *words, uint64_t nwords, uint64_t *cursor);

/* ---- instance ---- */
typedef struct tamco_instance tamco_instance;
tamco_instance *tamco_instance_from_file(const char *path, int qubo);
/* emulate an instance file whose lines are (i[k], j[k], K[k]) in order */
tamco_instance *tamco_instance_from_lines(uint64_t nlines, const uint32_t *i, const uint32_t *j,
                                          const double *K, int qubo);
void tamco_instance_free(tamco_instance *);
uint32_t tamco_instance_n(const tamco_instance *);
uint32_t tamco_instance_nnz(const tamco_instance *);
float tamco_instance_offset(const tamco_instance *);
/* copy out: row_ptr[n+1], col[nnz], J[nnz], bias[n] */
void tamco_instance_csr(const tamco_instance *, uint32_t *row_ptr, uint32_t *col, float *J, float *bias);
uint32_t tamco_instance_graph_nnz(const tamco_instance *);
uint32_t tamco_instance_graph_nodes(const tamco_instance *);
void tamco_instance_graph(const tamco_instance *, uint32_t *row_ptr, uint32_t *col);
/* suscept coefficients: nchi vectors of n doubles (row-major) */
void tamco_instance_set_suscept(tamco_instance *, uint32_t nchi, const double *w);

/* ---- elementary functions on states ---- */
float tamco_energy_ref(const tamco_instance *, const int8_t *s);
float tamco_delta_energy(const tamco_instance *, const int8_t *s, uint32_t i);
/* one Metropolis sweep consuming f32 uniforms from a word stream; returns the summed accepted dE */
float tamco_sweep_feed(const tamco_instance *, int8_t *s, float beta, const uint32_t *words,
                       uint64_t nwords, uint64_t *cursor);
/* one Metropolis sweep with Philox production uniforms for replica id rho at sweep t */
float tamco_sweep_philox(const tamco_instance *, int8_t *s, float beta, uint64_t seed, uint32_t t,
                         uint32_t rho);
int64_t tamco_overlap(const int8_t *a, const int8_t *b, uint32_t n);
double tamco_suscept(const tamco_instance *, uint32_t k, const int8_t *overlap);
void tamco_pack_bytes(const int8_t *s, uint32_t n, uint8_t *out);   /* ceil(n/8) bytes */
void tamco_pack_u64(const int8_t *s, uint32_t n, uint64_t *out);    /* ceil(n/64) words */
/* Houdayer move with the seed given as index k into the ascending list of q<0 sites.
 * returns cluster size, 0 if no q<0 site (no move), -1 if the seed lies outside the graph.
 * cluster_out (n bytes, nullable) receives 1 for sites in the cluster. */
int64_t tamco_houdayer_k(const tamco_instance *, int8_t *s1, int8_t *s2, uint32_t k, uint8_t *cluster_out);
/* number of q<0 sites */
uint32_t tamco_num_negative_overlap(const int8_t *s1, const int8_t *s2, uint32_t n);
/* Exchange pass on one chain.  energies[B] are slot energies; perm[B] in/out maps slot->state id;
 * words: uniform stream; variant 0 serial (dE bookkeeping), 1 threaded (stale dE).
 * accept_out[B-1] receives the swap decisions, nacc[B] is incremented like num_acceptances. */
void tamco_exchange(uint32_t B, const float *betas, const float *energies, uint32_t *perm, int variant,
                    const uint32_t *words, uint64_t nwords, uint64_t *cursor, uint8_t *accept_out,
                    uint32_t *nacc);
/* One exchange pass on one chain with the decisions given (swaps[j] != 0: slots j, j+1 swap, in order
 * j = 0..B-2), then update_round_trips.  tags [B], nacc [B], hist [B][2] and rt are updated in place. */
void tamco_chain_step(uint32_t B, const uint8_t *swaps, uint8_t *tags, uint32_t *nacc, uint32_t *hist, uint32_t *rt);
/* geometric beta schedule cast to f32 (tamc-core/src/sa.rs:13-22 via src/pt.rs:246-256) */
void tamco_beta_geometric(float beta_min, float beta_max, uint32_t num_beta, float *out);
/* expf used by the restatement (the reference's f32::exp) */
float tamco_expf(float x);

/* ---- PT-ICM runner ---- */
typedef struct {
  uint32_t num_sweeps;
  double warmup_fraction;
  uint32_t num_betas;
  const float *betas;
  uint32_t lo_beta_idx;
  int32_t icm;
  uint32_t num_chains;
  int32_t threaded;          /* 1: parallel_pt_loop semantics (stale exchange dE, per-slot streams) */
  int32_t sample_enabled;
  uint32_t sample;
  int32_t sample_states_enabled;
  uint32_t sample_states;
  uint32_t sample_limiting;
  int32_t rng_mode;          /* TAMCO_RNG_* */
  uint64_t seed;
  /* TAMCO_RNG_FEED: threaded -> num_chains*num_betas streams indexed c*B+b; serial -> 1 stream */
  uint32_t num_streams;
  const uint32_t *const *stream_words;
  const uint64_t *stream_lens;
  const int8_t *init_states; /* nullable, [C][B][n] slot order */
  /* TAMCO_RNG_PHILOX only: 0 = sweep colour class by colour class (the GPU's production order),
   * 1 = the reference's sequential order (0..n-1) */
  uint32_t sweep_order;
  /* resuming from a full PTState (src/pt.rs:397-408), all nullable, slot order: cached energies
   * [C][B], num_acceptances [C][B], round_trips [C], round-trip tags [C][B], diffusion_hist [C][B][2] */
  const float *init_energies;
  const uint32_t *init_nacc;
  const uint32_t *init_rt;
  const uint8_t *init_tags;
  const uint32_t *init_hist;
  /* TAMCO_RNG_PHILOX only: 1 = the production path's tracked-energy accounting instead of the
   * reference's f32 accumulation: a sweep's accepted dE are summed exactly (2^-32 fixed point) and
   * added to the f32 energy once; an ICM move adds the cluster's exact energy change to the tracked
   * energy (less the QUBO self-entry drift) instead of recomputing energy_ref.  For +-J and QUBO
   * couplings both give the same bits; for Gaussian couplings they differ in the last bits. */
  uint32_t energy_fixed;
} tamco_pt_params;

typedef struct {
  uint32_t n, B, C, T, meas_init, nchi, npairs;
  uint32_t n_meas;            /* measurement events */
  float *e;                   /* [B][n_meas*C] */
  int32_t *q;                 /* [B][n_meas*npairs] */
  float *suscept;             /* [B][nchi][n_meas*npairs] */
  uint32_t n_samp_slots, n_samp_events, nbytes;
  uint8_t *samples;           /* [n_samp_slots][n_samp_events*C][nbytes] */
  uint32_t n_gs, gs_words;
  uint32_t *gs_steps;
  float *gs_energies;
  uint64_t *gs_states;        /* [n_gs][gs_words] */
  uint32_t *acceptance_counts;/* [B] summed over chains */
  int8_t *final_states;       /* [C][B][n] slot order */
  float *final_energies;      /* [C][B] cached energies */
  uint32_t *final_ids;        /* [C][B] physical replica id at each slot */
  uint32_t *num_acceptances;  /* [C][B] */
  uint32_t *round_trips;      /* [C] */
  uint8_t *tags;              /* [C][B] PTRoundTrip ordinal */
  uint32_t *diffusion_hist;   /* [C][B][2] */
  uint64_t icm_moves;         /* number of non-empty Houdayer moves */
  uint64_t attempts;          /* Metropolis attempts performed */
  int32_t error;              /* 0 ok; 1 stream exhausted; 2 ICM seed outside graph; 3 invalid args */
} tamco_pt_result;

tamco_pt_result *tamco_pt_run(const tamco_instance *, const tamco_pt_params *);
void tamco_pt_result_free(tamco_pt_result *);

/* ---- SA runner ---- */
typedef struct {
  uint32_t num_betas;
  const float *betas;
  uint32_t num_replicas;
  uint32_t threads;          /* reference SaParams.threads: partitions of num_replicas/threads */
  int32_t rng_mode;          /* XOSHIRO or PHILOX */
  uint64_t seed;
  const int8_t *init_states; /* nullable [R][n] */
} tamco_sa_params;

typedef struct {
  uint32_t n, R;             /* R = replicas actually simulated */
  float min_energy;
  float *energies;           /* [R] energy_ref of the final states */
  int8_t *final_states;      /* [R][n] */
  uint64_t attempts;
  int32_t error;
} tamco_sa_result;

tamco_sa_result *tamco_sa_run(const tamco_instance *, const tamco_sa_params *);
void tamco_sa_result_free(tamco_sa_result *);

#ifdef __cplusplus
}
#endif
#endif
