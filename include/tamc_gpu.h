/*
 * tamc_gpu.h -- C ABI of the MI355X (gfx950) sampling engine for tamc's hot path:
 * Metropolis sweeps over sparse couplings, parallel-tempering replica exchange, Houdayer
 * isoenergetic cluster moves (ICM) and simulated annealing.
 *
 * The reference (USCqserver/tamc) has no FFI: its boundary is the `tamc` CLI plus the Rust runner
 * API.  Each entry point below names the reference item it replaces (paths relative to that tree).
 * All entry points take plain pointers and sizes, return a tamc_status, never throw or abort, and
 * report details through tamc_last_error() (thread-local).  Inputs are copied; results are
 * allocated by the library and released with the matching *_free function.
 *
 * Spins are int8 (+1/-1).  Bit-packed states use the reference's convention: bit set <=> spin -1,
 * least significant bit first.
 */
#ifndef TAMC_GPU_H
#define TAMC_GPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  TAMC_OK = 0,
  TAMC_E_INVALID_ARG = 1, /* reference: panics on bad beta order, missing lo_beta with icm, ... */
  TAMC_E_IO = 2,          /* reference: "Unable to read adjancency from instance file" */
  TAMC_E_DEVICE = 3,      /* HIP runtime error or no gfx950 device */
  TAMC_E_NOMEM = 4,
  TAMC_E_UNSUPPORTED = 5, /* shape outside what the kernels accept (see tamc_last_error) */
  TAMC_E_STREAM = 6,      /* lock-step uniform stream exhausted */
  TAMC_E_PENDING = 7      /* tamc_server_poll: the request has not finished yet */
} tamc_status;

/* Last error message of the calling thread ("" if none). */
const char *tamc_last_error(void);
/* Library ABI version (major*10000 + minor*100 + patch). */
int tamc_version(void);
/* Initialise the HIP device (done lazily by every entry point that needs it). */
tamc_status tamc_init(void);

/* ------------------------------------------------------------------ instances ---- */
typedef struct tamc_instance tamc_instance;

/* BqmIsingInstance::from_instance_file (src/ising.rs:286-324) over read_adjacency_list
 * (src/util.rs:69-109): "i j K" lines, quirks included (exponents truncated, duplicates overwrite). */
tamc_status tamc_instance_from_file(const char *path, int qubo, tamc_instance **out);
/* Same semantics from in-memory lines (i[k], j[k], K[k]) in file order. */
tamc_status tamc_instance_from_lines(uint64_t nlines, const uint32_t *i, const uint32_t *j, const double *K,
                                     int qubo, tamc_instance **out);
/* The fields of BqmIsingInstance (src/ising.rs:257-263) from caller arrays, for callers that already
 * hold one (from new_zero_bias, src/ising.rs:266-285, from_instance_file, or built by hand):
 * coupling_vecs as CSR -- row i is entries row_ptr[i] .. row_ptr[i+1]-1 in the row's own order,
 * which is the order the fields are summed in (QUBO self entries (i, K_ii/2) included, as
 * src/ising.rs:308-310 pushes them) -- plus bias, offset and suscept_coefs.  The coupling CsMat
 * (the ICM graph, src/pt.rs:343-346) is the off-diagonal pattern of coupling_vecs, as both
 * reference constructors build it.  That pattern must be symmetric (j in row i <=> i in row j for
 * i != j): a Houdayer cluster is a connected component of an undirected graph.  The arrays are
 * copied. */
typedef struct {
  uint32_t n;
  uint32_t nnz;
  const uint32_t *row_ptr;  /* [n+1]: row_ptr[0] = 0, non-decreasing, row_ptr[n] = nnz */
  const uint32_t *col;      /* [nnz], each < n */
  const float *J;           /* [nnz] */
  const float *bias;        /* [n]; null = zero biases (new_zero_bias) */
  float offset;
  uint32_t nchi;
  const double *suscept;    /* [nchi][n] suscept_coefs; null when nchi = 0 */
  int32_t qubo;             /* reported by tamc_instance_get_info; changes nothing else */
} tamc_instance_desc;
tamc_status tamc_instance_create(const tamc_instance_desc *desc, tamc_instance **out);
/* BqmIsingInstance::with_suscept (src/ising.rs:388-403): nchi vectors of n coefficients (row-major).
 * Like the reference's builder, this is for set-up: an instance must not change while a call that
 * uses it runs, and both suscept setters fail with TAMC_E_INVALID_ARG while server requests on the
 * instance are queued or resident. */
tamc_status tamc_instance_set_suscept(tamc_instance *inst, uint32_t nchi, const double *w);
/* read_txt_vec (src/util.rs:112-139) on one file, appended as one more suscept vector.  Like the
 * reference, a length mismatch is reported on stdout and the vector kept; runs then fail. */
tamc_status tamc_instance_add_suscept_file(tamc_instance *inst, const char *path);

typedef struct {
  uint32_t n;          /* BqmIsingInstance::size */
  uint32_t nnz;        /* entries of coupling_vecs (QUBO self entries included) */
  uint32_t graph_nnz;  /* nonzeros of the coupling CsMat (ICM graph) */
  uint32_t graph_nodes;/* petgraph Csr node count = max index + 1 */
  uint32_t nchi;
  float offset;
  int32_t qubo;
} tamc_instance_info;
tamc_status tamc_instance_get_info(const tamc_instance *inst, tamc_instance_info *info);
/* Copy out coupling_vecs in CSR form: row_ptr[n+1], col[nnz], J[nnz], bias[n]. */
tamc_status tamc_instance_get_csr(const tamc_instance *inst, uint32_t *row_ptr, uint32_t *col, float *J,
                                  float *bias);
/* Copy out the ICM graph (the coupling CsMat pattern, src/pt.rs:343-346) in CSR form:
 * row_ptr[graph_nodes+1], col[graph_nnz]. */
tamc_status tamc_instance_get_graph(const tamc_instance *inst, uint32_t *row_ptr, uint32_t *col);
/* BqmIsingInstance::induced_subgraph_instance_sorted (src/ising.rs:327-355): the instance induced
 * by the ascending site list nodes[count]; sites are renumbered 0..count-1, the offset is 0. */
tamc_status tamc_instance_induced_subgraph(const tamc_instance *inst, const uint32_t *nodes, uint32_t count,
                                           tamc_instance **out);
void tamc_instance_destroy(tamc_instance *inst);

/* --------------------------------------------- elementary device functions ---- */
/* Instance::energy_ref (src/ising.rs:438-454) of nstates states [nstates][n], reference order. */
tamc_status tamc_energy(const tamc_instance *inst, uint32_t nstates, const int8_t *states, float *energies);
/* Instance::delta_energy (src/ising.rs:468-480) of every site of every state: out[nstates][n]. */
tamc_status tamc_delta_energy(const tamc_instance *inst, uint32_t nstates, const int8_t *states, float *out);
/* IsingState::overlap (src/ising.rs:159-165) of state pairs (a[k], b[k]). */
tamc_status tamc_overlap(uint32_t n, uint32_t npairs, const int8_t *a, const int8_t *b, int64_t *q);
/* IsingState::write_to_bytes (src/ising.rs:96-121): out[nstates][ceil(n/8)]. */
tamc_status tamc_pack_bytes(uint32_t n, uint32_t nstates, const int8_t *states, uint8_t *out);
/* houdayer_cluster_move (src/pt.rs:31-77) with the seed given as index k[p] into the ascending list
 * of q<0 sites of pair p.  s1/s2 [npairs][n] are updated in place; cluster_size[p] = 0 when the pair
 * has no q<0 site; clusters (nullable) [npairs][n] receives 1 for flipped sites. */
tamc_status tamc_houdayer(const tamc_instance *inst, uint32_t npairs, int8_t *s1, int8_t *s2, const uint32_t *k,
                          int64_t *cluster_size, uint8_t *clusters);

/* --------------------------------------------------------------- PT-ICM ---- */
/* PtIcmParams (src/pt.rs:259-272) with BetaOptions and lo_beta resolved by tamc_pt_resolve. */
typedef struct {
  uint32_t num_sweeps;
  double warmup_fraction;
  uint32_t num_betas;
  const float *betas;       /* non-decreasing (PtIcmRunner::new panics otherwise) */
  uint32_t lo_beta_idx;     /* first beta slot that takes part in ICM */
  int32_t icm;
  uint32_t num_replica_chains;
  uint32_t threads;         /* >1 selects parallel_pt_loop semantics: stale exchange dE */
  int32_t sample_enabled;   /* Option<u32> sample */
  uint32_t sample;
  int32_t sample_states_enabled;
  uint32_t sample_states;
  uint32_t sample_limiting; /* Option<u8>, 0 when None */
  uint64_t seed;            /* production Philox key; results depend only on (instance, params, seed) */
  uint32_t sweep_order;     /* TAMC_ORDER_COLOUR (production default) or TAMC_ORDER_SEQUENTIAL */
} tamc_pt_params;

/* Site order of a production sweep.  COLOUR: colour classes of a greedy colouring in site order
 * (colour(i) = the smallest colour not taken by a lower-numbered coupled site), swept class by
 * class; a valid Metropolis sweep with the same stationary distribution as the reference's.
 * SEQUENTIAL: classes = levels (level(i) = 1 + max level of lower-numbered coupled sites), which
 * reproduces the reference's sequential site order exactly (tamc-core/src/metropolis.rs:77-83). */
enum { TAMC_ORDER_COLOUR = 0, TAMC_ORDER_SEQUENTIAL = 1 };

/* PtIcmRunner::new (src/pt.rs:307-351): geometric betas (tamc-core/src/sa.rs:13-22, cast to f32) or an
 * explicit array, the non-decreasing check and lo_beta_idx.  lo_num_beta < 0 and lo_beta NaN mean None.
 * betas_out must hold num_beta (geometric) or narr (array) floats. */
tamc_status tamc_pt_resolve(int32_t geometric, float beta_min, float beta_max, uint32_t num_beta, const float *arr,
                            uint32_t narr, float lo_beta, int64_t lo_num_beta, int32_t icm, float *betas_out,
                            uint32_t *num_betas_out, uint32_t *lo_beta_idx_out);

/* PtIcmMinResults + PtIcmThermalSamples + final Vec<PTState> (src/pt.rs:96-227, tamc-core/src/pt.rs:18-25). */
typedef struct {
  uint32_t n, B, C, T, meas_init, nchi, npairs;
  uint32_t n_meas;             /* measurement events */
  float *e;                    /* [B][n_meas*C]            PtIcmThermalSamples.e */
  int32_t *q;                  /* [B][n_meas*npairs]       PtIcmThermalSamples.q */
  float *suscept;              /* [B][nchi][n_meas*npairs] PtIcmThermalSamples.suscept */
  uint32_t n_samp_slots, n_samp_events, nbytes;
  uint8_t *samples;            /* [n_samp_slots][n_samp_events*C][nbytes] */
  uint32_t n_gs, gs_words;
  uint32_t *gs_steps;          /* MinResults.gs_time_steps */
  float *gs_energies;          /* MinResults.gs_energies */
  uint64_t *gs_states;         /* [n_gs][gs_words] MinResults.gs_states */
  uint32_t *acceptance_counts; /* [B] count_acc (src/pt.rs:585-591) */
  int8_t *final_states;        /* [C][B][n] slot order */
  float *final_energies;       /* [C][B] tracked energies */
  uint32_t *final_ids;         /* [C][B] physical replica id (chain*B + initial slot) */
  uint32_t *num_acceptances;   /* [C][B] */
  uint32_t *round_trips;       /* [C] */
  uint8_t *tags;               /* [C][B] PTRoundTrip ordinal: None, MinBeta, MaxBeta, MinBetaRefl, MaxBetaRefl */
  uint32_t *diffusion_hist;    /* [C][B][2] */
  uint64_t icm_moves;
  uint64_t attempts;           /* Metropolis spin-flip attempts performed */
  double timing_us;            /* sweep loop time, as MinResults.timing (src/pt.rs:445-454) */
  double device_ms;            /* HIP-event time of the kernels */
} tamc_pt_result;

/* PtIcmRunner::run / run_parallel (src/pt.rs:354-412): production path.  Sweeps in an order that
 * is equivalent to the reference's sequential site order, Philox uniforms keyed by
 * (seed, sweep, group of four mutually uncoupled sites, replica), exchange and ICM on the device.
 * init_states (nullable) [C][B][n] replaces generate_init_state (src/pt.rs:506-519). */
tamc_status tamc_pt_run(const tamc_instance *inst, const tamc_pt_params *params, const int8_t *init_states,
                        tamc_pt_result **out);

/* Lock-step correctness mode: the same loop, but every random number comes from host-supplied u32
 * word streams consumed in the reference's order (rand 0.8 conversions on the device).
 * threads > 1: num_streams = C*B, stream c*B+b feeds slot b of chain c (exchange on stream c*B,
 *              ICM choices on stream 0) -- parallel_pt_loop, src/pt.rs:414-457.
 * threads <= 1: one stream for everything, in pt_loop order (src/pt.rs:459-504). */
tamc_status tamc_pt_run_lockstep(const tamc_instance *inst, const tamc_pt_params *params, const int8_t *init_states,
                                 uint32_t num_streams, const uint32_t *const *stream_words,
                                 const uint64_t *stream_lens, tamc_pt_result **out);
void tamc_pt_result_free(tamc_pt_result *r);

/* Vec<PTState> (tamc-core/src/pt.rs:18-25) as a run's starting point, for PtIcmRunner::run(Some(states))
 * (src/pt.rs:397-408), which continues the given states, counters and round-trip tags.  Every array
 * is in slot order ([chain][beta slot]); the final_* fields of a tamc_pt_result are exactly these
 * arrays, so one run's result resumes in the next.  A null counter array means the value
 * PTState::new gives (tamc-core/src/pt.rs:27-38): zeros, and tags MinBeta at slot 0, MaxBeta at
 * slot B-1, None elsewhere. */
typedef struct {
  const int8_t *states;            /* [C][B][n] spins (+1/-1): PTState.states[b].arr */
  const float *energies;           /* [C][B] cached IsingState energies; null = energy_ref of each state */
  const uint32_t *num_acceptances; /* [C][B] */
  const uint32_t *round_trips;     /* [C] */
  const uint8_t *tags;             /* [C][B] PTRoundTrip ordinals (0..4, as tamc_pt_result.tags) */
  const uint32_t *diffusion_hist;  /* [C][B][2] */
} tamc_pt_state;

/* PtIcmRunner::run(Some(states)) (src/pt.rs:397-408): tamc_pt_run from a full PTState.  acceptance
 * counts in the result include the imported num_acceptances, as count_acc (src/pt.rs:585-591)
 * sums the carried-over counters. */
tamc_status tamc_pt_run_resume(const tamc_instance *inst, const tamc_pt_params *params, const tamc_pt_state *state,
                               tamc_pt_result **out);

/* ------------------------------------------------------------ batched PT ---- */
typedef struct {
  uint32_t waves_per_group;  /* 0 = chosen by the autotuner when `autotune` is set, else from the
                                instances' level widths */
  uint32_t sweeps_per_launch;/* 0 = all sweeps in one launch */
  int32_t use_graph;         /* capture the launch sequence in a hipGraph */
  int32_t autotune;          /* choose every setting left at 0 by timing the candidates on a 24-sweep
                                prefix of the same batch (measurements off) and keeping the fastest;
                                choices are cached per batch shape.  Candidates: 4, 8 and 16 waves per
                                group x both task tilings, then both ICM growth orders, then teams
                                against none, then one launch for all sweeps against four launches
                                of a quarter each (captured in a hipGraph) */
  uint32_t task_tiling;      /* how a colour level's tasks (4-site RNG groups x plane words) are laid
                                over a group's waves: 1 = dealt to the waves in turn, 2 = one
                                contiguous run per wave; 0 = tuned with `autotune`, else 1.  Results
                                do not depend on it */
  uint32_t icm_growth;       /* how ICM clusters grow: 1 = each thread joins a fixed stride of sites
                                per round, 2 = a wave closes a compact 64-site tile of the coupling
                                graph at a time; 0 = tuned with `autotune`, else tiles for instances
                                whose rows hold more than 4 couplings or 4 couplings on at most 4096
                                sites, strides otherwise.  The clusters do not depend on it */
  uint32_t teams;            /* a large group's Metropolis sweep shared by a team of work groups of
                                one launch, all resident at once (small groups then share work
                                groups, one after another): 0 = where a cost model predicts a
                                shorter launch (with `autotune`: kept only if it times faster),
                                1 = never, K >= 2 = every eligible group as a team of K (at most
                                16).  Results do not depend on it */
} tamc_batch_options;

typedef struct {
  double device_ms;          /* HIP-event time of all sweep launches */
  double wall_ms;
  uint64_t attempts;
  uint32_t waves_per_group, sweeps_per_launch, launches, groups;
  double phase_ms[7];        /* slowest work group: ICM overlap words, ICM seeds, ICM cluster growth,
                                ICM flip, exchange, sweep, measurement time */
  double icm_passes;         /* slowest work group: ICM cluster-growth passes, summed over sweeps */
  double prep_ms;            /* host time from the call to the first sweep launch (planning,
                                allocation, uploads, state initialisation) */
  double post_ms;            /* host time after the last sweep (downloads, result assembly) */
  double groups_busy_ms;     /* sum over work groups of their own phase time: divided by
                                device_ms times the compute units it gives how busy the mean
                                compute unit was (one group per compute unit) */
  uint32_t task_tiling;      /* tiling used (1 dealt in turn, 2 contiguous runs) */
  int32_t graph;             /* 1 when the launches ran as a captured hipGraph */
  uint32_t icm_growth;       /* ICM growth used (1 strided, 2 compact tiles, 0 the per-instance
                                default) */
  uint32_t teams;            /* groups that ran as teams of work groups */
  uint32_t work_groups;      /* work groups of the sweep launch */
} tamc_batch_stats;

/* Many independent PT-ICM solves (one per instance) in one persistent kernel; each result is
 * identical to tamc_pt_run on the same inputs.  results[count] receives one result per solve. */
tamc_status tamc_pt_run_batch(uint32_t count, const tamc_instance *const *insts, const tamc_pt_params *params,
                              const tamc_batch_options *opts, tamc_pt_result **results, tamc_batch_stats *stats);
/* tamc_pt_run_batch with initial states: init_states[k] is null or [C][B][n] spins (+-1) of solve k,
 * as for tamc_pt_run (src/pt.rs:397-408). */
tamc_status tamc_pt_run_batch_init(uint32_t count, const tamc_instance *const *insts, const tamc_pt_params *params,
                                   const int8_t *const *init_states, const tamc_batch_options *opts,
                                   tamc_pt_result **results, tamc_batch_stats *stats);
/* The plain path: the same solves, one instance per kernel launch, one after another. */
tamc_status tamc_pt_run_plain(uint32_t count, const tamc_instance *const *insts, const tamc_pt_params *params,
                              tamc_pt_result **results, tamc_batch_stats *stats);

/* ------------------------------------------------------------- serving ---- */
/* A request loop with continuous batching.  The reference has no server: its callers run many
 * PtIcmRunner solves concurrently from rayon (src/gla.rs:152-162, src/bin/tamc-pt-opt.rs:68-71);
 * a server takes the place of those concurrent runner calls.  Requests are admitted while device
 * memory (a pool reserved at create) and work-group slots allow; admitted requests stay resident
 * on the device and advance in time slices that all resident requests share, so requests that
 * arrive while others run join at the next slice.  Each result is identical to tamc_pt_run on
 * the same (instance, params, seed).  All functions are thread-safe; the instance passed to
 * tamc_server_submit must stay alive until the request has finished or been cancelled, and it is
 * immutable meanwhile: tamc_instance_set_suscept and tamc_instance_add_suscept_file fail with
 * TAMC_E_INVALID_ARG while a request on it is queued or resident. */
typedef struct tamc_server tamc_server;

typedef struct {
  uint64_t pool_bytes;       /* device memory reserved for request state (0 = 4 GiB) */
  uint32_t max_wave_slots;   /* waves of resident work groups, summed (0 = 16 per compute unit) */
  double slice_ms;           /* target device time of one slice (0 = 4 ms) */
  uint32_t waves_per_group;  /* 0 = per request from its level width (4, 8 or 16) */
} tamc_server_options;

typedef struct {
  double queued_ms;          /* submit -> admitted */
  double total_ms;           /* submit -> result ready on the host */
  uint32_t slices;           /* slices the request took part in */
} tamc_request_times;

typedef struct {
  uint64_t submitted, completed, failed, cancelled;
  uint32_t queued, resident, resident_waves;
  uint64_t slices;
  double device_busy_ms;     /* sum of slice device times */
  uint64_t attempts;         /* Metropolis attempts of completed requests */
  uint64_t pool_used, pool_bytes;
  double slice_wall_ms;      /* sum of slice wall-clock times (launch to the last stream's end) */
  double complete_ms;        /* host time downloading and assembling finished requests */
} tamc_server_stats;

tamc_status tamc_server_create(const tamc_server_options *opts, tamc_server **out);
/* Validates the request like tamc_pt_run and queues it; *ticket identifies it. */
tamc_status tamc_server_submit(tamc_server *srv, const tamc_instance *inst, const tamc_pt_params *params,
                               const int8_t *init_states, uint64_t *ticket);
/* wait_ms: 0 = do not wait, < 0 = wait until done.  TAMC_OK and *out = the result once the request
 * has finished (the ticket is then released; free the result with tamc_pt_result_free);
 * TAMC_E_PENDING while it runs; the request's own error if it failed. */
tamc_status tamc_server_poll(tamc_server *srv, uint64_t ticket, int32_t wait_ms, tamc_pt_result **out,
                             tamc_request_times *times);
/* Drops a queued or resident request (its device memory is released at the next slice boundary). */
tamc_status tamc_server_cancel(tamc_server *srv, uint64_t ticket);
tamc_status tamc_server_get_stats(tamc_server *srv, tamc_server_stats *stats);
/* Cancels what is left, stops the loop and releases the device pool. */
void tamc_server_destroy(tamc_server *srv);

/* ---------------------------------------------------------------------- SA ---- */
/* SaParams (src/sa.rs:21-26) with BetaOptions resolved to the schedule array. */
typedef struct {
  uint32_t num_betas;
  const float *betas;   /* one sweep per entry, non-decreasing (src/sa.rs:101-104) */
  uint32_t num_replicas;
  uint32_t threads;     /* num_replicas > 1: replicas = threads*(num_replicas/threads) (src/sa.rs:147,161) */
  uint64_t seed;
} tamc_sa_params;

/* AnnealMinResults + final states (src/sa.rs:28-65, 110-165). */
typedef struct {
  uint32_t n, R;
  float min_energy;
  float *energies;      /* [R] energy_ref of the final states, reference order */
  int8_t *final_states; /* [R][n] */
  uint64_t attempts;
  double timing_us;
  double device_ms;
} tamc_sa_result;

/* SaRunner::run / run_parallel: sweeps in the reference's sequential site order, one replica per
 * lane, Philox uniforms keyed by (seed, sweep, site quad, replica). */
tamc_status tamc_sa_run(const tamc_instance *inst, const tamc_sa_params *params, tamc_sa_result **out);
/* count independent SA solves in one launch. */
tamc_status tamc_sa_run_batch(uint32_t count, const tamc_instance *const *insts, const tamc_sa_params *params,
                              tamc_sa_result **results, tamc_batch_stats *stats);
void tamc_sa_result_free(tamc_sa_result *r);

/* SA requests to a tamc_server (SaRunner::run / run_parallel, src/sa.rs:110-165), served beside its
 * PT-ICM requests: the restart waves of every resident SA solve advance by whole sweeps of their
 * schedules in each slice, and their spins stay on the device in between.  Each result is identical to tamc_sa_run on the
 * same (instance, params, seed).  Poll an SA ticket with tamc_server_poll_sa (tamc_server_poll
 * refuses it, and the other way round); free its result with tamc_sa_result_free. */
tamc_status tamc_server_submit_sa(tamc_server *srv, const tamc_instance *inst, const tamc_sa_params *params,
                                  uint64_t *ticket);
tamc_status tamc_server_poll_sa(tamc_server *srv, uint64_t ticket, int32_t wait_ms, tamc_sa_result **out,
                                tamc_request_times *times);

/* No counterpart in the reference: tamc_sa_run for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  It runs
 * the same kernel and gives the same results as tamc_sa_run, and calls no hipMalloc, hipFree or
 * other allocating routine: the outputs are caller device buffers, energies [R] floats and states
 * [R][n] int8 (+1/-1), with R the restarts tamc_sa_run keeps; the schedule's betas are staged in a
 * static device scratch of 512 KiB per GPU (at most 131072 betas, else TAMC_E_INVALID_ARG), which
 * every call on that GPU shares and which keeps the last call's contents.  The instance must
 * already be on the device (any earlier call that ran it uploads it), else TAMC_E_INVALID_ARG.
 * stream is a hipStream_t (null: the default stream); the work is queued on it and the call returns
 * without waiting.  Calls that overlap on one GPU race on the scratch.  The restarts' bit planes
 * stay in LDS, so the instance may have at most 20224 sites (else TAMC_E_UNSUPPORTED). */
tamc_status tamc_sa_run_noalloc(const tamc_instance *inst, const tamc_sa_params *params, float *energies,
                                int8_t *states, void *stream);

/* --------------------------------------------------------------- outputs ---- */
/* bincode-1 serialisation of PtIcmThermalSamples (src/pt.rs:116-125, written at src/pt.rs:652). */
tamc_status tamc_write_thermal_samples(const tamc_pt_result *r, const float *betas, uint8_t compression_level,
                                       const char *path);
/* bincode-1 serialisation of AnnealState (src/sa.rs:67-89) for the states whose energy is within
 * 1e-8 of the minimum (src/sa.rs:255-270). */
tamc_status tamc_write_anneal_state(const tamc_sa_result *r, const char *path);
/* The same structures as Python pickles (src/pt.rs:650 and src/sa.rs:265 write serde-pickle
 * when the sample file name ends in .pkl): protocol 3, structs as dicts keyed by field name, Vecs
 * as lists. */
tamc_status tamc_write_thermal_samples_pickle(const tamc_pt_result *r, const float *betas, uint8_t compression_level,
                                              const char *path);
tamc_status tamc_write_anneal_state_pickle(const tamc_sa_result *r, const char *path);

/* ------------------------------------------------------- tamc-pt-opt ---- */
/* One temperature-set update of pt_optimize_beta (src/bin/tamc-pt-opt.rs:57-167), in f32 with the
 * reference's operation order.  Inputs per instance k: dif_hists[k][nt][2], the runs' diffusion
 * histograms summed over chains, and round_trips[k], summed over chains.  momentum[nt] is updated
 * in place; new_betas[nt] receives the next set; err the mean |log10| change; taus[k] the round-trip
 * times.  *updated = 0 when the eta normalisation is below f32 epsilon: the reference then keeps
 * the set (new_betas = beta_vec) and leaves the momentum unchanged.  Fails with
 * TAMC_E_INVALID_ARG where the reference's bisection panics. */
tamc_status tamc_pt_opt_update(uint32_t nt, const float *beta_vec, float *momentum, uint32_t count,
                               const uint32_t *dif_hists, const uint32_t *round_trips, uint32_t num_sweeps,
                               uint32_t num_chains, float alpha, float m, float *new_betas, float *err,
                               float *taus, int *updated);

/* ----------------------------------------------------------- percolation ---- */
/* Site percolation (random_percolation, src/percolation.rs:11-75; the percolate binary,
 * src/bin/percolate.rs): num_samples independent samples on the graph of an instance read in Ising
 * form (edges = couplings, node weights = diagonal entries).  In sample k site i is occupied with
 * probability p: rand's Bernoulli rule (v < p * 2^64, always for p = 1) on a 64-bit draw from
 * Philox4x32-10 keyed by seed, counter {k, i, 0, 4}.  max_cluster[k] = size of the largest
 * connected cluster of occupied sites (0 when none is occupied, where the reference panics);
 * percolates[k] = 1 when an occupied site of negative weight is connected to an occupied site of
 * positive weight (the reference's depth-first search finds a path).  Any number of sites (above
 * 16384 the labels live in device scratch, samples run in batches). */
tamc_status tamc_percolate(const tamc_instance *inst, double p, uint32_t num_samples, uint64_t seed,
                           uint32_t *max_cluster, uint8_t *percolates);

/* ------------------------------------------------------------------- GLA ---- */
/* Graph partitions for GLA (GlaParams::read_graph_partition, src/gla.rs:35-59): one partition per
 * line of whitespace-separated site indices (read_u32_lines, src/util.rs:142-172: a line that does
 * not start with an index is skipped; reading stops at the first token that is not one).  Fails
 * with TAMC_E_INVALID_ARG on an index >= n or an index listed twice, as the reference does.
 * Partition p holds sites[offsets[p] .. offsets[p+1]), sorted ascending (src/gla.rs:88-91). */
typedef struct {
  uint32_t num_partitions;
  uint32_t *offsets; /* [num_partitions + 1] */
  uint32_t *sites;   /* [offsets[num_partitions]] */
} tamc_partition;
tamc_status tamc_partition_from_file(const tamc_instance *inst, const char *path, tamc_partition **out);
/* The same checks on caller arrays (partition p = sites[offsets[p] .. offsets[p+1]), any order). */
tamc_status tamc_partition_create(const tamc_instance *inst, uint32_t num_partitions, const uint32_t *offsets,
                                  const uint32_t *sites, tamc_partition **out);
void tamc_partition_free(tamc_partition *p);

/* run_gla (src/gla.rs:122-196): every partition's induced sub-instance is solved by PT-ICM with the
 * same parameters, all in one batched launch (sub-solve k uses seed params->seed + k); each
 * sub-solve's last ground state is taken, and the couplings between partitions are summed over the
 * edges of the partition graph (src/gla.rs:73-120, 175-190), in the reference's edge order.
 * Mirrors GlaResults (src/gla.rs:62-70); final_state stays empty as in the reference, the joined
 * state is returned beside it. */
typedef struct {
  uint32_t instance_size, num_partitions, num_boundaries;
  float *partition_energies;  /* [num_partitions] last ground-state energy of each sub-solve */
  uint32_t *boundary_pairs;   /* [num_boundaries][2] (source partition, target partition) */
  float *boundary_energies;   /* [num_boundaries] sum of K s_i s_j over the edges between them */
  float bulk_energy;          /* sum of partition_energies (f32, in order) */
  float boundary_energy;      /* sum of boundary_energies (f32, in order) */
  float total_energy;         /* bulk_energy + boundary_energy */
  int8_t *joined_state;       /* [instance_size] the sub-solves' ground states joined (+-1) */
  uint64_t attempts;          /* Metropolis attempts of all sub-solves */
  double device_ms;           /* kernel time of the batched launch */
} tamc_gla_result;
tamc_status tamc_gla_run(const tamc_instance *inst, const tamc_partition *parts, const tamc_pt_params *params,
                         tamc_gla_result **out);
void tamc_gla_result_free(tamc_gla_result *r);
/* GlaResults as write_data writes it (src/util.rs:26-36): pickle when the path ends in .pkl, else
 * bincode-1. */
tamc_status tamc_write_gla_results(const tamc_gla_result *r, const char *path);

#ifdef __cplusplus
}
#endif
#endif
