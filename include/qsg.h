/*
 * qsg.h -- C ABI of the MI355X (gfx950) particle-tracking path.
 *
 * One handle = one transport problem whose read-only tables (mesh planes,
 * cross sections) and particle/tally state live in device memory.  Each
 * entry point returns QSG_OK or an error code; qsg_last_error() gives the
 * message (thread-local).  A handle is not re-entrant; distinct handles are
 * independent.  Host buffers are caller-owned and only accessed during the
 * call.  No C++ exceptions cross this boundary.
 *
 * Reference interfaces replaced (paths relative to the reference's src/):
 *   qsg_create / qsg_create_from_deck   getParameters  Parameters.cc:80
 *                                       + initMC       initMC.cc:53
 *   qsg_cycle_init                      cycleInit      main.cc:96-121
 *   qsg_cycle_begin_host                the end of cycleInit (main.cc:117-120): the
 *                                       processing vaults the host filled
 *   qsg_cycle_tracking                  cycleTracking  main.cc:138-307
 *                                       (CycleTrackingGuts, CycleTracking.cc:15)
 *   qsg_cycle_finalize                  cycleFinalize  main.cc:310-324
 *                                       (Tallies::CycleFinalize, Tallies.cc:25)
 *   qsg_get_scalar_flux                 Tallies::_scalarFluxDomain (Tallies.hh:298)
 *   qsg_set_particles / qsg_get_census  ParticleVaultContainer processing /
 *                                       processed vaults (ParticleVaultContainer.hh:43-61)
 *   qsg_destroy                         MonteCarlo::~MonteCarlo  MonteCarlo.cc:113
 */
#ifndef QSG_H
#define QSG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QSG_OK 0
#define QSG_ERR_ARG 1      /* bad argument or malformed deck            */
#define QSG_ERR_DEVICE 2   /* HIP runtime failure                        */
#define QSG_ERR_CAPACITY 3 /* a particle buffer would overflow           */
#define QSG_ERR_STATE 4    /* call out of order (e.g. finalize twice)    */
#define QSG_ERR_NUMERIC 5  /* runaway history / production overflow      */

typedef struct qsg_problem_s* qsg_problem;

/* Balance tallies of one cycle (Tallies.hh:36-100 Balance). */
typedef struct qsg_balance {
    uint64_t start, source, rr, split;
    uint64_t absorb, scatter, fission, produce;
    uint64_t collision, escape, census, num_segments;
    uint64_t end;
} qsg_balance;

/* Size information of a problem. */
typedef struct qsg_info {
    int32_t nx, ny, nz, n_groups;
    int32_t n_materials, n_steps;
    uint64_t n_cells;
    uint64_t n_particles_target;
    uint64_t source_per_cycle;
    uint64_t particle_capacity;
    uint64_t device_bytes;
    double dt;
    double source_weight;
    int32_t cycle;      /* cycles completed so far */
    int32_t lds_tables; /* 1 if the cross-section tables are staged in LDS */
} qsg_info;

/* Particle state in structure-of-arrays form (host pointers, n entries). */
typedef struct qsg_particles {
    double *x, *y, *z, *vx, *vy, *vz;
    double *energy, *weight, *time_to_census, *num_mean_free_paths;
    uint64_t* seed;
    int32_t* cell; /* global cell id i + nx*(j + ny*k) */
} qsg_particles;

/* Kernel launch shape of the tracking pass (chosen by the autotuner). */
typedef struct qsg_tuning {
    int32_t block_size;   /* threads per workgroup, 64..256, multiple of 64 */
    int32_t waves_per_cu; /* persistent waves launched per compute unit    */
    int32_t refill_min;   /* idle lanes (1..64) that trigger a wave refill  */
    int32_t chunk;        /* queue entries a wave claims at a time          */
    int32_t stack_entries; /* per-wave secondary stack entries, 0 or 512..65536;
                              0 sends secondaries through extra passes      */
    int32_t sort_interval; /* sort the population by cell every n cycles
                              (0: never)                                     */
    int32_t waves_per_simd; /* register budget of the tracking kernels: 2, 3
                               or 4 resident waves per SIMD                  */
    int32_t graphs;        /* batched form: 1 replays each step's launches
                              from captured HIP graphs, 0 launches them one
                              by one                                         */
} qsg_tuning;

const char* qsg_last_error(void);

/* argv-style construction: same flags and deck grammar as the qs driver. */
int qsg_create(int argc, const char* const* argv, int device, qsg_problem* out);
/* Deck given as text (optional extra argv applied first, like the CLI). */
int qsg_create_from_deck(const char* deck_text, int argc, const char* const* argv, int device,
                         qsg_problem* out);

/* Per-problem sizing.  No reference counterpart: the reference's vaults grow
 * on the host.  particle_capacity is the number of entries of each of the
 * problem's particle buffers (processing, census, population-control output
 * and the two secondary buffers); 0 selects the default, 2 x nParticles + one
 * cycle's source + 65536.  A cycle that would need more stops with
 * QSG_ERR_CAPACITY and leaves the problem failed (see below). */
typedef struct qsg_options {
    uint64_t particle_capacity;
} qsg_options;
/* qsg_create_from_deck with options (opts may be NULL: the defaults). */
int qsg_create_from_deck_opts(const char* deck_text, int argc, const char* const* argv, int device,
                              const qsg_options* opts, qsg_problem* out);
int qsg_destroy(qsg_problem p);

int qsg_get_info(qsg_problem p, qsg_info* info);
int qsg_set_tuning(qsg_problem p, const qsg_tuning* t);
int qsg_get_tuning(qsg_problem p, qsg_tuning* t);

/* A problem whose cycle stopped part way (a device error flag: a buffer
 * overflow, a runaway history; or a HIP failure) is failed: its device state
 * is no longer a cycle boundary, so every later call that would advance or
 * read its particles returns QSG_ERR_STATE.  qsg_problem_error gives the
 * code and message it failed with (code 0: not failed).  Errors found before
 * any device work (arguments, call order, a population that cannot fit)
 * leave the problem usable. */
int qsg_problem_error(qsg_problem p, int* code, char* message, size_t message_len);

/* Per-cycle phases, in order. */
int qsg_cycle_init(qsg_problem p);
/* In place of qsg_cycle_init, for a caller that keeps the reference's own
 * cycleInit (MC_SourceNow, PopulationControl, RouletteLowWeightParticles on
 * the host, main.cc:96-121): the n given particles become this cycle's
 * processing population and no device source or population control runs.
 * Then qsg_cycle_tracking, qsg_get_census / qsg_get_scalar_flux and
 * qsg_cycle_finalize as usual; the cycle's balance reports start = n,
 * source = rr = split = 0 (the caller booked those) and the tracking
 * counters.  Particle cells are global ids (i + nx*(j + ny*k)). */
int qsg_cycle_begin_host(qsg_problem p, const qsg_particles* in, uint64_t n);
int qsg_cycle_tracking(qsg_problem p);
/* Writes this cycle's balance and scalar-flux sum, then zeroes the flux. */
int qsg_cycle_finalize(qsg_problem p, qsg_balance* balance, double* flux_sum);

/* Runs n cycles; per-cycle results into arrays of length n (may be NULL).
 * tracking_seconds (may be NULL) receives the device time of the tracking
 * phase of each cycle.  The same results as n calls of the three phases.
 * Each cycle is enqueued before the host waits for the previous one: the
 * set-up reads the carried census count on the device, so the device goes
 * from cycle to cycle without waiting for the host.  A cycle whose first
 * tracking pass leaves secondaries (a wave stack overflowed, or
 * stack_entries is 0) or reports an error makes every kernel enqueued after
 * it return at once; the host then finishes that cycle with further passes
 * and enqueues the next one again.  The arrays are complete when the call
 * returns (on an error, entries of the failing cycle may be unset). */
int qsg_run(qsg_problem p, int n_cycles, qsg_balance* balances, double* flux_sums,
            double* tracking_seconds);

/* Device timings accumulated since creation (or the last qsg_reset_stats),
 * measured with HIP events on the problem's stream. */
typedef struct qsg_stats {
    double track_kernel_seconds; /* sum over tracking-pass kernel launches    */
    double tracking_seconds;     /* cycle tracking phases, first to last pass
                                    (a batch: each whole step)               */
    double init_seconds;         /* source + population control; under
                                    qsg_run all device time between two
                                    cycles' tracking (counter copy, finalize
                                    and set-up)                              */
    double finalize_seconds;     /* fluence + flux reset (+ counter readback);
                                    under qsg_run the last cycle's only      */
    uint64_t track_launches;     /* tracking-pass kernel launches            */
    uint64_t segments;           /* segments tracked                         */
    uint64_t collisions, facet_crossings, particle_loads, cycles;
} qsg_stats;
int qsg_get_stats(qsg_problem p, qsg_stats* stats);
int qsg_reset_stats(qsg_problem p);

/* Scalar flux [n_cells][n_groups] accumulated so far in this cycle. */
int qsg_get_scalar_flux(qsg_problem p, double* out, uint64_t n);

/* Replace the processing population (after qsg_cycle_init, before
 * qsg_cycle_tracking).  Used to track a given set of particles. */
int qsg_set_particles(qsg_problem p, const qsg_particles* in, uint64_t n);
/* Number of census particles after qsg_cycle_tracking. */
int qsg_get_census_count(qsg_problem p, uint64_t* n);
int qsg_get_census(qsg_problem p, qsg_particles* out, uint64_t n);
/* Processing population after qsg_cycle_init. */
int qsg_get_processing_count(qsg_problem p, uint64_t* n);
int qsg_get_processing(qsg_problem p, qsg_particles* out, uint64_t n);

/* ---- Batched form ----
 * Many independent problems (on one device, each at a cycle boundary) advance
 * one cycle per qsg_batch_step through shared kernel launches.  Every problem
 * keeps its own buffers, tallies and RNG streams, so its results are
 * bit-identical to running it alone.  The problem handles stay valid and
 * continue from the batch's state; they must not be advanced on their own
 * while they belong to a batch.  No reference counterpart: the reference runs
 * one problem per process. */
typedef struct qsg_batch_s* qsg_batch;
int qsg_batch_create(qsg_problem* problems, int n, qsg_batch* out);
int qsg_batch_destroy(qsg_batch b);
/* balances / flux_sums: arrays of n (may be NULL); inactive problems get 0.
 * A problem whose cycle raises a device error flag (QSG_ERR_CAPACITY,
 * QSG_ERR_NUMERIC) is taken out of the batch and failed; every other problem
 * finishes the step as usual and its entries are valid.  The step then
 * returns the first such problem's code, and qsg_batch_problem_error names
 * each failed problem.  Any other failure (a HIP error) fails the whole
 * batch: its later steps return QSG_ERR_STATE. */
int qsg_batch_step(qsg_batch b, qsg_balance* balances, double* flux_sums);
/* code: 0, or the error a problem was taken out of the batch with. */
int qsg_batch_problem_error(qsg_batch b, int index, int* code);
/* A failed problem cannot be made active again (QSG_ERR_STATE). */
int qsg_batch_set_active(qsg_batch b, int index, int active);
int qsg_batch_set_tuning(qsg_batch b, const qsg_tuning* t);
int qsg_batch_get_stats(qsg_batch b, qsg_stats* stats);

/* ---- Serving loop ----
 * Solve requests (deck text) are queued, admitted in arrival order while their
 * device footprint fits hbm_budget_bytes (and at most max_active run, 0 = no
 * cap), kept resident on the device, advanced together one cycle per
 * qsg_server_step (requests take and free slots of one running batch), and retired
 * after their nSteps cycles.  A request's per-cycle tallies equal those of the
 * same deck run alone.  A request whose cycle fails (a buffer overflow, a
 * runaway history) is failed alone; the others continue.  No reference
 * counterpart.
 * Request states: 0 queued, 1 running, 2 done, 3 failed, 4 cancelled. */
typedef struct qsg_server_s* qsg_server;
typedef struct qsg_server_stats {
    uint64_t submitted, admitted, completed, rejected, steps, segments;
    uint64_t bytes_in_use, active, queued;
} qsg_server_stats;
int qsg_server_create(int device, uint64_t hbm_budget_bytes, int max_active, qsg_server* out);
int qsg_server_destroy(qsg_server s);
int qsg_server_submit(qsg_server s, const char* deck_text, uint64_t* request_id);
/* qsg_server_submit with per-request sizing (opts may be NULL). */
int qsg_server_submit_opts(qsg_server s, const char* deck_text, const qsg_options* opts, uint64_t* request_id);
int qsg_server_step(qsg_server s, int* n_active, int* n_queued);
int qsg_server_status(qsg_server s, uint64_t id, int* state, int* cycles_done);
/* The first n completed cycles of a request (also of a running or cancelled
 * one).  A failed request returns the code it failed with (the message in
 * qsg_last_error). */
int qsg_server_result(qsg_server s, uint64_t id, qsg_balance* balances, double* flux_sums, int n);
/* Stops a queued or running request between steps (every running request
 * is then at a cycle boundary): its slot and device memory are freed at once
 * and its state becomes 4; the cycles it completed stay readable. */
int qsg_server_cancel(qsg_server s, uint64_t id);
/* Forgets a request (a running one is cancelled first). */
int qsg_server_release(qsg_server s, uint64_t id);
int qsg_server_set_tuning(qsg_server s, const qsg_tuning* t);
int qsg_server_get_stats(qsg_server s, qsg_server_stats* stats);

/* CORAL2 correctness checks (coralBenchmarkCorrectness, main.cc:73 ->
 * CoralBenchmark.cc:17-228) over the balance tallies accumulated since
 * creation and, for the fluence test, the per-cell fluence accumulated on the
 * device when the deck sets coralBenchmark. */
typedef struct qsg_coral_report {
    int32_t benchmark;      /* the deck's coralBenchmark (1 or 2; 0: none)     */
    int32_t ratio_pass;     /* BalanceRatioTest      (CoralBenchmark.cc:46)  */
    int32_t event_pass;     /* BalanceEventTest      (CoralBenchmark.cc:119) */
    int32_t missing_pass;   /* MissingParticleTest   (CoralBenchmark.cc:152) */
    int32_t fluence_pass;   /* FluenceTest           (CoralBenchmark.cc:177) */
    double ratio_error;     /* largest of the six relative ratio errors      */
    double ratio_tolerance; /* 0.01 (problem 1) or 0.011 (problem 2)         */
    double event_ratio;     /* |facet crossings / collisions - 1|            */
    uint64_t gains, losses; /* start+source+produce+split, absorb+census+
                               escape+rr+fission, summed over all cycles      */
    double fluence_max_percent_diff;
} qsg_coral_report;
int qsg_coral_check(qsg_problem p, qsg_coral_report* report);
/* Census energy spectrum (EnergySpectrum, EnergySpectrum.cc:12-59): over all
 * cycles so far, the census particles of each cycle counted per energy group
 * (counts[G]: energies above eMax); energies receives the group bounds
 * (n_groups + 1 values, may be NULL).  n = n_groups + 1.  Kept only when the
 * command line or deck sets energySpectrum (-e); the qs driver writes it to
 * <energySpectrum>.dat at the end of the run, as gameOver does (main.cc:93).
 * (crossSectionsOut (-S) is written by qsg_create / qsg_create_from_deck,
 * as initMC does, initMC.cc:72.) */
int qsg_get_energy_spectrum(qsg_problem p, double* energies, uint64_t* counts, uint64_t n);

/* Cumulative per-cell fluence (n_cells values, global cell order); only for
 * decks that set coralBenchmark. */
int qsg_get_fluence(qsg_problem p, double* fluence, uint64_t n_cells);

/* The tracking kernels' nearest-facet search (MCT_Nearest_Facet, MCT.cc:87-137
 * with the retry of MCT.cc:487-536) on n given rays of problem p: xyz and dir
 * hold [n][3] positions and direction cosines, cell the global cell ids
 * (i + nx*(j + ny*k)).  Returns the distance to the nearest facet (clamped at
 * 0), the facet (0..23) and the position after any nudge toward the cell
 * centre.  A diagnostic: the device function is the one cycle tracking uses. */
int qsg_nearest_facet_probe(qsg_problem p, uint64_t n, const double* xyz, const double* dir, const int32_t* cell,
                            double* distance, int32_t* facet, double* xyz_out);

/* No counterpart in the reference.  qsg_nearest_facet_probe for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): the same kernel and results, with
 * the rays and results staged through one fixed device array per GPU
 * (5.5 MiB, held until the process ends) instead of a per-call allocation.
 * At most QSG_PROBE_NOALLOC_MAX_RAYS rays per call; more is QSG_ERR_ARG.
 * Calls on one GPU share its array and are serialised by the library (each
 * returns after its results are copied out). */
#define QSG_PROBE_NOALLOC_MAX_RAYS 65536
int qsg_nearest_facet_probe_noalloc(qsg_problem p, uint64_t n, const double* xyz, const double* dir,
                                    const int32_t* cell, double* distance, int32_t* facet, double* xyz_out);

/* Device FP64 builtins on host arrays (rounding checks).
 * op: 0 sqrt(a) 1 a/b 2 log(a) 3 sin(a) 4 cos(a) 5 a*b 6 (double)(uint64 bits of a)
 *     7 the sine and 8 the cosine of one sincos(a) call */
int qsg_math_probe(int op, const double* a, const double* b, double* out, uint64_t n, int device);

#ifdef __cplusplus
}
#endif

#endif /* QSG_H */
