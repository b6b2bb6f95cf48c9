/*
 * prism_md.h -- C ABI of the MI355X (gfx950) classical MD step.
 *
 * This is the drop-in boundary for the reference's fused classical step:
 * the Rust host engine `NhsAmberFusedEngine` (crates/prism-nhs/src/fused_engine.rs)
 * launching the phase-split CUDA kernels of
 * crates/prism-gpu/src/kernels/nhs_amber_fused.cu through cudarc.  Every entry
 * point below names the reference interface it replaces.
 *
 * Conventions (mirroring the reference's C-ABI helper libraries,
 * crates/prism-nhs/src/cuda/deterministic_reduce.cu:388-435):
 *   - every function returns int: 0 = OK, negative = error class (PRISM_MD_E_*);
 *   - preconditions are checked in argument order before any device work;
 *   - a handle owns all of its device memory, allocated in prism_md_create;
 *     stepping and force evaluation never allocate;
 *   - all device work is ordered on the handle's HIP stream;
 *   - one host thread per handle; different handles are independent;
 *   - device-side faults (constraint non-convergence, non-finite state,
 *     capacity overflow) are latched in a device status word and returned as an
 *     error code at the next synchronising call; there is no silent fallback and
 *     no device trap.
 *
 * Units are the reference's (nhs_amber_fused.cu:66-71): Angstrom, ps, kcal/mol,
 * amu, elementary charge; 418.4 converts kcal/mol/amu to A^2/ps^2,
 * kB = 0.001987204 kcal/mol/K, Coulomb constant 332.0636.
 */
#ifndef PRISM_MD_H
#define PRISM_MD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PRISM_MD_ABI_VERSION 2  /* 2: energies out[2] is the work ledger W */

/* ---- return codes ------------------------------------------------------ */
#define PRISM_MD_OK               0
#define PRISM_MD_E_INVALID       -1  /* bad argument / topology / parameter   */
#define PRISM_MD_E_CAPACITY      -2  /* cell, tile, list or batch capacity    */
#define PRISM_MD_E_HIP           -3  /* HIP runtime error                     */
#define PRISM_MD_E_NONFINITE     -4  /* NaN/Inf in positions, velocities, F   */
#define PRISM_MD_E_CONSTRAINT    -5  /* SHAKE/RATTLE did not converge         */
#define PRISM_MD_E_NOMEM         -6  /* device allocation failed at create    */
#define PRISM_MD_E_UNSUPPORTED   -7  /* feature not available in this build   */
#define PRISM_MD_E_NODEVICE      -8  /* no HIP device visible                 */
#define PRISM_MD_E_DIVERGED      -9  /* a coordinate beyond 1000 A (heartbeat_check,
                                        housekeeping.cu:27-45)                */

/* ---- force-term mask (prism_md_compute_forces) -------------------------- */
#define PRISM_MD_TERM_BOND       0x01u /* compute_bond_force        nhs_amber_fused.cu:354-386 */
#define PRISM_MD_TERM_ANGLE      0x02u /* compute_angle_force       nhs_amber_fused.cu:388-448 */
#define PRISM_MD_TERM_DIHEDRAL   0x04u /* compute_dihedral_force    nhs_amber_fused.cu:524-625 */
#define PRISM_MD_TERM_NONBONDED  0x08u /* shifted LJ+Coulomb        nhs_amber_fused.cu:3104-3313 */
#define PRISM_MD_TERM_EXCEPTION  0x10u /* exact 1-4 exceptions      nhs_amber_fused.cu:3316-3470 */
#define PRISM_MD_TERM_ALL        0x1Fu
/* energy slots in E_terms_out[PRISM_MD_N_TERMS]: bond, angle, dihedral, nonbonded, exception */
#define PRISM_MD_N_TERMS 5

/* ---- pair-list modes (prism_md_build_pairs) ------------------------------ */
/* Reference-compatible: the exact set of the reference's cell scheme --
 * 10 A cells, grid origin = initial bounding-box minimum - 10 A,
 * n = ceil((extent + 20)/10) per axis, truncating cell assignment with clamp,
 * 27-cell stencil, keep !(r2 > 144.0f) with r2 = (dx*dx + dy*dy) + dz*dz in
 * FP32 without contraction, 1-2/1-3 exclusions dropped, 1-4 partners kept.
 * (fused_engine.rs:12285-12330, nhs_amber_fused.cu:11690-11728,
 *  nhs_neighbor_warp.cuh:10-119, fused_engine.rs:16674). */
#define PRISM_MD_PAIRS_REFERENCE 0
/* Intent: every non-excluded pair with FP32 r2 <= 144.0f (brute force). */
#define PRISM_MD_PAIRS_BRUTE     1

/* ---- state transfer flags ------------------------------------------------ */
#define PRISM_MD_HOST_PTR   0
#define PRISM_MD_DEVICE_PTR 1

/* ---- topology ------------------------------------------------------------
 * Field meanings follow PrismPrepTopology (crates/prism-nhs/src/input.rs:1027-1182)
 * after the harmonic lowering to the U = k*delta^2 convention
 * (input.rs:467-538, called at fused_engine.rs:10882) and HMR
 * (input.rs:2482-2565).  All arrays are host pointers, read only during
 * prism_md_create / prism_md_batch_admit. */
typedef struct PrismMdTopologyV1 {
    int32_t        n_atoms;
    const float   *positions;              /* [3*n_atoms] A                        */
    const float   *masses;                 /* [n_atoms] amu                        */
    const float   *charges;                /* [n_atoms] e                          */
    const float   *sigma;                  /* [n_atoms] A      (LjParam.sigma)     */
    const float   *epsilon;                /* [n_atoms] kcal/mol (LjParam.epsilon) */

    int32_t        n_bonds;
    const int32_t *bond_atoms;             /* [2*n_bonds] i, j                     */
    const float   *bond_r0;                /* A                                    */
    const float   *bond_k;                 /* kcal/mol/A^2, U = k (r - r0)^2       */

    int32_t        n_angles;
    const int32_t *angle_atoms;            /* [3*n_angles] i, j (vertex), k        */
    const float   *angle_theta0;           /* rad                                  */
    const float   *angle_k;                /* kcal/mol/rad^2, U = k (t - t0)^2     */

    int32_t        n_dihedrals;
    const int32_t *dihedral_atoms;         /* [4*n_dihedrals] i, j, k, l           */
    const int32_t *dihedral_periodicity;   /* [n_dihedrals]                        */
    const float   *dihedral_phase;         /* rad                                  */
    const float   *dihedral_k;             /* kcal/mol, U = k (1 + cos(n phi - g)) */
    const uint8_t *dihedral_improper;      /* identity flag only; may be NULL      */

    /* Topological exclusions, full and symmetric, CSR over atoms: 1-2, 1-3 and
     * 1-4 partners (PrismPrepTopology.exclusions).  Partners that also appear
     * as exceptions are evaluated by the exception term, the rest are skipped
     * (fused_engine.rs:11718-11731). */
    const int32_t *exclusion_offsets;      /* [n_atoms + 1]                        */
    const int32_t *exclusion_atoms;        /* [exclusion_offsets[n_atoms]]         */

    /* Exact scaled 1-4 exceptions (NonbondedExceptionParam, input.rs:593-611):
     * no cutoff, no shift. */
    int32_t        n_exceptions;
    const int32_t *exception_atoms;        /* [2*n_exceptions]                     */
    const float   *exception_charge_product; /* e^2                                */
    const float   *exception_sigma;        /* A                                    */
    const float   *exception_epsilon;      /* kcal/mol                             */

    /* SHAKE/RATTLE hydrogen clusters (HCluster, input.rs:625-642): one heavy
     * atom and 1..3 hydrogens, each atom in at most one cluster. */
    int32_t        n_clusters;
    const int32_t *cluster_central;        /* [n_clusters]                         */
    const int32_t *cluster_hydrogens;      /* [3*n_clusters], -1 = unused slot     */
    const float   *cluster_bond_lengths;   /* [3*n_clusters] A                     */
    const int32_t *cluster_n_hydrogens;    /* [n_clusters]                         */
    const float   *cluster_inv_mass_central; /* [n_clusters] 1/amu                 */
    const float   *cluster_inv_mass_h;     /* [n_clusters] 1/amu                   */
} PrismMdTopologyV1;

/* ---- run parameters ---------------------------------------------------------
 * The reference keeps these in the device ProtocolState
 * (crates/prism-gpu/src/kernels/protocol_state.cuh) written by the
 * protocol_director kernel (protocol_director.cu:58-160).  Here they are host
 * parameters evaluated per step on the device from this descriptor. */
typedef struct PrismMdParamsV1 {
    double   dt_ps;               /* base timestep, ps (0.002, 0.004 with HMR)        */
    double   temperature_K;       /* start temperature of the schedule               */
    double   gamma_ps;            /* gamma_base, 1/ps; 0 selects exact NVE kicks      */
    double   cutoff_A;            /* potential-shift cutoff, 10 A                     */
    double   skin_A;              /* pair-list skin, 2 A (list cutoff 12 A)           */
    int32_t  rebuild_every;       /* forced rebuild interval in steps (20)            */
    int32_t  com_remove_every;    /* centre-of-mass velocity removal interval (100)   */
    uint64_t seed;                /* Philox4x32-10 key for the Langevin noise         */
    uint32_t stream_id;           /* second Philox key word (system / replica id)     */
    int32_t  constraint_max_iter; /* SHAKE/RATTLE iteration limit (100)               */
    double   shake_tol;           /* relative |r2-d2|/d2 tolerance (FP64 iteration)   */
    double   rattle_tol;          /* relative |r.v|/(|r| max(|v|,1)) tolerance        */
    /* Temperature/friction schedule (protocol_director.cu:58-123).  With all
     * step marks 0 the temperature is constant; with equilibration_steps 0
     * gamma is constant. */
    double   end_temperature_K;
    int32_t  cold_hold_end, ramp_end, warm_hold_end, ramp_down_end;
    int32_t  warm_tail;           /* 1: hold end_temperature after ramp_down_end      */
    int32_t  equilibration_steps; /* gamma ramp length                                */
    double   equilibration_gamma_ps; /* gamma at step 0 of the ramp                   */
    int32_t  resort_every;        /* re-sort atoms into blocks every Nth rebuild (8);
                                   * other rebuilds refresh block boxes and rows only */
    float    sort_cell_A;         /* edge of the Hilbert sort cells that order atoms into
                                   * 64-atom blocks (2.0; 0 selects 2.0)            */
    int32_t  chunk_tiles;         /* row tiles per nonbonded work unit, 8..32; 0 picks
                                   * 8 below 100,000 atoms, 24 above                */
    int32_t  reserved[5];
} PrismMdParamsV1;

typedef struct PrismMdStatsV1 {
    uint64_t steps_done;          /* completed MD steps                               */
    uint64_t pair_rebuilds;       /* cell-list / tile-list rebuilds performed          */
    uint64_t force_evaluations;   /* force evaluations performed                      */
    int32_t  n_tiles;             /* cluster-pair tiles in the current list           */
    int32_t  n_blocks;            /* 64-atom spatial blocks                           */
    int32_t  status;              /* latched device status bits (0 = clean)           */
    int32_t  max_shake_sweeps;    /* most SHAKE sweeps any cluster needed so far       */
    int32_t  max_rattle_sweeps;   /* most RATTLE sweeps any cluster needed so far      */
    int32_t  reserved;
} PrismMdStatsV1;

typedef struct PrismMdHandle PrismMdHandle;

/* Capacity of one slot of a batch or serving pool.  A system fits when
 * n_atoms < atoms and every other count is <= its capacity (exclusions: the
 * total length of the exclusion lists, both directions). */
typedef struct PrismMdSlotCapsV1 {
    int32_t atoms;
    int32_t bonds, angles, dihedrals, exceptions;
    int32_t exclusions;
    int32_t reserved[2];
} PrismMdSlotCapsV1;

/* Fills *params with the defaults: dt 0.004 ps (HMR), 300 K, gamma_base
 * 10/ps (fused_engine.rs:13314) with the equilibration friction ramp from
 * 1000/ps over the first 10000 steps (fused_engine.rs:15370-15378), cutoff
 * 10 A, skin 2 A, rebuild every 20 steps, re-sort every 8th rebuild, COM
 * removal every 100 steps, seed 0x50524953, 100 constraint iterations,
 * shake_tol 1e-10, rattle_tol 1e-12, constant temperature.
 * Replaces the constructor defaults of fused_engine.rs:13314-13366. */
int prism_md_default_params(PrismMdParamsV1 *params);

/* ABI version (PRISM_MD_ABI_VERSION).  No reference counterpart beyond the
 * 109-parameter kernel ABI pin, fused_engine.rs:7650. */
int prism_md_abi_version(void);

/* Build the engine: validate and lower the topology, upload it, allocate every
 * buffer, set Maxwell-Boltzmann velocities (COM-free) and build the first
 * pair list.  hip_stream may be NULL (a private stream is created).
 * Replaces NhsAmberFusedEngine::new / new_on_stream_with_nonbonded_policy,
 * fused_engine.rs:10771, :10871-13500, and init_rng_states,
 * nhs_amber_fused.cu:11308-11317. */
int prism_md_create(const PrismMdTopologyV1 *topology, const PrismMdParamsV1 *params,
                    void *hip_stream, PrismMdHandle **out);

/* Release every device buffer and the private stream, if any. */
int prism_md_destroy(PrismMdHandle *h);

/* Positions / velocities as [3*n_atoms] float arrays (host or device pointer
 * per flags).  Either pointer may be NULL to leave that field unchanged.
 * Setting positions invalidates the pair list and the carried force.
 * Replaces the engine's position/velocity upload and
 * get_positions / get_velocities, fused_engine.rs:24117, :20939. */
int prism_md_set_state(PrismMdHandle *h, const float *x, const float *v, int flags);
int prism_md_get_state(PrismMdHandle *h, float *x, float *v, int flags);

/* Rebuild the cell hash and tile list at the current positions, then
 * materialise the atom-pair list of the given mode for parity checks.  This
 * diagnostic call allocates device staging for the list (hipMalloc/hipFree
 * of 8 bytes per pair) on each call; prism_md_build_pairs_static does not.
 * Replaces rebuild_neighbor_lists, fused_engine.rs:16615-16748. */
int prism_md_build_pairs(PrismMdHandle *h, int mode);

/* prism_md_build_pairs without any device allocation, for callers that
 * cannot allocate (code that runs where hipMalloc is forbidden: tight serving
 * loops, harnesses that pre-plan all memory).  No counterpart in the
 * reference.  The pairs are staged in one static device array of the library
 * (a counter and 4,194,304 pairs, 32 MiB, held by each GPU the library runs
 * on until the process ends) instead of memory allocated per call; a list of
 * more pairs is refused with PRISM_MD_E_CAPACITY (n_out of get_pairs is then
 * 0).  prism_md_get_pairs reads the result as after prism_md_build_pairs.
 * Every call on one GPU shares that array, which keeps the last call's
 * contents: calls that overlap on one GPU (two host threads, two handles on
 * different streams) race.  Calls in sequence are safe; each clears the
 * counter it uses. */
int prism_md_build_pairs_static(PrismMdHandle *h, int mode);

/* Copy the pair list of the last prism_md_build_pairs as unique pairs i < j
 * (atom ids of the topology), sorted by (i, j).  *n_out receives the total
 * count; at most cap pairs are written.  Returns PRISM_MD_E_CAPACITY if
 * cap is too small (n_out is still set). */
int prism_md_get_pairs(PrismMdHandle *h, int32_t *i_out, int32_t *j_out,
                       size_t cap, size_t *n_out);

/* Evaluate the selected terms at the current positions.  F_out: [3*n_atoms]
 * host floats (may be NULL), the fixed-point accumulated force converted to
 * float; E_terms_out: [PRISM_MD_N_TERMS] host doubles (may be NULL).
 * Replaces the FORCE_N / FORCE_NP1 phases of prism_amber_phase_body_v1,
 * nhs_amber_fused.cu:2981-3484, and the per-atom PE sum of
 * energy_monitor.cu:24-99. */
int prism_md_compute_forces(PrismMdHandle *h, unsigned term_mask,
                            float *F_out, double *E_terms_out);

/* Same evaluation, raw 64-bit fixed-point force accumulators [3*n_atoms]
 * (scale 2^PRISM_MD_FORCE_FIXED_BITS per kcal/mol/A) for bitwise checks. */
#define PRISM_MD_FORCE_FIXED_BITS 24
int prism_md_compute_forces_fixed(PrismMdHandle *h, unsigned term_mask,
                                  int64_t *F_fixed_out);

/* Advance n_steps constrained Langevin OVRVO steps (phase order of
 * fused_engine.rs:6223-6235 plus the post-step clamp and final RATTLE,
 * fused_engine.rs:17498-17800).  noise == NULL draws the Langevin normals from
 * Philox4x32-10 keyed by (seed, stream_id); otherwise noise is a host array
 * [n_steps][2][n_atoms][3] of standard normals consumed by the first and the
 * second OU half-step of each step.
 * Replaces NhsAmberFusedEngine::step / run, fused_engine.rs:18963-19520,
 * :23911, and the captured replay of persistent_engine.rs:3207-3300. */
int prism_md_step(PrismMdHandle *h, uint64_t n_steps, const float *noise);

/* out[0] = potential energy PE (sum of all terms, kcal/mol), out[1] = kinetic
 * energy KE (kcal/mol), out[2] = W, the cumulative non-Hamiltonian work
 * (kcal/mol) since the system was created or admitted: the kinetic-energy
 * change of every Ornstein-Uhlenbeck half-step, of the RATTLE projections that
 * follow a bath update (after OU_N, and the final one), of the 1000 kcal/mol/A
 * kick clamp and of the velocity safety clamp, exactly the entries the
 * reference books into d_external_work per step (nhs_amber_fused.cu:1713-1788,
 * :2245-2270, :11148-11186; routing fused_engine.rs:19273-19277, :17698-17740)
 * and reduces with nhs_reduce_external_work_components (:11097-11126).  The
 * pre-drift RATTLE, SHAKE and pre-OU_NP1 RATTLE go to the reference's separate
 * constraint planes and are not in W.  Sums reduced deterministically in FP64.
 * Replaces prism_energy_monitor_* (energy_monitor.cu:24-99), the
 * nhs_diagnostic_phase_energy_v1 sums (nhs_amber_fused.cu:14840-14934) and the
 * W_ext window read (fused_engine.rs:10700-10708).  ABI version 1 returned the
 * kinetic temperature in out[2]; it is now prism_md_kinetic_temperature. */
int prism_md_energies(PrismMdHandle *h, double out[3]);

/* Kinetic temperature (K) over 3N - N_constraints - 3 degrees of freedom
 * (the reference's temperature diagnostic, nhs_amber_fused.cu:14840-14934). */
int prism_md_kinetic_temperature(PrismMdHandle *h, double *T);

int prism_md_get_stats(PrismMdHandle *h, PrismMdStatsV1 *stats);

/* FIRE energy minimisation (Bitzek et al., PRL 97, 170201 (2006)) of the
 * current positions with the engine's own force kernels, for at most
 * max_iter iterations or until max |F| < f_tol (kcal/mol/A); afterwards every
 * X-H bond is set to its constraint length and the pair list and carried
 * force are rebuilt.  Velocities are unchanged.  e_final / f_max may be NULL.
 * No direct counterpart: the reference relaxes unminimised inputs with its
 * equilibration friction ramp (protocol_director.cu:108-122). */
int prism_md_minimize(PrismMdHandle *h, int max_iter, double f_tol, double *e_final, double *f_max);

/* Human-readable description of the last error on this handle (never NULL);
 * with h == NULL, the reason of this thread's last failed prism_md_create. */
const char *prism_md_last_error(PrismMdHandle *h);

/* ---- diagnostics used by the benchmark ------------------------------------
 * Number of unique non-excluded, non-exception pairs with FP32 r2 < cutoff^2
 * at the current positions (the algorithmic pair count of the nonbonded
 * kernel). */
int prism_md_count_pairs(PrismMdHandle *h, uint64_t *n_pairs);

/* Launch the nonbonded force kernel `iters` times back to back on the handle's
 * stream at the current positions and report the mean launch duration in ms,
 * measured with HIP events on that stream. */
int prism_md_time_nonbonded(PrismMdHandle *h, int iters, double *mean_ms);

/* ---- batched ensemble ------------------------------------------------------
 * Many independent systems of uneven size in one engine: each phase of the
 * step is one launch over all systems, and batch_step replays a captured
 * hipGraph of graph_steps steps.  Every system keeps its own parameters,
 * Philox key (seed, stream_id), schedule, rebuild decisions and fault record;
 * its trajectory is bitwise the one it would follow alone.  The reference's
 * batching precedent is the legacy descriptor + flatten design
 * (crates/prism-gpu/src/kernels/amber_simd_batch.cu:8-22, :822-900); its
 * production engine runs one system per engine (replicas disabled,
 * fused_engine.rs:18307). */
typedef struct PrismMdBatch PrismMdBatch;
int prism_md_batch_create(const PrismMdTopologyV1 *const *topologies, const PrismMdParamsV1 *params,
                          int32_t n_systems, void *hip_stream, PrismMdBatch **out);
int prism_md_batch_destroy(PrismMdBatch *b);
int prism_md_batch_size(PrismMdBatch *b, int32_t *n_systems, int64_t *n_atoms);
/* n_steps steps of every active system; graph_steps > 0 captures that many
 * steps once into a hipGraph and replays it (the remainder runs eagerly).
 * A system whose step latches a fault (constraint, non-finite state,
 * capacity) halts at the end of that step and keeps its state there; the
 * other systems run all n_steps, and the call returns the fault of the first
 * active faulted system.  Retiring or deactivating the faulted system clears
 * the error for later calls (its stats keep the fault bits). */
int prism_md_batch_step(PrismMdBatch *b, uint64_t n_steps, int32_t graph_steps);
/* Joint FIRE minimisation of all active systems (see prism_md_minimize). */
int prism_md_batch_minimize(PrismMdBatch *b, int max_iter, double f_tol, double *f_max);
/* Inactive systems keep their state and are skipped by every kernel. */
int prism_md_batch_set_active(PrismMdBatch *b, int32_t system, int active);
int prism_md_batch_set_state(PrismMdBatch *b, int32_t system, const float *x, const float *v, int flags);
int prism_md_batch_get_state(PrismMdBatch *b, int32_t system, float *x, float *v, int flags);
int prism_md_batch_compute_forces(PrismMdBatch *b, int32_t system, unsigned term_mask, float *F_out,
                                  double *E_terms_out);
int prism_md_batch_energies(PrismMdBatch *b, int32_t system, double out[3]);
int prism_md_batch_kinetic_temperature(PrismMdBatch *b, int32_t system, double *T);
int prism_md_batch_get_stats(PrismMdBatch *b, int32_t system, PrismMdStatsV1 *stats);
const char *prism_md_batch_last_error(PrismMdBatch *b);

/* ---- serving pool ----------------------------------------------------------
 * A batch whose slots start empty: each slot reserves device memory for the
 * given capacities, so a pool is sized to an HBM budget up front.  Systems
 * enter a free slot with batch_admit (topology lowered into the slot's ranges,
 * Maxwell-Boltzmann velocities, own lists and carried force; no allocation,
 * the captured graph stays valid) and leave with batch_retire; coordinates,
 * velocities, lists and noise counters stay on the device in between.  A
 * system's trajectory does not depend on its slot or on the other systems.
 * Analogue of the reference's stream-ordered pool and batch scheduler
 * (crates/prism-nhs/src/vram_pool.rs:1-40, crates/prism-nhs/src/batch_scheduler.rs:1-16). */
int prism_md_slot_caps_for(const PrismMdTopologyV1 *t, PrismMdSlotCapsV1 *caps);
int prism_md_pool_create(const PrismMdSlotCapsV1 *caps, int32_t n_slots, void *hip_stream, PrismMdBatch **out);
int prism_md_batch_admit(PrismMdBatch *b, int32_t slot, const PrismMdTopologyV1 *t, const PrismMdParamsV1 *p);
int prism_md_batch_retire(PrismMdBatch *b, int32_t slot);
int prism_md_batch_slot_info(PrismMdBatch *b, int32_t slot, PrismMdSlotCapsV1 *caps, int32_t *occupied);
int prism_md_batch_device_bytes(PrismMdBatch *b, uint64_t *bytes);

/* ---- step limits and per-step latency ----------------------------------------
 * A system stops advancing when its step count reaches end_step (0 removes
 * the limit): later steps and graph replays leave it exactly as it is, so a
 * job of N steps ends at step N even when it shares replays of many steps
 * with other jobs.  Counterpart of the reference's per-run step budget
 * (NhsAmberFusedEngine::run(n), fused_engine.rs:23911).
 *
 * Per-step latency comes from the device: the integrator's last workgroup
 * stamps the end of every step with the device's constant-rate wall clock,
 * and a system's step latency is the time from the end of the previous step
 * to the end of its own, on the device clock -- including whatever ran
 * between them (other launches, admissions and retirements of other jobs,
 * host gaps).  Latencies are kept per system from its admission (its first
 * step, with no previous step of its own, is not counted); the most recent
 * 4,096 step ends are buffered on the device, so a single step call of more
 * steps keeps only its last 4,096.  No reference counterpart: the reference's
 * step timings are profiler receipts outside its tree (README.md:105-107). */
int prism_md_set_step_limit(PrismMdHandle *h, uint64_t end_step);
int prism_md_query_latency(PrismMdHandle *h, double *p50_ms, double *p99_ms, uint64_t *samples);
int prism_md_batch_set_step_limit(PrismMdBatch *b, int32_t system, uint64_t end_step);
int prism_md_batch_query_latency(PrismMdBatch *b, int32_t system, double *p50_ms, double *p99_ms,
                                 uint64_t *samples);
/* The latencies themselves (ms, in step order, at most cap); *n = how many exist. */
int prism_md_batch_latency_samples(PrismMdBatch *b, int32_t system, float *ms, size_t cap, size_t *n);

/* ---- checkpoint / restore ---------------------------------------------------
 * The dynamic state of a system -- positions, velocities, the positions and
 * block boxes of the last pair-list build, the sorted atom order, the step
 * counter and rebuild bookkeeping, the carried centre-of-mass velocity -- as a
 * self-describing host buffer.  The Langevin noise is counter-based (Philox
 * keyed by seed and stream_id, counter = atom and step), so no generator state
 * is stored.  Restoring into any engine or batch slot that holds the same
 * topology and run parameters (checked by digest; E_INVALID otherwise) rebuilds
 * the same tile rows and carried force, and the trajectory continues bitwise
 * as if it had not been interrupted.  save needs cap >= checkpoint_size.
 * Replaces capture_process_checkpoint_fields / restore_process_checkpoint_fields
 * (fused_engine.rs:24719, :25076). */
int prism_md_checkpoint_size(PrismMdHandle *h, size_t *bytes);
int prism_md_checkpoint_save(PrismMdHandle *h, void *buf, size_t cap, size_t *written);
int prism_md_checkpoint_restore(PrismMdHandle *h, const void *buf, size_t len);
int prism_md_batch_checkpoint_size(PrismMdBatch *b, int32_t system, size_t *bytes);
int prism_md_batch_checkpoint_save(PrismMdBatch *b, int32_t system, void *buf, size_t cap, size_t *written);
int prism_md_batch_checkpoint_restore(PrismMdBatch *b, int32_t system, const void *buf, size_t len);

/* Sample the nonbonded kernel inside prism_md_step: with every > 0, HIP
 * events on the handle's stream bracket the nonbonded launch of every
 * every-th step from now on (at most 1024 samples per prism_md_step call).
 * Each call returns the mean duration (ms) and count of the samples taken
 * since sampling was last (re)started; every = 0 stops sampling, and a call
 * with a changed interval restarts the count.  No reference counterpart (the
 * reference's timings are nsys receipts outside its tree, README.md:105-107). */
int prism_md_kernel_timing(PrismMdHandle *h, int every, double *mean_ms, uint64_t *samples);

#ifdef __cplusplus
}
#endif
#endif /* PRISM_MD_H */
