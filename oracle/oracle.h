/*
 * oracle.h -- CPU FP64 restatement of the reference's classical MD step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation: only tests, the smoke check and the benchmark's CPU-baseline
 * leg may load it.  No product code links or calls it.
 *
 * Every function restates a specific part of the reference
 * (crates/prism-gpu/src/kernels/nhs_amber_fused.cu, nhs_neighbor_warp.cuh,
 * protocol_director.cu, housekeeping.cu; crates/prism-nhs/src/fused_engine.rs)
 * in double precision; the citations are on the definitions in oracle.cpp.
 * Parity pins (what checks each function) are listed in oracle.cpp's header.
 */
#ifndef PRISM_ORACLE_H
#define PRISM_ORACLE_H

#include <stdint.h>
#include "../include/prism_md.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Forces F[3n] (kcal/mol/A) and per-term energies E[5] (bond, angle, dihedral,
 * nonbonded, exception) at positions x[3n] (double).  decide_f32 = 1 decides
 * cutoff membership from the FP32 r2 of the float-rounded positions, as the
 * reference does; 0 decides in FP64.  Returns 0 or a negative error. */
int oracle_forces(const PrismMdTopologyV1 *t, const double *x, unsigned term_mask, int decide_f32,
                  double *F, double *E);

/* Pair list.  mode 0 = reference cell scheme on the grid defined by the
 * positions x_grid (the creation-time coordinates), mode 1 = brute force.
 * Writes unique pairs i<j sorted by (i, j) when cap suffices; returns the
 * pair count, or a negative error (PRISM_MD_E_CAPACITY when a cell exceeds
 * 512 atoms, an atom reaches 1024 neighbours, or an axis needs > 32 cells). */
int64_t oracle_pairs(const PrismMdTopologyV1 *t, const float *x, const float *x_grid, int mode,
                     int32_t *out_i, int32_t *out_j, int64_t cap);

/* n_steps of constrained Langevin OVRVO from step index step0 (0-based, the
 * reference's timestep counter).  x, v: [3n] double, updated in place.
 * noise: [n_steps][2][n][3] standard normals or NULL for Philox noise.
 * work: accumulated non-Hamiltonian work, may be NULL: the kinetic-energy
 * change of the OU half-steps, the RATTLE after OU_N and the final RATTLE,
 * the force clamp and the velocity clamp (the reference's d_external_work).
 * Returns 0, PRISM_MD_E_CONSTRAINT or PRISM_MD_E_NONFINITE. */
int oracle_run(const PrismMdTopologyV1 *t, const PrismMdParamsV1 *p, double *x, double *v, int64_t step0,
               int64_t n_steps, const float *noise, double *work);

/* Potential (sum of the five terms, FP64 cutoff decision) and kinetic energy. */
int oracle_energies(const PrismMdTopologyV1 *t, const double *x, const double *v, double *pe, double *ke);

/* Philox4x32-10 (Salmon et al., SC'11). */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Three standard normals for (atom, step, purpose) of the Langevin stream. */
void oracle_langevin_normals(uint64_t seed, uint32_t stream, uint32_t atom, uint64_t step, uint32_t purpose,
                             double out[3]);

/* Temperature and friction the director assigns to schedule step `step`. */
void oracle_schedule(const PrismMdParamsV1 *p, uint32_t step, double *temperature, double *gamma, int *hold);

/* One SHAKE / RATTLE solve on a single cluster (positions and velocities of
 * the cluster's atoms indexed as in the full arrays).  Returns iterations used
 * or -1 on non-convergence / invalid geometry. */
int oracle_shake_cluster(const PrismMdTopologyV1 *t, int cluster, double *x, const double *x_start, double *v,
                         double dt, double tol, int max_iter);
int oracle_rattle_cluster(const PrismMdTopologyV1 *t, int cluster, const double *x, double *v, double tol,
                          int max_iter);

#ifdef __cplusplus
}
#endif
#endif
