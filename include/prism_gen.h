/*
 * prism_gen.h -- seeded synthetic AMBER-typed chain generator.
 *
 * No counterpart in the reference: the reference reads prepared proteins from
 * PrismPrepTopology JSON (crates/prism-nhs/src/input.rs:1027-1182).  This
 * generator emits exactly the fields the step consumes, in the layout of
 * prism_md.h, for the synthetic workloads: a self-avoiding heavy-atom chain
 * (1.53 A bonds, ~111 degree angles) confined to a sphere at 0.1 atoms/A^3,
 * 1-2 hydrogens per heavy atom at 1.09 A, hydrogen mass repartitioning
 * (input.rs:2482-2565), uniform random force-field parameters, exact scaled
 * 1-4 exceptions (1/2 LJ, 1/1.2 Coulomb).
 */
#ifndef PRISM_GEN_H
#define PRISM_GEN_H

#include "prism_md.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct PrismGenParamsV1 {
    int32_t  n_atoms;          /* total atoms, heavy + hydrogen (>= 3)          */
    uint64_t seed;             /* generator seed                                */
    double   density;          /* atoms / A^3 inside the confining sphere (0.1) */
    double   heavy_bond_A;     /* 1.53                                          */
    double   h_bond_A;         /* 1.09                                          */
    double   heavy_angle_deg;  /* 111                                           */
    double   hmr_factor;       /* 3.0 (H 1.008 -> 3.024); 1.0 disables HMR      */
    double   charge_max;       /* 0.8: q ~ U(-0.8, 0.8) then neutralised        */
    double   sigma_min, sigma_max;     /* 1.0, 3.5 A                            */
    double   epsilon_min, epsilon_max; /* 0.01, 0.2 kcal/mol                    */
    double   bond_k_min, bond_k_max;   /* 300, 500 kcal/mol/A^2                 */
    double   angle_k_min, angle_k_max; /* 40, 80 kcal/mol/rad^2                 */
    double   torsion_vn_max;   /* 2.0 kcal/mol: V_n ~ U(0, 2), k = V_n/2        */
    double   scale14_lj;       /* 0.5                                           */
    double   scale14_qq;       /* 1/1.2                                         */
    int32_t  n_candidates;     /* torsion candidates per growth step (24)       */
    int32_t  reserved[7];
} PrismGenParamsV1;

typedef struct PrismGenStatsV1 {
    int32_t n_heavy;
    int32_t n_hydrogen;
    double  sphere_radius_A;
    double  min_nonbonded_distance_A; /* over pairs beyond 1-4, capped at 4 A  */
    double  mean_nearest_nonbonded_A;
    double  total_charge;
    double  total_mass;
} PrismGenStatsV1;

typedef struct PrismGenHandle PrismGenHandle;

int prism_gen_default_params(PrismGenParamsV1 *params);
int prism_gen_create(const PrismGenParamsV1 *params, PrismGenHandle **out);
/* Pointer into the generator's storage; valid until prism_gen_destroy. */
const PrismMdTopologyV1 *prism_gen_topology(const PrismGenHandle *g);
int prism_gen_stats(const PrismGenHandle *g, PrismGenStatsV1 *stats);
int prism_gen_destroy(PrismGenHandle *g);

/* Maxwell-Boltzmann velocities at temperature T (A/ps), sigma =
 * sqrt(418.4 kB T / m), centre-of-mass velocity removed, seeded.
 * Same role as generate_maxwell_boltzmann_velocities, fused_engine.rs:16066-16116
 * (the reference's StdRng stream is not reproduced). */
int prism_gen_maxwell_boltzmann(const float *masses, int32_t n_atoms, double temperature_K,
                                uint64_t seed, float *v_out);

#ifdef __cplusplus
}
#endif
#endif /* PRISM_GEN_H */
