/*
 * prism_topology.h -- PrismPrepTopology JSON ingest and export.
 *
 * Reads the reference's on-disk topology format, PrismPrepTopology
 * (crates/prism-nhs/src/input.rs:1027-1182, loaded by
 * PrismPrepTopology::load, input.rs:1970-1993), applies the same validation
 * (PrismPrepTopology::validate, input.rs:1995-2300), the harmonic lowering to
 * the U = k*delta^2 convention (lower_harmonic_parameters, input.rs:467-538,
 * with HarmonicParameterContract::validate_for_topology, input.rs:300-323) and
 * hydrogen mass repartitioning (apply_hmr, input.rs:2482-2565), and exposes
 * the result as the PrismMdTopologyV1 view that prism_md_create and
 * prism_md_batch_admit take.  Host code only; no device work.
 *
 * Fields of the JSON that the classical step does not read (residue and chain
 * metadata, aromatic targets, lambda-fragment, periodic-cell, water and
 * provenance contracts) are parsed as JSON and otherwise ignored.
 *
 * Every function returns 0 or a negative PRISM_MD_E_* code; the reason of the
 * last failure on the calling thread is available from prism_topo_last_error().
 */
#ifndef PRISM_TOPOLOGY_H
#define PRISM_TOPOLOGY_H

#include "prism_md.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Harmonic source conventions (HarmonicEnergyConvention, input.rs:258-275). */
#define PRISM_TOPO_CONVENTION_FROM_CONTRACT 0 /* require harmonic_parameter_contract      */
#define PRISM_TOPO_CONVENTION_AMBER_NO_HALF 1 /* U = k d^2: coefficients used as stored   */
#define PRISM_TOPO_CONVENTION_OPENMM_HALF   2 /* U = k d^2 / 2: coefficients halved       */

typedef struct PrismTopoLoadOptionsV1 {
    /* Source convention of the bond and angle coefficients.  FROM_CONTRACT
     * (the reference's production rule) refuses a topology without a valid
     * harmonic_parameter_contract; the other values are an explicit caller
     * statement used when the file carries none, and are refused when it
     * carries one that disagrees. */
    int32_t harmonic_convention;
    /* 1: repartition hydrogen masses unless the file records hmr_applied =
     * true (apply_hmr: only a verified prior application suppresses it). */
    int32_t apply_hmr;
    double  hmr_factor;         /* 3.0 (nhs_rt_full.rs:15928)                   */
    int32_t reserved[8];
} PrismTopoLoadOptionsV1;

typedef struct PrismTopoInfoV1 {
    int32_t n_atoms;
    int32_t n_residues;
    int32_t n_chains;
    int32_t hmr_in_file;        /* -1 key absent, 0 false, 1 true               */
    int32_t hmr_applied_here;   /* clusters repartitioned by this load          */
    int32_t hmr_skipped;        /* clusters left alone (central atom < 1 amu)   */
    int32_t harmonic_convention;/* PRISM_TOPO_CONVENTION_* that was applied      */
    int32_t has_exception_contract;
    double  recommended_timestep_fs; /* 0 when absent                            */
    int32_t reserved[8];
} PrismTopoInfoV1;

typedef struct PrismTopoHandle PrismTopoHandle;

int prism_topo_default_options(PrismTopoLoadOptionsV1 *opt);
/* Parse and lower a PrismPrepTopology JSON document from a file or a buffer. */
int prism_topo_load_json(const char *path, const PrismTopoLoadOptionsV1 *opt, PrismTopoHandle **out);
int prism_topo_parse_json(const char *text, size_t len, const PrismTopoLoadOptionsV1 *opt, PrismTopoHandle **out);
/* Pointer into the handle's storage; valid until prism_topo_destroy. */
const PrismMdTopologyV1 *prism_topo_view(const PrismTopoHandle *t);
int prism_topo_info(const PrismTopoHandle *t, PrismTopoInfoV1 *info);
int prism_topo_destroy(PrismTopoHandle *t);
const char *prism_topo_last_error(void);

/* Write a lowered topology (for example the synthetic generator's) as a
 * PrismPrepTopology JSON document: AMBER_NO_HALF harmonic contract, exact
 * exception contract with its pair and strict-exclusion digests
 * (nonbonded_exception_digest_sha256 / strict_exclusion_digest_sha256,
 * input.rs:1934-1968), hmr_applied as given.  Exclusion partners that are
 * also exceptions are written as exceptions only (the reference refuses an
 * exception that overlaps a strict exclusion, input.rs:2178-2183). */
int prism_topo_write_json(const PrismMdTopologyV1 *t, int32_t hmr_applied, const char *path);

/* SHA-256 (FIPS 180-4) of a byte buffer as 64 lower-case hex digits plus NUL;
 * exported for the digests above. */
int prism_topo_sha256_hex(const void *data, size_t len, char out_hex[65]);

#ifdef __cplusplus
}
#endif
#endif /* PRISM_TOPOLOGY_H */
