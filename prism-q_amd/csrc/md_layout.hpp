// md_layout.hpp -- device data layout shared by the host engine and the HIP
// kernels of the classical MD step.
//
// One engine holds S >= 1 independent systems packed into shared arenas:
//   atoms     [atom_begin, atom_begin + n_atoms) per system; atom_begin is a
//             multiple of BLOCK so the spatially sorted slots of a system
//             occupy whole 64-slot blocks [block_begin, block_begin + n_blocks);
//   groups    constraint groups (one heavy atom with its hydrogens, or a lone
//             atom), padded per system to whole 64-lane waves;
//   terms     bonds / angles / dihedrals / exceptions with arena atom indices;
//   tiles     per i-block a row of up to MAX_ROW j-blocks (the cluster-pair
//             list), with per-row exclusion bit masks.
// Every kernel is launched over a fixed arena extent and skips padding, so a
// captured graph stays valid while systems come and go.
#pragma once
#include <stdint.h>

namespace prism {

constexpr int BLOCK = 64;            // atoms per spatial block (= wavefront)
constexpr int MAX_ROW = 512;         // j-blocks per i-block row (full list)
constexpr int MAX_MASKS = 64;        // exclusion-mask tiles per row
constexpr int SUB = 16;              // i-atoms per nonbonded wave (block quarter)
constexpr int QUARTERS = BLOCK / SUB;
constexpr int QBOX_PER_BLOCK = QUARTERS + 1;  // current boxes per block: its quarters', then its own
constexpr int CHUNK_MIN = 8;         // tiles per nonbonded work unit: per system, in
constexpr int CHUNK_MAX = 32;        // [CHUNK_MIN, CHUNK_MAX] (SysDesc::chunk)
constexpr int CHUNKS_PER_ROW = MAX_ROW / CHUNK_MIN;
constexpr int UNITS_PER_BLOCK = QUARTERS * CHUNKS_PER_ROW;
constexpr int NB_CLAIM_GROUPS = 8;   // unit sub-lists per XCD range, each with its claim counter
constexpr int NB_CLAIM_COUNTERS = 8 * NB_CLAIM_GROUPS;
constexpr int NB_CLAIM_STRIDE = 32;  // counters 128 bytes apart
constexpr int FORCE_FIXED_BITS = 24; // int64 fixed point: 2^24 per kcal/mol/A
constexpr int ENERGY_FIXED_BITS = 24;
constexpr int TERM_CHUNK = 256;     // terms per live flag (term_live)
constexpr float DUMMY_FAR = 1.0e7f;  // padding slots sit far from every atom

constexpr double COULOMB = 332.0636;         // nhs_amber_fused.cu:66
constexpr double KCAL_TO_AMU_A2_PS2 = 418.4; // nhs_amber_fused.cu:69
constexpr double KB = 0.001987204;           // nhs_amber_fused.cu:70
constexpr float MAX_FORCE = 1000.0f;         // nhs_amber_fused.cu:71
constexpr float MAX_COORD = 1000.0f;         // divergence bound of heartbeat_check, housekeeping.cu:41

// status bits latched on the device (SysState.status)
enum : int {
    ST_CONSTRAINT = 1,   // SHAKE/RATTLE did not converge
    ST_NONFINITE = 2,    // NaN/Inf in state or forces
    ST_ROW_OVERFLOW = 4, // more than MAX_ROW tiles in a row
    ST_MASK_OVERFLOW = 8,
    ST_EXCEPTION_COINCIDENT = 16,
    ST_DIVERGED = 32,    // a coordinate beyond MAX_COORD
    ST_BARRIER = 64,     // a grid barrier (step seam, nonbonded rows) timed out
    // any of these halts the system's integration at the end of the step
    ST_HALT = ST_CONSTRAINT | ST_NONFINITE | ST_ROW_OVERFLOW | ST_MASK_OVERFLOW | ST_EXCEPTION_COINCIDENT | ST_DIVERGED |
              ST_BARRIER,
};

// Immutable per-system description (host-written).
struct SysDesc {
    int atom_begin, n_atoms;
    int block_begin, n_blocks;
    int wave_begin, n_waves;        // constraint-group waves
    int active;
    int rebuild_every;
    int com_every;
    int constraint_max_iter;
    int sort_bits;                  // max bits per axis of the Hilbert sort key
    int resort_every;               // full re-sort on every Nth rebuild
    float dt;
    float cutoff, cutoff2, list_cutoff;
    float half_skin2;               // (skin/2)^2 displacement trigger
    int chunk;                      // row tiles per nonbonded work unit (CHUNK_MIN..CHUNK_MAX)
    // derived on the host from the fields above, for the nonbonded kernel
    float inv_rc, inv_rc2;          // 1 / cutoff, 1 / cutoff^2
    float reach2;                   // (cutoff + 0.001)^2: bound of the box culls
    int pad_desc;
    double shake_tol, rattle_tol;
    double inv_total_mass;
    // schedule (protocol_director.cu:58-123)
    float t_start, t_end, gamma_base, gamma_eq;
    int cold_hold_end, ramp_end, warm_hold_end, ramp_down_end;
    int warm_tail, equilibration_steps;
    uint32_t seed_lo, stream;
    int noise_explicit;             // 1: read normals from the explicit array
    float sort_cell;                // Hilbert sort cell edge, A
    // term ranges of this system in the global term arrays (energy sums)
    int bond_begin, n_bonds, angle_begin, n_angles;
    int dih_begin, n_dih, exc_begin, n_exc;
    unsigned long long end_step;    // step count at which the system stops advancing (0: no limit)
};

// Mutable per-system state (device-written).
struct SysState {
    unsigned long long step;        // index of the step being executed
    long long rebuild_step;         // step whose positions the pair list was built at (-1: before step 0)
    int rebuild_flag;               // set by the integrator or the host, consumed at the end of the step
    int resort_flag;                // set by the host: the next rebuild re-sorts (positions replaced)
    int status;                     // ST_* bits
    int n_rebuilds;
    double com_v[3];                // centre-of-mass velocity for the next COM step
    double energy[8];               // FP64 term energies of the last evaluation (5 used)
    unsigned long long force_evals;
    int fault_atom;                 // first atom of the first failing group (-1: none)
    int fault_kind;                 // FAULT_* of that failure
    unsigned long long fault_step;
    int pad_sweeps[2];
    int halted;                     // set at the end of a step that latched an ST_HALT bit: the
                                    // integrator leaves the system as it was after that step
    int pad_state;
};

// failure kinds recorded with the first fault
enum : int {
    FAULT_RATTLE_1 = 1,     // RATTLE after the first OU half-step
    FAULT_RATTLE_2 = 2,     // RATTLE before the drift
    FAULT_SHAKE = 3,        // SHAKE of the drifted positions
    FAULT_RATTLE_3 = 4,     // RATTLE before the second OU half-step
    FAULT_RATTLE_4 = 5,     // final RATTLE
    FAULT_NONFINITE = 6,    // non-finite position or velocity
    FAULT_DIVERGED = 7,     // |x| > MAX_COORD in some coordinate after the drift
};

// A rebuild re-sorts the atoms into blocks when the host asked for it or on
// every resort_every-th rebuild; the other rebuilds keep the blocks and only
// refresh their boxes and the rows.
__host__ __device__ inline bool resort_due(const SysDesc &D, const SysState &S) {
    return S.resort_flag || D.resort_every <= 1 || S.n_rebuilds % D.resort_every == 0;
}

// A system that has reached its step limit keeps its state and does not
// advance; the integrator treats it like a halted one.
__host__ __device__ inline bool parked(const SysDesc &D, const SysState &S) {
    return D.end_step != 0 && S.step >= D.end_step;
}

// Global engine counters (device-written).
struct EngineCounters {
    unsigned int done_groups;       // step-end handshake of the integrator: groups of workgroups
                                    // done (low 16 bits) + 1 << 16 per group that published
    int any_dirty;                  // some system rebuilds its rows this step
    int n_units;                    // nonbonded work units claimed by the row builder
    int any_rebuild;                // some system raised its rebuild flag (cleared with any_dirty);
                                    // the rebuild kernels return at once while it is 0
    unsigned seam_gen;              // bumped by the step seam's grid barrier once per launch
    int pad_c;
    unsigned long long step_ends;   // steps completed by the engine (all systems together)
    // the nonbonded kernel's rows barrier (zeroed at every step end, where
    // any_dirty is cleared), each word on a cache line of its own: the waiting workgroups
    // poll nb_go, and neither the arrivals nor the unit claims above queue
    // behind those reads
    alignas(128) unsigned nb_arrive;  // row-building workgroups done
    alignas(128) unsigned nb_go;      // set by the last of them
};
constexpr int TS_RING = 4096;       // device wall-clock stamps of the most recent step ends

struct Group {                      // constraint group
    int atom[4];                    // central, h0, h1, h2 (-1 unused); atom[0] < 0 => padding
};

struct GroupParam {
    float d[3];                     // X-H bond lengths
    float inv_mass_c, inv_mass_h;
    int n_h;
    int pad[2];
};

struct Bond { int i, j; float r0, k; };
struct Angle { int i, j, k; float theta0, kf; int pad[3]; };
struct Dihedral { int i, j, k, l; int n; float cos_phase, sin_phase, kf; };  // phase as cos, sin (host, FP64)
struct Exception14 { int i, j; float qq, sigma, eps; int pad[3]; };

}  // namespace prism
