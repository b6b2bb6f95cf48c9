// kernels.hpp -- device pointer bundle and launch wrappers of the MD kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/prism_md.h"
#include "md_layout.hpp"

namespace prism {

struct DevPtrs {
    // atoms, arena (topology) order
    float4 *x, *v, *xs, *x_build;
    long long *f;                   // [A][4] fixed-point force accumulators
    const float *mass, *charge, *sigma, *eps;
    const int *atom_sys;            // system of each arena atom (-1 padding)
    // sorted slots (same arena index space)
    int *perm;                      // slot -> atom, -1 padding
    int *slot_of;                   // atom -> slot
    float4 *xq;                     // sorted (x, y, z, q sqrt(332.0636))
    float2 *lj;                     // sorted (sigma/2, sqrt(24 eps))
    // blocks
    float4 *bb_c, *bb_h;            // build-time box centre / half extent
    float4 *qbox;                   // [B][QBOX_PER_BLOCK][2] current boxes of each block's quarters and of
                                    // the block: centre, half extent
    const int *block_sys;           // system of each block (-1 padding)
    int *row_n, *row_j, *row_m;     // [B], [B][MAX_ROW], [B][MAX_ROW]
    int *unit_count;                // [B] nonbonded work units of each block's row
    int *unit_off;                  // [B] first unit of each block in the unit list
    int *units;                     // [B * UNITS_PER_BLOCK] (block << 11 | quarter << 9 | first tile)
    float4 *unit_f;                 // [B * UNITS_PER_BLOCK][SUB] per-unit i-forces (x, y, z, energy)
    // bonded contributions: one slot per (term, atom position), 3 FP32 each,
    // in one of two layouts (term_major):
    //   0  atom-major: the slots of atom a are [inc_off[a], inc_off[a+1]) and
    //      dst_* give each term position's slot -- the gather reads contiguous
    //      slots, the terms' stores scatter;
    //   1  term-major: bonds, angles, dihedrals, exceptions, each term's
    //      positions adjacent, so a wave of terms stores one contiguous range;
    //      inc_idx[inc_off[a] .. inc_off[a+1]) lists atom a's slots.
    // Large arenas use term-major (their bonded stores are bandwidth-bound),
    // small ones atom-major (their gather is latency-bound).  The forces are
    // the same integers either way.
    float *contrib;                 // [n_slots][3]
    const int *inc_off;             // [A + 1]
    const int *dst_bond, *dst_angle, *dst_dih, *dst_exc;  // atom-major slot per term position: [n][arity]
    const int *inc_idx;             // term-major: [n_slots] slots of each atom
    int term_major;
    // per TERM_CHUNK consecutive terms (bonds, angles, dihedrals, exceptions
    // concatenated): 1 when one of them belongs to an active system.  The
    // bonded threads of other chunks return at once -- a serving pool's
    // arena is mostly free slots and padding.
    const unsigned char *term_live;
    // 1 for a serving pool: most chunks are free there, so a bonded thread
    // tests its chunk's flag before loading its record; elsewhere both load
    // together (one wait instead of two for a live term)
    int terms_sparse;
    unsigned *hs_group;             // [second-half workgroups / HS_GROUP] step-end handshake counters
    double *term_e;                 // per-term energy (bonds, angles, dihedrals, exceptions)
    double *atom_e;                 // per-atom nonbonded energy of the last gather
    unsigned long long *masks;      // [B][MAX_MASKS][64] per partner block: skip bits of each i-lane
    int *pblk;                      // [B][MAX_MASKS] partner blocks of each block, ascending
    int *n_pblk;                    // [B]
    // partner lists over arena atoms
    const int *skip_off, *skip_idx; // exclusions U exception partners (nonbonded skip)
    const int *ex_off, *ex_idx;     // exclusions minus exception partners (pair list)
    // constraint groups
    const Group *groups;
    const GroupParam *gparams;
    const int *wave_sys;            // system of each group wave (-1 padding)
    // terms
    const Bond *bonds;
    const Angle *angles;
    const Dihedral *dih;
    const Exception14 *exc;
    int n_bonds, n_angles, n_dih, n_exc;
    // systems
    const SysDesc *sys;
    SysState *sst;
    int n_sys;
    int n_atoms_arena, n_blocks_arena, n_waves_arena;
    // scratch
    unsigned int *keys_a, *keys_b;
    int *vals_a, *vals_b;
    double *com_partial;            // [waves][4][3]: momentum sums of each 16-group quarter of a group wave
    double *work_g;                 // [waves * 64] per constraint group: cumulative non-Hamiltonian
                                    // work (bath, clamps), written only by the group's own lane
    int *sweeps;                    // [waves][2]: most SHAKE / RATTLE sweeps so far, per group wave
                                    // (written only by that wave; reduced on the host)
    EngineCounters *counters;
    unsigned *nb_claim;             // [NB_CLAIM_COUNTERS * NB_CLAIM_STRIDE] unit-claim counters of the
                                    // nonbonded kernel (zero between launches: each launch's last
                                    // claimer resets its counter)
    unsigned long long *step_ts;    // [TS_RING] device wall clock at the end of step k (slot k % TS_RING)
    int rows_in_nonbonded;          // 1: the step's nonbonded launch builds the rows (single-system engines)
    int *rows_total;                // [B] unit totals of the row-building workgroups of that launch
};

struct RefGrid {                    // reference cell grid of the pair-list parity mode
    float ox, oy, oz;
    int nx, ny, nz;
};

// Rebuild of the systems whose rebuild flag is set: re-sort or box refresh,
// then rows and the unit list.  Inside a step the flags are consumed by
// k_second_half (lists built at x_(s+1) after SHAKE of step s); after a
// rebuild between steps launch_consume(step_offset = -1) does it (at x_s).
// with_bonded: the sort's launch also evaluates every bonded term and
// exception (all terms, no energies) into the contribution slots, as
// launch_bonded(P, PRISM_MD_TERM_ALL, false) does; the MD step uses it.
void launch_sort(const DevPtrs &P, bool with_bonded, hipStream_t s);
// launch_sort forms the quarter boxes (P.qbox) of every block whose system
// does not re-sort this step; the rows pass (launch_rows, or the step's
// nonbonded launch) those of a system it re-sorted; launch_quarter_boxes
// forms them alone.  launch_rows builds the rows of a rebuild between steps.
void launch_rows(const DevPtrs &P, hipStream_t s);
void launch_quarter_boxes(const DevPtrs &P, hipStream_t s);
void launch_consume(const DevPtrs &P, int step_offset, hipStream_t s);
// slot_of / xq / lj of system s's n_slots sorted slots from its uploaded perm
void launch_restore_slots(const DevPtrs &P, int s, int n_slots, hipStream_t st);
// start / stop: optional HIP events recorded with the launch, bracketing the
// kernel's execution (sampled timing).  boxes: form the quarter boxes first
// (false only where they are current, as for a repeated timing launch).
void launch_nonbonded(const DevPtrs &P, bool energy, hipStream_t s, hipEvent_t start = nullptr,
                      hipEvent_t stop = nullptr, bool boxes = true);
// The MD step's evaluation: nonbonded partials and the bonded contribution
// slots (written beforehand, see launch_sort) added straight into the force
// accumulators, which k_first_half zeroed; no gather follows.  The quarter
// boxes are the step's launch_sort's (and launch_rows').  With
// rows_in_nonbonded the launch first builds the rows and the unit list of a
// rebuild step (launch_rows' work, the list in block order) behind grid
// barriers, and the step launches no rows kernel.
void launch_nonbonded_direct(const DevPtrs &P, hipStream_t s, hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
void launch_bonded(const DevPtrs &P, unsigned term_mask, bool energy, hipStream_t s);
void launch_first_half(const DevPtrs &P, const float *noise, hipStream_t s);
void launch_second_half(const DevPtrs &P, const float *noise, hipStream_t s);
// The second half of a step and the first half of the next in one launch
// (no explicit noise); usable when seam_fits(P): every workgroup resident.
bool seam_fits(const DevPtrs &P);
void launch_seam(const DevPtrs &P, hipStream_t s);
void launch_zero_forces(const DevPtrs &P, hipStream_t s);
// Sum the nonbonded unit partials and the bonded slots of every atom, in fixed
// order, into the int64 fixed-point accumulator (one writer per atom).
void launch_gather(const DevPtrs &P, unsigned term_mask, bool energy, hipStream_t s);
// Per-system FP64 energy sums of the last evaluation into SysState.energy.
void launch_energy_reduce(const DevPtrs &P, hipStream_t s);
// ke_out[S][RED_STRIDE] = (KE, COM velocity, cumulative work); com_sys: system
// whose carried COM velocity is also set (-2 all, -1 none).
constexpr int RED_STRIDE = 5;
void launch_kinetic(const DevPtrs &P, double *ke_out, int com_sys, hipStream_t s);
void launch_emit_pairs(const DevPtrs &P, int mode, RefGrid g, int *out_i, int *out_j, unsigned long long *count,
                       unsigned long long cap, hipStream_t s);
void launch_count_pairs(const DevPtrs &P, unsigned long long *count, hipStream_t s);
int fire_partial_blocks(const DevPtrs &P);
void launch_fire_partials(const DevPtrs &P, const float4 *vmin, double *part, hipStream_t s);
void launch_fire_update(const DevPtrs &P, float4 *vmin, float cv, float cf, float dt, float max_step, hipStream_t s);
void launch_snap_hydrogens(const DevPtrs &P, hipStream_t s);

}  // namespace prism
