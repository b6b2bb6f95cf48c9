// rows.hpp -- the cluster-pair rows of the pair-list rebuild, one wave per
// i-block, in registers and global memory only (no LDS), so that any kernel
// can run it: k_rows, and the MD step's nonbonded kernel ahead of its units.
//
// Replaces the neighbour-row build of nhs_build_neighbor_list_warp
// (nhs_neighbor_warp.cuh:10-119) for the block-tile list of this build.
#pragma once
#include <climits>

#include "device_common.hpp"

namespace prism {

// Index of block q among a block's partner blocks (ascending; lane l holds
// the l-th, the n_pb valid ones first), -1 if absent.  A lower-bound search
// of seven fixed steps with every lane active: each lane searches its own q
// through cross-lane reads, which need all lanes to take part.
__device__ inline int pb_slot(int pb_reg, int n_pb, int q) {
    int lo = 0, hi = n_pb;
    for (int step = 0; step < 7; ++step) {
        const int mid = (lo + hi) >> 1;
        const int v = __shfl(pb_reg, min(mid, 63));
        if (lo < hi) {
            if (v < q) lo = mid + 1;
            else hi = mid;
        }
    }
    const int v = __shfl(pb_reg, min(lo, 63));  // outside the test below: every lane reads
    return lo < n_pb && v == q ? lo : -1;
}

// Up to SKIP_BATCH skip partners of one atom at once, e .. e + SKIP_BATCH - 1
// below e1: their sorted slots (-1 past e1).  The index loads are issued
// together, then the slot loads: two memory round trips per batch instead of
// two per partner.
constexpr int SKIP_BATCH = 16;
__device__ inline void skip_slots(const DevPtrs &P, int e, int e1, int (&ps)[SKIP_BATCH]) {
    int idx[SKIP_BATCH];
#pragma unroll
    for (int u = 0; u < SKIP_BATCH; ++u) idx[u] = e + u < e1 ? P.skip_idx[e + u] : -1;
#pragma unroll
    for (int u = 0; u < SKIP_BATCH; ++u) ps[u] = idx[u] >= 0 ? P.slot_of[idx[u]] : -1;
}

// One wave per i-block b: the row of j-blocks whose build-time boxes lie
// within the list cutoff (full list, both directions, ascending j) and the
// mask slot of each tile whose j-block holds a skipped pair (or is the block
// itself).  The partner blocks and their mask words are rebuilt only when the
// atoms were re-sorted (resort); each lane owns column `lane` of the words.
// Returns the row length.
__device__ inline int build_row(const DevPtrs &P, int b, bool resort, int lane) {
    const int s = P.block_sys[b];
    const SysDesc &D = P.sys[s];
    SysState &S = P.sst[s];
    // the partner blocks first: the row's mask slots are looked up among them
    int pb_reg = INT_MAX, n_pb = 0;
    if (resort) {
        // the block itself (self pairs) and every block holding a skip
        // partner of one of its atoms, found by repeated wave minimum
        const int a = P.perm[b * BLOCK + lane];
        const int e0 = a >= 0 ? P.skip_off[a] : 0, e1 = a >= 0 ? P.skip_off[a + 1] : 0;
        int cur = -1;
        for (;;) {
            int cand = cur < b ? b : INT_MAX;
            for (int e = e0; e < e1; e += SKIP_BATCH) {
                int ps[SKIP_BATCH];
                skip_slots(P, e, e1, ps);
#pragma unroll
                for (int u = 0; u < SKIP_BATCH; ++u) {
                    const int q = ps[u] / BLOCK;
                    if (ps[u] >= 0 && q > cur && q < cand) cand = q;
                }
            }
            cand = wave_min_int(cand);
            if (cand == INT_MAX) break;
            if (lane == n_pb) pb_reg = cand;
            ++n_pb;
            cur = cand;
        }
        if (n_pb > MAX_MASKS) {
            if (lane == 0) atomicOr(&S.status, ST_MASK_OVERFLOW);
            n_pb = MAX_MASKS;
        }
        // one 64-bit word per (partner block, i-lane): the self bit or zero,
        // then a bit per skip partner.  Each lane owns column `lane` of the
        // words, so plain stores and loads suffice: one thread's accesses to
        // one address keep their program order.
        unsigned long long *mrow = P.masks + (size_t)b * MAX_MASKS * 64;
        const int mb = pb_slot(pb_reg, n_pb, b);
        for (int m = 0; m < n_pb; ++m) mrow[m * 64 + lane] = m == mb ? 1ull << lane : 0ull;
        // (a lane's consecutive partners in one block are OR-ed into a
        // register first: one update per run instead of one per partner)
        const int n_e = e1 - e0, n_e_max = -wave_min_int(-n_e);
        int run_m = -1;
        unsigned long long run = 0ull;
        for (int k0 = 0; k0 < n_e_max; k0 += SKIP_BATCH) {  // every lane in every round: pb_slot reads across lanes
            int ps[SKIP_BATCH];
            skip_slots(P, e0 + k0, e1, ps);
#pragma unroll
            for (int u = 0; u < SKIP_BATCH; ++u) {
                // (rounds past n_e_max search too, with no partner: -1)
                const int m = pb_slot(pb_reg, n_pb, ps[u] >= 0 ? ps[u] / BLOCK : -1);
                if (ps[u] >= 0 && m >= 0) {
                    if (m != run_m) {
                        if (run_m >= 0) mrow[run_m * 64 + lane] |= run;
                        run_m = m;
                        run = 0ull;
                    }
                    run |= 1ull << (ps[u] % BLOCK);
                }
            }
        }
        if (run_m >= 0) mrow[run_m * 64 + lane] |= run;
        if (lane < n_pb) P.pblk[(size_t)b * MAX_MASKS + lane] = pb_reg;
        if (lane == 0) P.n_pblk[b] = n_pb;
    } else {
        n_pb = P.n_pblk[b];
        if (lane < n_pb) pb_reg = P.pblk[(size_t)b * MAX_MASKS + lane];
    }
    // the row: j-blocks of the system within reach, with their mask slots
    const float4 ci = P.bb_c[b], hi = P.bb_h[b];
    const float reach = D.list_cutoff + 0.01f;
    const float reach2 = reach * reach;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int count = 0;
    const int jend = D.block_begin + D.n_blocks;
    for (int j0 = D.block_begin; j0 < jend; j0 += 64) {
        const int jb = j0 + lane;
        bool ok = false;
        if (jb < jend) {
            const float4 cj = P.bb_c[jb], hj = P.bb_h[jb];
            const float dx = fmaxf(fabsf(ci.x - cj.x) - hi.x - hj.x, 0.0f);
            const float dy = fmaxf(fabsf(ci.y - cj.y) - hi.y - hj.y, 0.0f);
            const float dz = fmaxf(fabsf(ci.z - cj.z) - hi.z - hj.z, 0.0f);
            ok = dx * dx + dy * dy + dz * dz <= reach2;
        }
        const unsigned long long m = __ballot(ok);
        const int pos = count + __popcll(m & lt);
        const int slot = pb_slot(pb_reg, n_pb, jb);
        if (ok && pos < MAX_ROW) {
            P.row_j[(size_t)b * MAX_ROW + pos] = jb;
            P.row_m[(size_t)b * MAX_ROW + pos] = slot;
        }
        count += __popcll(m);
    }
    if (count > MAX_ROW) {
        if (lane == 0) atomicOr(&S.status, ST_ROW_OVERFLOW);
        count = MAX_ROW;
    }
    if (lane == 0) {
        P.row_n[b] = count;
        P.unit_count[b] = QUARTERS * ((count + D.chunk - 1) / D.chunk);
    }
    return count;
}

// The unit count of block b on a step where some system rebuilds (one wave):
// a rebuilding system's block rebuilds its row (and, after a re-sort, its
// masks and its quarters' current boxes, which the sort kernel leaves to this
// pass); any other active block keeps its count; an inactive block has none.
__device__ inline int rows_count(const DevPtrs &P, int b, int lane) {
    const int s = P.block_sys[b];
    if (s < 0 || !P.sys[s].active) return 0;
    if (!P.sst[s].rebuild_flag) return P.unit_count[b];
    const bool resort = resort_due(P.sys[s], P.sst[s]);
    if (resort) quarter_boxes(P, b, lane);
    const int tiles = build_row(P, b, resort, lane);
    const int ch = P.sys[s].chunk;
    return QUARTERS * ((tiles + ch - 1) / ch);
}

// The count rows_count returned, read back once the rows are built.
__device__ inline int rows_count_built(const DevPtrs &P, int b) {
    const int s = P.block_sys[b];
    return s >= 0 && P.sys[s].active ? P.unit_count[b] : 0;
}

// Block b's cnt units at [off, off + cnt) of the unit list: quarter q's
// chunks in order, each stored with its first tile, so decoding needs no
// descriptor.
__device__ inline void write_units(const DevPtrs &P, int b, int off, int cnt, int lane) {
    if (cnt == 0) return;
    if (lane == 0) P.unit_off[b] = off;
    const int per_q = cnt / QUARTERS;  // chunks per quarter
    const int ch = P.sys[P.block_sys[b]].chunk;
    for (int k = lane; k < cnt; k += 64) P.units[off + k] = (b << 11) | ((k / per_q) << 9) | ((k % per_q) * ch);
}

// rows_count, then a range of the unit list claimed with one atomic add (the
// list's order follows the claims, no result depends on it).
__device__ inline void rows_block(const DevPtrs &P, int b, int lane) {
    const int cnt = rows_count(P, b, lane);
    if (cnt == 0) return;
    int off = 0;
    if (lane == 0) off = atomicAdd(&P.counters->n_units, cnt);
    write_units(P, b, __shfl(off, 0), cnt, lane);
}

}  // namespace prism
