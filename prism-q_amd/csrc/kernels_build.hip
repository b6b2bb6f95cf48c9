// kernels_build.hip -- pair-list rebuild: per-system Hilbert cell hash with a
// stable LSD radix sort, 64-atom spatial blocks, cluster-pair tile rows and
// exclusion bit masks.
//
// Replaces reset_cell_counts / build_cell_list / nhs_build_neighbor_list_warp
// (nhs_amber_fused.cu:11690-11795, nhs_neighbor_warp.cuh:10-119) driven by
// rebuild_neighbor_lists (fused_engine.rs:16615-16748).  The reference keeps a
// per-atom row of <= 1024 neighbour ids at 12 A, rebuilt after every SHAKE; this
// build keeps block-pair tiles at cutoff + skin and rebuilds when an atom has
// moved more than skin/2 or after rebuild_every steps.
#include "bonded.hpp"
#include "device_common.hpp"
#include "rows.hpp"

namespace prism {

namespace {

constexpr int SORT_THREADS = 1024;
constexpr int SORT_WAVES = SORT_THREADS / 64;

// 3-D Hilbert index of cell (x, y, z), `bits` bits per axis, after J. Skilling,
// "Programming the Hilbert curve", AIP Conf. Proc. 707, 381 (2004): the
// axes-to-transpose transform followed by bit interleaving.  Unlike Morton
// order the curve has no jumps, so 64 consecutive atoms form compact blocks.
__device__ inline unsigned int hilbert_key(unsigned int x, unsigned int y, unsigned int z, int bits) {
    unsigned int X[3] = {x, y, z};
    const unsigned int M = 1u << (bits - 1);
    for (unsigned int Q = M; Q > 1; Q >>= 1) {
        const unsigned int P = Q - 1;
        for (int i = 0; i < 3; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                const unsigned int t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    X[1] ^= X[0];
    X[2] ^= X[1];
    unsigned int t = 0;
    for (unsigned int Q = M; Q > 1; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1;
    for (int i = 0; i < 3; ++i) X[i] ^= t;
    unsigned int key = 0;
    for (int b = bits - 1; b >= 0; --b)
        for (int i = 0; i < 3; ++i) key = (key << 1) | ((X[i] >> b) & 1u);
    return key;
}

// Box of one block from the current positions of its atoms; the build-time
// positions are recorded for the displacement trigger.
__device__ inline void refresh_block(const DevPtrs &P, int b, int lane) {
    const int a = P.perm[b * BLOCK + lane];
    float l[3], h[3];
    if (a >= 0) {
        const float4 p = P.x[a];
        P.x_build[a] = p;
        l[0] = h[0] = p.x; l[1] = h[1] = p.y; l[2] = h[2] = p.z;
    } else {
        l[0] = l[1] = l[2] = 3.0e38f;
        h[0] = h[1] = h[2] = -3.0e38f;
    }
    for (int c = 0; c < 3; ++c) {
        l[c] = wave_min(l[c]);
        h[c] = wave_max(h[c]);
    }
    if (lane == 0) {
        P.bb_c[b] = make_float4(0.5f * (l[0] + h[0]), 0.5f * (l[1] + h[1]), 0.5f * (l[2] + h[2]), 0.0f);
        P.bb_h[b] = make_float4(0.5f * (h[0] - l[0]), 0.5f * (h[1] - l[1]), 0.5f * (h[2] - l[2]), 0.0f);
    }
}

// Sorted slot `slot` (the i-th of its system) holding arena atom a (-1:
// padding): inverse permutation, position with the scaled charge, and the
// Lorentz-Berthelot halves (sigma/2, sqrt(24 eps)).  Padding slots sit far
// apart outside every cutoff.
__device__ inline void fill_slot(const DevPtrs &P, int slot, int i, int a) {
    if (a >= 0) {
        const float qscale = (float)sqrt(COULOMB);
        const float4 p = P.x[a];
        P.slot_of[a] = slot;
        P.xq[slot] = make_float4(p.x, p.y, p.z, P.charge[a] * qscale);
        P.lj[slot] = make_float2(0.5f * P.sigma[a], sqrtf(24.0f * P.eps[a]));
    } else {
        P.xq[slot] = make_float4(DUMMY_FAR + 64.0f * (float)i, DUMMY_FAR, DUMMY_FAR, 0.0f);
        P.lj[slot] = make_float2(0.0f, 0.0f);
    }
}

// Slots of system s rebuilt from an uploaded permutation (checkpoint restore).
__global__ __launch_bounds__(256) void k_restore_slots(DevPtrs P, int s) {
    const SysDesc &D = P.sys[s];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D.n_blocks * BLOCK) return;
    const int slot = D.block_begin * BLOCK + i;
    fill_slot(P, slot, i, P.perm[slot]);
}

// Workgroups [0, n_sys): one per system, a full re-sort when the system's
// rebuild is due and resort_due() holds.  Workgroups [n_sys, bonded_wg0): one
// wave per block, every step the current boxes of the block's quarters (not
// of a system that re-sorts: the rows pass boxes those after the sort) and a box
// refresh of blocks whose system rebuilds without re-sorting.  Workgroups from bonded_wg0 on (only in the MD step's launch):
// the bonded terms and exceptions of the step's force evaluation, one per
// thread -- they read the positions the sort leaves untouched and write
// contribution slots that only the gather reads, and riding here saves the
// step a kernel launch.  Nothing here writes the flags it reads; they are
// consumed at the end of the step (k_second_half) or by k_consume after a
// rebuild between steps.
template <bool TM>
__global__ __launch_bounds__(SORT_THREADS) void k_sort(DevPtrs P, int bonded_wg0) {
    __shared__ float red[6][SORT_WAVES];
    __shared__ unsigned int hist[256];
    __shared__ unsigned int running[256];
    __shared__ unsigned int wave_cnt[SORT_WAVES][256];
    const int tid = threadIdx.x, lane = tid & 63, wave = wave_id();
    if ((int)blockIdx.x >= bonded_wg0) {
        bonded_term<false, TM>(P, PRISM_MD_TERM_ALL, ((int)blockIdx.x - bonded_wg0) * SORT_THREADS + tid);
        return;
    }
    if ((int)blockIdx.x >= P.n_sys) {
        // one wave per block: every step the quarters' current boxes for the
        // nonbonded culls -- unless the block's system re-sorts this step,
        // which moves atoms between slots, and the rows pass boxes them after the
        // sort -- then, on a rebuild without re-sort, the box refresh
        const int b = ((int)blockIdx.x - P.n_sys) * SORT_WAVES + wave;
        if (b >= P.n_blocks_arena) return;
        const int s = P.block_sys[b];
        if (s < 0) return;
        const SysDesc &D = P.sys[s];
        const SysState &S = P.sst[s];
        if (!D.active) return;
        const bool rebuild = P.counters->any_rebuild != 0 && S.rebuild_flag;
        const bool resort = rebuild && resort_due(D, S);
        if (!resort) quarter_boxes(P, b, lane);
        if (!rebuild || resort) return;
        refresh_block(P, b, lane);
        if (lane == 0 && b == D.block_begin) {
            atomicExch(&P.counters->any_dirty, 1);
            atomicExch(&P.counters->n_units, 0);  // every active block re-claims its units
        }
        return;
    }
    if (P.counters->any_rebuild == 0) return;  // nothing to re-sort this step: one load
    const int s = blockIdx.x;
    const SysDesc &D = P.sys[s];
    const SysState &S = P.sst[s];
    if (!D.active || !S.rebuild_flag || !resort_due(D, S)) return;
    const int n = D.n_atoms, a0 = D.atom_begin;

    // ---- bounding box of the current positions
    float lo[3] = {3.0e38f, 3.0e38f, 3.0e38f}, hi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    for (int i = tid; i < n; i += SORT_THREADS) {
        const float4 p = P.x[a0 + i];
        lo[0] = fminf(lo[0], p.x); lo[1] = fminf(lo[1], p.y); lo[2] = fminf(lo[2], p.z);
        hi[0] = fmaxf(hi[0], p.x); hi[1] = fmaxf(hi[1], p.y); hi[2] = fmaxf(hi[2], p.z);
    }
    for (int c = 0; c < 3; ++c) {
        lo[c] = wave_min(lo[c]);
        hi[c] = wave_max(hi[c]);
    }
    if (lane == 0)
        for (int c = 0; c < 3; ++c) { red[c][wave] = lo[c]; red[3 + c][wave] = hi[c]; }
    __syncthreads();
    int bits = 1;
    const float inv_cell = 1.0f / D.sort_cell;
    for (int c = 0; c < 3; ++c) {
        lo[c] = 3.0e38f;
        hi[c] = -3.0e38f;
        for (int w = 0; w < SORT_WAVES; ++w) {
            lo[c] = fminf(lo[c], red[c][w]);
            hi[c] = fmaxf(hi[c], red[3 + c][w]);
        }
        const unsigned int cells = (unsigned int)fminf((hi[c] - lo[c]) * inv_cell, 1023.0f) + 1u;
        bits = max(bits, 32 - __clz(cells));  // bits to hold cell indices 0..cells-1
    }
    bits = min(bits, D.sort_bits);
    const unsigned int cmax = (1u << bits) - 1u;

    // ---- keys: Hilbert index of sort cells relative to the box minimum
    for (int i = tid; i < n; i += SORT_THREADS) {
        const float4 p = P.x[a0 + i];
        const unsigned int cx = min(cmax, (unsigned int)fmaxf(0.0f, (p.x - lo[0]) * inv_cell));
        const unsigned int cy = min(cmax, (unsigned int)fmaxf(0.0f, (p.y - lo[1]) * inv_cell));
        const unsigned int cz = min(cmax, (unsigned int)fmaxf(0.0f, (p.z - lo[2]) * inv_cell));
        P.keys_a[a0 + i] = hilbert_key(cx, cy, cz, bits);
        P.vals_a[a0 + i] = i;
    }
    __syncthreads();

    // ---- stable LSD radix sort, 8-bit digits, one workgroup.  Each wave owns
    // one contiguous segment of the keys; a pass counts every wave's digits,
    // turns the counts into (digit, wave) offsets -- digit-major, wave-minor,
    // which keeps the pass stable -- and then every wave scatters its segment
    // in order with its own running offsets (a 64-key chunk at a time, ranks
    // inside the chunk from ballots).  Four workgroup barriers per pass.
    unsigned int *kin = P.keys_a + a0, *kout = P.keys_b + a0;
    int *vin = P.vals_a + a0, *vout = P.vals_b + a0;
    const int total_bits = 3 * bits;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int seg = (n + SORT_WAVES - 1) / SORT_WAVES;
    const int s0 = min(n, wave * seg), s1 = min(n, s0 + seg);
    for (int shift = 0; shift < total_bits; shift += 8) {
        for (int k = 0; k < 4; ++k) wave_cnt[wave][lane + 64 * k] = 0;
        __syncthreads();
        for (int i = s0 + lane; i < s1; i += 64) atomicAdd(&wave_cnt[wave][(kin[i] >> shift) & 255u], 1u);
        __syncthreads();
        if (tid < 256) {  // exclusive scan of the digit totals: waves 0..3, 64 digits each
            unsigned int tot = 0;
            for (int w = 0; w < SORT_WAVES; ++w) tot += wave_cnt[w][tid];
            unsigned int inc = tot;
            for (int o = 1; o < 64; o <<= 1) {
                const unsigned int y = __shfl_up(inc, o);
                if (lane >= o) inc += y;
            }
            if (lane == 63) hist[wave] = inc;  // wave totals, wave < 4
            running[tid] = inc - tot;
        }
        __syncthreads();
        if (tid < 256) {
            unsigned int acc = running[tid];
            for (int w = 0; w < wave; ++w) acc += hist[w];
            for (int w = 0; w < SORT_WAVES; ++w) {
                const unsigned int c = wave_cnt[w][tid];
                wave_cnt[w][tid] = acc;  // where wave w's keys with digit tid start
                acc += c;
            }
        }
        __syncthreads();
        for (int base = s0; base < s1; base += 64) {
            const int i = base + lane;
            const bool valid = i < s1;
            const unsigned int key = valid ? kin[i] : 0u;
            const int val = valid ? vin[i] : 0;
            const unsigned int d = (key >> shift) & 255u;
            unsigned long long peers = __ballot(valid);
            if (!valid) peers = ~peers;
            for (int b = 0; b < 8; ++b) {
                const bool bit = (d >> b) & 1u;
                const unsigned long long m = __ballot(bit);
                peers &= bit ? m : ~m;
            }
            const int rank = __popcll(peers & lt);
            const unsigned int first = wave_cnt[wave][d];
            if (valid) {
                kout[first + rank] = key;
                vout[first + rank] = val;
            }
            wave_lds_fence();  // every lane has read its digit's offset
            if (valid && rank == 0) wave_cnt[wave][d] = first + (unsigned int)__popcll(peers);
            wave_lds_fence();
        }
        __syncthreads();
        unsigned int *tk = kin; kin = kout; kout = tk;
        int *tv = vin; vin = vout; vout = tv;
    }

    // ---- gather into sorted slots
    const int slot0 = D.block_begin * BLOCK;
    for (int i = tid; i < D.n_blocks * BLOCK; i += SORT_THREADS) {
        const int a = i < n ? a0 + vin[i] : -1;
        P.perm[slot0 + i] = a;
        fill_slot(P, slot0 + i, i, a);
    }
    __syncthreads();

    // ---- block boxes over real atoms
    for (int b = wave; b < D.n_blocks; b += SORT_WAVES) refresh_block(P, D.block_begin + b, lane);
    if (tid == 0) {
        atomicExch(&P.counters->any_dirty, 1);
        atomicExch(&P.counters->n_units, 0);  // every active block re-claims its units
    }
}

constexpr int ROW_WAVES = 2;
// LDS that k_rows reserves and does not use: with it a compute unit holds at
// most two of its workgroups, so the blocks claim their units roughly in
// block order and each XCD's eighth of the unit list stays a compact region.
// (Without it every block is in flight at once, claims come in completion
// order, and the ensemble's nonbonded kernel was measured 13% slower.)
constexpr size_t ROWS_LDS_RESERVE = 56 * 1024;
// One wave per block.  On a batch's MD step and on a rebuild outside the
// step (a single system's step builds its rows inside its nonbonded launch,
// in block order): blocks of a
// rebuilding system rebuild their row; every active block then claims a range
// of the unit list with one atomic add and writes its units there
// (rows_block).  The claim order varies, but no result depends on it: the
// gather reads each block's units through unit_off in chunk order.
__global__ __launch_bounds__(64 * ROW_WAVES) void k_rows(DevPtrs P) {
    const int b = blockIdx.x * ROW_WAVES + wave_id();
    if (P.counters->any_dirty == 0 || b >= P.n_blocks_arena) return;
    rows_block(P, b, threadIdx.x & 63);
}

// The quarters' boxes alone, for the force evaluations outside the step (whose
// positions were set or moved since the last k_rows).
__global__ __launch_bounds__(256) void k_qbox(DevPtrs P) {
    const int b = blockIdx.x * 4 + wave_id();
    if (b >= P.n_blocks_arena) return;
    const int s = P.block_sys[b];
    if (s < 0 || !P.sys[s].active) return;
    quarter_boxes(P, b, threadIdx.x & 63);
}

// Consume the rebuild flags after a rebuild between steps (the integrator
// consumes them at the end of a step).
__global__ __launch_bounds__(256) void k_consume(DevPtrs P, int step_offset) {
    for (int s = threadIdx.x; s < P.n_sys; s += blockDim.x) {
        SysState &S = P.sst[s];
        if (!P.sys[s].active || !S.rebuild_flag) continue;
        S.rebuild_flag = 0;
        S.resort_flag = 0;
        S.rebuild_step = (long long)S.step + step_offset;
        S.n_rebuilds += 1;
    }
    if (threadIdx.x == 0) {
        P.counters->any_dirty = 0;
        P.counters->any_rebuild = 0;
        P.counters->nb_arrive = 0u;
        P.counters->nb_go = 0u;
    }
}

}  // namespace

void launch_sort(const DevPtrs &P, bool with_bonded, hipStream_t st) {
    const int refresh = (P.n_blocks_arena + SORT_WAVES - 1) / SORT_WAVES;
    const int wg0 = P.n_sys + refresh;
    const int terms = with_bonded ? P.n_bonds + P.n_angles + P.n_dih + P.n_exc : 0;
    const int bonded = (terms + SORT_THREADS - 1) / SORT_THREADS;
    if (P.term_major) hipLaunchKernelGGL(k_sort<true>, dim3(wg0 + bonded), dim3(SORT_THREADS), 0, st, P, wg0);
    else hipLaunchKernelGGL(k_sort<false>, dim3(wg0 + bonded), dim3(SORT_THREADS), 0, st, P, wg0);
}

void launch_rows(const DevPtrs &P, hipStream_t st) {
    const int grid = (P.n_blocks_arena + ROW_WAVES - 1) / ROW_WAVES;
    hipLaunchKernelGGL(k_rows, dim3(grid), dim3(64 * ROW_WAVES), ROWS_LDS_RESERVE, st, P);
}

void launch_quarter_boxes(const DevPtrs &P, hipStream_t st) {
    hipLaunchKernelGGL(k_qbox, dim3((P.n_blocks_arena + 3) / 4), dim3(256), 0, st, P);
}

void launch_restore_slots(const DevPtrs &P, int s, int n_slots, hipStream_t st) {
    hipLaunchKernelGGL(k_restore_slots, dim3((n_slots + 255) / 256), dim3(256), 0, st, P, s);
}

void launch_consume(const DevPtrs &P, int step_offset, hipStream_t st) {
    hipLaunchKernelGGL(k_consume, dim3(1), dim3(256), 0, st, P, step_offset);
}

}  // namespace prism
