// kernels_force.hip -- force evaluation: nonbonded cluster-pair tiles,
// bonded terms and exact 1-4 exceptions.  Every partial force is converted to
// int64 fixed point (2^24 per kcal/mol/A) before it is summed, so the total is
// independent of summation order and bitwise reproducible: the MD step's
// evaluation adds each work unit's partial into the atom's accumulator with an
// integer atomic; the evaluations outside the step store per-unit partials
// and a gather sums them per atom in a fixed order.
//
// Replaces the FORCE_N / FORCE_NP1 gather of prism_amber_phase_body_v1
// (nhs_amber_fused.cu:2981-3484) with its term functions compute_bond_force
// (:354-386), compute_angle_force (:388-448), compute_dihedral_force
// (:524-625), compute_nonbonded_force (:643-717) and
// ordinary_cutoff_potential_offset_v1 (:631-640).
#include <hip/hip_ext.h>

#include "bonded.hpp"
#include "device_common.hpp"
#include "rows.hpp"

namespace prism {

namespace {

constexpr int NB_WAVES = 4;
constexpr int RING = 128;  // survivor ring per wave: two batches of 64

// Cutoff test of the nonbonded pair: r2 = fma(dz, dz, fma(dy, dy, dx*dx)) in
// FP32 with d = x_i - x_j.  The oracle decides pair membership with the same
// rounding (fmaf), so the evaluated pair set is identical on both sides.
__device__ inline float pair_r2(float dx, float dy, float dz) { return fmaf(dz, dz, fmaf(dy, dy, dx * dx)); }

// Bit patterns of positive floats order like the floats: one unsigned compare
// tests 0.01 <= r2 < cutoff2 (the reference's skip, nhs_amber_fused.cu:3159).
__device__ inline bool in_shell(float r2, unsigned lo_bits, unsigned span) {
    return (unsigned)(__float_as_uint(r2) - lo_bits) < span;
}

// The survivors of a tile sit in LDS padded to a multiple of 4 with dummy
// j-atoms far outside the cutoff, so every lane's j index is valid.  Kept
// pairs have r2 >= 0.01 (never denormal), rejected ones are zeroed by the
// select, so the raw v_rsq_f32 is used without denormal scaling.
template <bool ENERGY, bool MASKED, class PAR>
__device__ inline void tile_pairs(const float4 xi, const float2 li, int n_iter, int js, const float4 *jpos,
                                  const PAR *jpar, unsigned long long excl, unsigned lo_bits, unsigned span,
                                  float inv_rc, float inv_rc2, float &fx, float &fy, float &fz, float &e) {
#pragma unroll 2
    for (int k = 0; k < n_iter; ++k) {
        const int jj = 4 * k + js;
        const float4 q = jpos[jj];
        const PAR pr = jpar[jj];
        const float dx = xi.x - q.x, dy = xi.y - q.y, dz = xi.z - q.z;
        const float r2 = pair_r2(dx, dy, dz);
        bool keep = in_shell(r2, lo_bits, span);
        if constexpr (MASKED) keep = keep && !((excl >> __float_as_int(pr.z)) & 1ull);
        const float inv_r = __builtin_amdgcn_rsqf(r2);
        const float inv_r2 = inv_r * inv_r;
        const float sig = li.x + pr.x;
        const float s2 = sig * sig * inv_r2;
        const float s6 = s2 * s2 * s2;
        const float eps24 = li.y * pr.y;  // 24 sqrt(eps_i eps_j)
        const float fc = (xi.w * q.w) * inv_r;  // 332.0636 q_i q_j / r
        float f = fmaf(eps24 * s6, fmaf(2.0f, s6, -1.0f), fc) * inv_r2;
        f = keep ? f : 0.0f;
        fx = fmaf(f, dx, fx);
        fy = fmaf(f, dy, fy);
        fz = fmaf(f, dz, fz);
        if (ENERGY) {
            const float c2 = sig * sig * inv_rc2;
            const float c6 = c2 * c2 * c2;
            const float v = (eps24 * (1.0f / 6.0f)) * ((s6 * s6 - s6) - (c6 * c6 - c6)) + (xi.w * q.w) * (inv_r - inv_rc);
            e += keep ? 0.5f * v : 0.0f;
        }
    }
}

constexpr int N_XCD = 8;

__device__ inline long long wave_xor_ll(long long v, int m) {
    const int lo = __shfl_xor((int)(v & 0xffffffffll), m), hi = __shfl_xor((int)(v >> 32), m);
    return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// The rows of a rebuild step inside a single system's MD-step nonbonded
// launch (instead of a k_rows launch on every step; a batch keeps k_rows,
// since re-sort steps of many systems at once were measured far slower
// here: a serving pool's p99 step latency 0.48 -> 1.68 ms).
// When the sort kernel marked the lists
// dirty, the first n_builders workgroups each take K consecutive blocks, one
// block per wave at a time: they build the rows (rows_count) and publish
// their unit totals (rows_total); after a grid barrier each builder adds up
// the totals before it, writes its blocks' units at those offsets -- the unit
// list in block order, whatever the timing, so each XCD's eighth of it is a
// compact region -- and the last builder stores n_units; after a second
// barrier every workgroup claims units.  Only the builders count in
// (nb_arrive, zeroed at every step end: phase p completes at p x n_builders
// arrivals); the last of a phase stores p in nb_go, which the others poll.
// The persistent grid is sized to the workgroups the device holds at once, so
// every builder runs; the wait is bounded all the same (ROWS_WAIT_TICKS of the
// constant-rate clock, then ST_BARRIER latches on every active system and
// the launch ends without its units).  A builder's writes are released at
// agent scope before it counts in, and every workgroup acquires after it sees
// nb_go.  Returns false on a timeout.
constexpr unsigned long long ROWS_WAIT_TICKS = 1ull << 26;  // ~0.67 s at 100 MHz

__device__ inline bool rows_barrier(const DevPtrs &P, bool builder, unsigned n_builders, unsigned phase) {
    __shared__ int ok;
    __syncthreads();  // the workgroup's writes are done; a builder releases them at agent scope once
    if (threadIdx.x == 0) {
        int good = 1;
        bool last = false;
        if (builder) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            last = atomicAdd(&P.counters->nb_arrive, 1u) == phase * n_builders - 1u;
            if (last) {
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
                __hip_atomic_store(&P.counters->nb_go, phase, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (!last) {
            const unsigned long long t0 = wall_clock64();
            while (__hip_atomic_load(&P.counters->nb_go, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < phase) {
                if (wall_clock64() - t0 > ROWS_WAIT_TICKS) {
                    good = 0;
                    for (int q = 0; q < P.n_sys; ++q)
                        if (P.sys[q].active) atomicOr(&P.sst[q].status, ST_BARRIER);
                    break;
                }
                __builtin_amdgcn_s_sleep(8);  // few reads of the one word while the rows are built
            }
        }
        // acquire only: a waiter wrote nothing the others read, and a full
        // fence would write its XCD's L2 back once per workgroup
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        ok = good;
    }
    __syncthreads();
    return ok;
}

// The rows phase above; false if a barrier timed out.
__device__ inline bool rows_in_launch(const DevPtrs &P, int wave, int lane) {
    const int B = P.n_blocks_arena;
    if (B == 0) return true;
    const int per = (int)gridDim.x * NB_WAVES;
    const int K = NB_WAVES * ((B + per - 1) / per);  // blocks per builder
    const unsigned n_builders = (unsigned)((B + K - 1) / K);
    const bool builder = blockIdx.x < n_builders;
    const int first = (int)blockIdx.x * K, end = min(B, first + K);
    __shared__ int wsum[NB_WAVES];
    if (builder) {
        int mine = 0;
        for (int b = first + wave; b < end; b += NB_WAVES) mine += rows_count(P, b, lane);
        if (lane == 0) wsum[wave] = mine;
        __syncthreads();
        if (threadIdx.x == 0) {
            int t = 0;
            for (int w = 0; w < NB_WAVES; ++w) t += wsum[w];
            P.rows_total[blockIdx.x] = t;
        }
    }
    if (!rows_barrier(P, builder, n_builders, 1u)) return false;
    if (builder) {
        // every wave adds the totals before this builder and scans its blocks'
        // counts itself, then writes the units of the blocks it built
        int base = 0;
        for (int w = lane; w < (int)blockIdx.x; w += 64) base += P.rows_total[w];
        for (int o = 32; o > 0; o >>= 1) base += __shfl_xor(base, o);
        int carry = base;
        for (int c0 = first; c0 < end; c0 += 64) {
            const int cn = c0 + lane < end ? rows_count_built(P, c0 + lane) : 0;
            int inc = cn;
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(inc, o);
                if (lane >= o) inc += t;
            }
            const int excl = carry + inc - cn;
            for (int k = wave; k < min(64, end - c0); k += NB_WAVES)
                write_units(P, c0 + k, __shfl(excl, k), __shfl(cn, k), lane);
            carry += __shfl(inc, 63);
        }
        if (blockIdx.x == n_builders - 1u && threadIdx.x == 0) P.counters->n_units = carry;
    }
    return rows_barrier(P, builder, n_builders, 2u);
}

// Persistent grid over the compact unit list.  A unit is (i-block b, quarter
// q, chunk c of D.chunk tiles of b's row).  Lane l works on i-atom
// b*64 + q*16 + (l & 15) against j-slot l >> 4: every instruction evaluates
// 16 i-atoms x 4 j-atoms.  Per tile the 64 j-atoms are culled against the
// quarter's current bounding box at the cutoff (P.qbox, written before the
// launch by quarter_boxes) and the survivors compacted into LDS.  The row is
// a full list (both directions), so each lane only accumulates its own
// i-force; the four j-slot partials are summed at the end of the unit and
// stored as the unit's partial (no atomics).
//
// DIRECT (the MD step's evaluation): instead of storing the partial for a
// separate gather, each unit converts it to int64 fixed point and adds it to
// the atom's force accumulator with an integer atomic; the first unit of a
// quarter also adds the quarter's bonded contribution slots.  Integer
// addition is associative, so the force is bitwise the one the gather forms
// (the same fixed-point terms, summed in another order), and the step loses
// a kernel and the partials' round trip through HBM.  The accumulators are
// zeroed by the integrator's first half when it consumes the carried force.
template <bool ENERGY, bool DIRECT, bool TM>
__global__ __launch_bounds__(64 * NB_WAVES) void k_nonbonded(DevPtrs P) {
    const int wave = wave_id(), lane = threadIdx.x & 63;
    const int il = lane & (SUB - 1), js = lane >> 4;
    // per wave: the survivors of a masked tile, and a ring of the survivors of
    // unmasked tiles evaluated 64 at a time
    __shared__ float4 jpos_s[NB_WAVES][64];
    __shared__ float4 jpar_s[NB_WAVES][64];
    __shared__ float4 rpos_s[NB_WAVES][RING];
    __shared__ float2 rpar_s[NB_WAVES][RING];
    float4 *jpos = jpos_s[wave];
    float4 *jpar = jpar_s[wave];
    float4 *rpos = rpos_s[wave];
    float2 *rpar = rpar_s[wave];
    if (DIRECT && P.rows_in_nonbonded && P.counters->any_dirty != 0 && !rows_in_launch(P, wave, lane)) return;
    const int n_units = P.counters->n_units;
    // Workgroups are dispatched round-robin over the XCDs (workgroup w on XCD
    // w % 8, each XCD with its own L2).  Each XCD takes one contiguous eighth
    // of the unit list -- blocks are space-filling-curve ordered, so that is a
    // compact region whose j-tiles stay in its L2 -- and its waves go round
    // robin over it.  (A shared claim counter was measured far slower:
    // device-scope atomics on one address serialise.)
    const bool by_xcd = gridDim.x % N_XCD == 0;
    const int xcd = by_xcd ? (int)(blockIdx.x % N_XCD) : 0;
    const int span_u = by_xcd ? (n_units + N_XCD - 1) / N_XCD : n_units;
    const int u_begin = xcd * span_u, u_end = min(n_units, u_begin + span_u);
    const int wave_x = (by_xcd ? (int)(blockIdx.x / N_XCD) : (int)blockIdx.x) * NB_WAVES + wave;  // wave within the range
    const int waves_x = (by_xcd ? (int)(gridDim.x / N_XCD) : (int)gridDim.x) * NB_WAVES;
    // Units are claimed, not assigned: the range is split into
    // NB_CLAIM_GROUPS interleaved sub-lists (unit u_begin + g + k G), each
    // with a counter that the range's waves of group g advance by one per
    // claim, so a wave that drew short units draws more.  (The static
    // round-robin it replaces left most waves idle behind the slowest.)  A
    // unit is computed whole by the wave that claims it, so no result
    // depends on the claim order.  Every wave of the group claims until a
    // claim comes back empty; the group's last claim overall (value
    // n_g + waves_g - 1) resets the counter to zero for the next launch.
    const int G = min(NB_CLAIM_GROUPS, waves_x);  // every group has a wave
    const int g = wave_x % G;
    const int n_g = u_end > u_begin + g ? (u_end - u_begin - g + G - 1) / G : 0;
    const int waves_g = (waves_x - g + G - 1) / G;
    // (+ mbcnt(0) = 0: the compiler takes the address for a per-lane one and
    // leaves the claim a plain returning atomic, whose result is first read
    // at the unit's end; rewritten as a wave-aggregated atomic it would be
    // read at once and the unit would start by waiting for it)
    unsigned *ctr = P.nb_claim + (size_t)(xcd * G + g) * NB_CLAIM_STRIDE + __builtin_amdgcn_mbcnt_lo(0u, 0u);
    // The next unit is claimed when a unit starts; the claim is read and the
    // next unit's code loaded once the unit's tiles are known, and the next
    // unit's row entries load before the unit's last batch, reductions and
    // force atomics -- so a unit starts with its row in registers, and no
    // unit start waits for a claim, a code or the previous atomics.
    unsigned k_pending = 0u;
    if (lane == 0) k_pending = atomicAdd(ctr, 1u);
    int k_cur = __builtin_amdgcn_readfirstlane((int)k_pending);
    int code_cur = k_cur < n_g ? __builtin_amdgcn_readfirstlane(P.units[u_begin + g + k_cur * G]) : 0;
    // row entries of the unit's tiles, up to CHUNK_MAX of them (entries past
    // the row are not used)
    auto row_at = [&](int code, int &jb, int &ms) {
        const int b = code >> 11, t0 = code & (MAX_ROW - 1);
        const size_t at = (size_t)b * MAX_ROW + min(t0 + (lane & (CHUNK_MAX - 1)), MAX_ROW - 1);
        jb = P.row_j[at];
        ms = P.row_m[at];
    };
    int nx_jb = 0, nx_ms = -1;
    if (k_cur < n_g) row_at(code_cur, nx_jb, nx_ms);
    for (;;) {
        if (k_cur >= n_g) {
            if (k_cur == n_g + waves_g - 1 && lane == 0) atomicExch(ctr, 0u);
            break;
        }
        const int u = u_begin + g + k_cur * G;
        const int code = code_cur;
        if (lane == 0) k_pending = atomicAdd(ctr, 1u);
        int next_code = 0;
        {
            const int b = code >> 11, q = (code >> 9) & 3, t0 = code & (MAX_ROW - 1);
            const int my_jb = nx_jb, my_ms = nx_ms;
            // the current box of each tile's j-block
            const float4 tb_c = P.qbox[((size_t)my_jb * QBOX_PER_BLOCK + QUARTERS) * 2];
            const float4 tb_h = P.qbox[((size_t)my_jb * QBOX_PER_BLOCK + QUARTERS) * 2 + 1];
            const int s = P.block_sys[b];
            const SysDesc &D = P.sys[s];
            const int nrow = P.row_n[b];
            const int t1 = min(nrow, t0 + D.chunk);
            const int c = t0 == 0 ? 0 : 1;  // first unit of the quarter or not
            const float reach2 = D.reach2;
            const unsigned lo_bits = __float_as_uint(0.01f);
            const unsigned span = __float_as_uint(D.cutoff2) - lo_bits;
            const float inv_rc2 = D.inv_rc2, inv_rc = D.inv_rc;

            const int slot_i = b * BLOCK + q * SUB + il;
            const float4 xi = P.xq[slot_i];
            const float2 li = P.lj[slot_i];
            const int ai = P.perm[slot_i];
            // the quarter's current box, formed once per evaluation
            // (quarter_boxes) rather than by each of its units
            const float4 qc = P.qbox[((size_t)b * QBOX_PER_BLOCK + q) * 2];
            const float4 qh = P.qbox[((size_t)b * QBOX_PER_BLOCK + q) * 2 + 1];
            const float cx = qc.x, cy = qc.y, cz = qc.z;
            const float hx = qh.x, hy = qh.y, hz = qh.z;

            // The unit's live tiles: a tile whose j-block's current box lies
            // beyond the cutoff of the quarter's current box holds no pair and
            // is skipped (about half of a row's tiles are, for a block
            // quarter: the row was built at cutoff + skin from build-time boxes).
            const int n_t = t1 - t0;
            const float tx = fmaxf(fabsf(cx - tb_c.x) - hx - tb_h.x, 0.0f);
            const float ty = fmaxf(fabsf(cy - tb_c.y) - hy - tb_h.y, 0.0f);
            const float tz = fmaxf(fabsf(cz - tb_c.z) - hz - tb_h.z, 0.0f);
            unsigned long long tiles = __ballot(lane < n_t && tx * tx + ty * ty + tz * tz <= reach2);
            k_cur = __builtin_amdgcn_readfirstlane((int)k_pending);
            if (k_cur < n_g) next_code = P.units[u_begin + g + k_cur * G];
            const bool any_tile = tiles != 0ull;
            float fx = 0.0f, fy = 0.0f, fz = 0.0f, e = 0.0f;
            // The ring holds `queued` survivors from `head` (0 or RING/2) on;
            // a full batch is always the contiguous [head, head + RING/2), and
            // appends wrap at most once (queued < RING/2 before an append of
            // at most 64).  Wave-uniform.
            int head = 0, queued = 0;
            // One live tile k with its j-atoms (xj, lj) already loaded: cull
            // the 64 j-atoms against the quarter's box, then evaluate a masked
            // tile's survivors at once or queue an unmasked tile's in the ring.
            auto tile = [&](const float4 xj, const float2 lj, int k) {
                const int ms = __builtin_amdgcn_readlane(my_ms, k);
                const float ex = fmaxf(fabsf(xj.x - cx) - hx, 0.0f);
                const float ey = fmaxf(fabsf(xj.y - cy) - hy, 0.0f);
                const float ez = fmaxf(fabsf(xj.z - cz) - hz, 0.0f);
                const bool active = ex * ex + ey * ey + ez * ez <= reach2;
                const unsigned long long m = __ballot(active);
                const int n_act = __popcll(m);
                if (ms >= 0 && n_act > 0) {
                    // a tile with skipped pairs: its survivors now, with their mask word
                    if (active) {
                        const int pos = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                        jpos[pos] = xj;
                        jpar[pos] = make_float4(lj.x, lj.y, __int_as_float(lane), 0.0f);
                    }
                    const int n_pad = (n_act + 3) & ~3;
                    if (lane >= n_act && lane < n_pad) {  // dummies: r2 ~ 1e14, outside every cutoff
                        jpos[lane] = make_float4(DUMMY_FAR, DUMMY_FAR, DUMMY_FAR, 0.0f);
                        jpar[lane] = make_float4(0.0f, 0.0f, __int_as_float(0), 0.0f);
                    }
                    wave_lds_fence();
                    const unsigned long long excl = P.masks[((size_t)b * MAX_MASKS + ms) * 64 + q * SUB + il];
                    tile_pairs<ENERGY, true>(xi, li, n_pad >> 2, js, jpos, jpar, excl, lo_bits, span, inv_rc, inv_rc2,
                                             fx, fy, fz, e);
                    wave_lds_fence();
                } else if (n_act > 0) {
                    if (active) {
                        const int pos = (head + queued + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u))) & (RING - 1);
                        rpos[pos] = xj;
                        rpar[pos] = lj;
                    }
                    queued += n_act;
                    if (queued >= RING / 2) {  // a full batch: 16 iterations of 16 i x 4 j
                        wave_lds_fence();
                        tile_pairs<ENERGY, false>(xi, li, RING / 8, js, rpos + head, rpar + head, 0ull, lo_bits, span,
                                                  inv_rc, inv_rc2, fx, fy, fz, e);
                        wave_lds_fence();  // read before the slots are written again
                        head ^= RING / 2;
                        queued -= RING / 2;
                    }
                }
            };
            // A software pipeline over the live tiles in two register sets
            // (A, B): the next tile's j-atoms load while this one computes, and
            // no registers rotate between tiles.  The next load is issued even
            // after the last tile (it reloads that tile, unused), so the wait
            // for the current tile never covers it.  The j-block index is
            // wave-uniform: each load is a scalar base plus the lane's offset.
            if (any_tile) {
                const unsigned ul = (unsigned)lane;
                int ka = __ffsll((long long)tiles) - 1, kb;
                tiles &= tiles - 1ull;
                const int ja = __builtin_amdgcn_readlane(my_jb, ka);
                float4 xa = (P.xq + (size_t)ja * BLOCK)[ul], xb;
                float2 la = (P.lj + (size_t)ja * BLOCK)[ul], lb;
                for (;;) {
                    bool more = tiles != 0ull;
                    kb = more ? __ffsll((long long)tiles) - 1 : ka;
                    tiles &= tiles - 1ull;
                    const int jnb = __builtin_amdgcn_readlane(my_jb, kb);
                    xb = (P.xq + (size_t)jnb * BLOCK)[ul];
                    lb = (P.lj + (size_t)jnb * BLOCK)[ul];
                    tile(xa, la, ka);
                    if (!more) break;
                    more = tiles != 0ull;
                    ka = more ? __ffsll((long long)tiles) - 1 : kb;
                    tiles &= tiles - 1ull;
                    const int jna = __builtin_amdgcn_readlane(my_jb, ka);
                    xa = (P.xq + (size_t)jna * BLOCK)[ul];
                    la = (P.lj + (size_t)jna * BLOCK)[ul];
                    tile(xb, lb, kb);
                    if (!more) break;
                }
            }
            code_cur = __builtin_amdgcn_readfirstlane(next_code);
            if (k_cur < n_g) row_at(code_cur, nx_jb, nx_ms);
            if (queued > 0) {  // the unit's last, partial batch, padded to whole iterations
                const int n_pad = (queued + 3) & ~3;
                if (lane >= queued && lane < n_pad) {
                    rpos[head + lane] = make_float4(DUMMY_FAR, DUMMY_FAR, DUMMY_FAR, 0.0f);
                    rpar[head + lane] = make_float2(0.0f, 0.0f);
                }
                wave_lds_fence();
                tile_pairs<ENERGY, false>(xi, li, n_pad >> 2, js, rpos + head, rpar + head, 0ull, lo_bits, span, inv_rc,
                                          inv_rc2, fx, fy, fz, e);
                wave_lds_fence();
            }
            // sum the four j-slot partials of each i-atom (lanes il, il+16, il+32,
            // il+48) in a fixed order and store the unit's partial; the gather
            // adds the partials of an atom's units in unit order
            fx += __shfl_xor(fx, 16); fy += __shfl_xor(fy, 16); fz += __shfl_xor(fz, 16);
            fx += __shfl_xor(fx, 32); fy += __shfl_xor(fy, 32); fz += __shfl_xor(fz, 32);
            if (ENERGY) {
                e += __shfl_xor(e, 16);
                e += __shfl_xor(e, 32);
            }
            if (!DIRECT) {
                if (js == 0) P.unit_f[(size_t)u * SUB + il] = make_float4(fx, fy, fz, ENERGY ? e : 0.0f);
            } else {
                long long add[3] = {to_fixed(fx, FORCE_FIXED_BITS), to_fixed(fy, FORCE_FIXED_BITS),
                                    to_fixed(fz, FORCE_FIXED_BITS)};
                if (c == 0) {  // the quarter's bonded slots, lane js taking slots js, js+4, ...
                    long long bs[3] = {0, 0, 0};
                    if (ai >= 0) {
                        for (int j = P.inc_off[ai] + js; j < P.inc_off[ai + 1]; j += 4) {
                            const float *cb = P.contrib + (size_t)(TM ? P.inc_idx[j] : j) * 3;
                            bs[0] += to_fixed(cb[0], FORCE_FIXED_BITS);
                            bs[1] += to_fixed(cb[1], FORCE_FIXED_BITS);
                            bs[2] += to_fixed(cb[2], FORCE_FIXED_BITS);
                        }
                    }
                    for (int k = 0; k < 3; ++k) {
                        bs[k] += wave_xor_ll(bs[k], 16);
                        bs[k] += wave_xor_ll(bs[k], 32);
                        add[k] += bs[k];
                    }
                }
                if (ai >= 0 && js < 3) {
                    const long long v = js == 0 ? add[0] : (js == 1 ? add[1] : add[2]);
                    atomicAdd((unsigned long long *)(P.f + (size_t)ai * 4 + js), (unsigned long long)v);
                }
            }
        }
    }
}

template <bool ENERGY, bool TM>
__global__ __launch_bounds__(256) void k_bonded(DevPtrs P, unsigned mask) {
    bonded_term<ENERGY, TM>(P, mask, blockIdx.x * blockDim.x + threadIdx.x);
}

// One wave per block quarter (16 sorted slots), four lanes per slot: lane
// (k, il) converts the nonbonded unit partials c = k, k+4, ... and the bonded
// contribution slots k, k+4, ... of slot il to int64 fixed point; the four
// lanes' sums are then combined.  Integer addition is associative, so the
// force is bitwise independent of this split.  For a given k the 16 lanes of a
// quarter read 256 contiguous bytes of each unit partial.
template <bool ENERGY, bool TM>
__global__ __launch_bounds__(256) void k_gather(DevPtrs P, unsigned mask) {
    const int lane = threadIdx.x & 63;
    const int il = lane & (SUB - 1), k = lane >> 4;
    const int quarter = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int slot = quarter * SUB + il;
    if (slot >= P.n_atoms_arena) return;  // whole waves: n_atoms_arena is a multiple of 64
    const int a = P.perm[slot];
    const int s = a >= 0 ? P.atom_sys[a] : -1;
    const bool live = s >= 0 && P.sys[s].active;
    long long acc[3] = {0, 0, 0};
    double e = 0.0;
    if (live) {
        if (mask & PRISM_MD_TERM_NONBONDED) {
            const int b = slot / BLOCK, q = (slot % BLOCK) / SUB;
            const int per_q = P.unit_count[b] / QUARTERS;
            const float4 *uf = P.unit_f + ((size_t)P.unit_off[b] + (size_t)q * per_q) * SUB + il;
#pragma unroll 2
            for (int c = k; c < per_q; c += 4) {
                const float4 p = uf[(size_t)c * SUB];
                acc[0] += to_fixed(p.x, FORCE_FIXED_BITS);
                acc[1] += to_fixed(p.y, FORCE_FIXED_BITS);
                acc[2] += to_fixed(p.z, FORCE_FIXED_BITS);
                if (ENERGY) e += p.w;
            }
        }
        for (int j = P.inc_off[a] + k; j < P.inc_off[a + 1]; j += 4) {
            const float *c = P.contrib + (size_t)(TM ? P.inc_idx[j] : j) * 3;
            acc[0] += to_fixed(c[0], FORCE_FIXED_BITS);
            acc[1] += to_fixed(c[1], FORCE_FIXED_BITS);
            acc[2] += to_fixed(c[2], FORCE_FIXED_BITS);
        }
    }
    for (int c = 0; c < 3; ++c) {
        acc[c] += wave_xor_ll(acc[c], 16);
        acc[c] += wave_xor_ll(acc[c], 32);
    }
    if (ENERGY) {
        e += __shfl_xor(e, 16);
        e += __shfl_xor(e, 32);
    }
    if (live && k == 0) {
        long long *f = P.f + (size_t)a * 4;
        f[0] = acc[0];
        f[1] = acc[1];
        f[2] = acc[2];
        if (ENERGY) P.atom_e[a] = e;
    }
}

// One workgroup per system: FP64 sums in a fixed order of the per-atom
// nonbonded energies and the per-term energies of the last evaluation.
__global__ __launch_bounds__(256) void k_energy_reduce(DevPtrs P) {
    const int s = blockIdx.x;
    if (s >= P.n_sys) return;
    const SysDesc &D = P.sys[s];
    __shared__ double red[256];
    const int begins[5] = {D.bond_begin, D.angle_begin, D.dih_begin, -1, D.exc_begin};
    const int counts[5] = {D.n_bonds, D.n_angles, D.n_dih, D.n_atoms, D.n_exc};
    const int bases[5] = {0, P.n_bonds, P.n_bonds + P.n_angles, 0, P.n_bonds + P.n_angles + P.n_dih};
    for (int term = 0; term < 5; ++term) {
        double acc = 0.0;
        for (int i = threadIdx.x; i < counts[term]; i += blockDim.x)
            acc += term == 3 ? P.atom_e[D.atom_begin + i] : P.term_e[bases[term] + begins[term] + i];
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) P.sst[s].energy[term] = red[0];
        __syncthreads();
    }
}

__global__ void k_zero_forces(DevPtrs P) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (size_t)P.n_atoms_arena * 4) P.f[i] = 0;
}

// (dx*dx + dy*dy) + dz*dz with dx = b - a, every operation rounded on its own.
// HIP's __fadd_rn / __fmul_rn are the plain operators here, and hipcc
// contracts a*b + c into an FMA by default, which rounds once instead of
// twice: a pair at FP32 r2 = 144.0 exactly then tested as > 144 and left the
// list (found at 250,000 atoms).  Contraction is switched off for this
// function, so the test rounds as the reference's list does.
__device__ inline float r2_list(float3 a, float3 b) {
#pragma clang fp contract(off)
    const float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
    return (dx * dx + dy * dy) + dz * dz;
}
__device__ inline bool in_list(const int *off, const int *idx, int a, int b) {
    for (int e = off[a]; e < off[a + 1]; ++e)
        if (idx[e] == b) return true;
    return false;
}
__device__ inline int ref_cell(float x, float o, int n) {  // nhs_amber_fused.cu:11704-11712
    const int c = (int)((x - o) * 0.1f);  // a subtract then a multiply: nothing to contract
    return max(0, min(c, n - 1));
}

// Pair enumeration over the tile rows.  mode 0 / 1: the parity pair list
// (reference cell scheme / brute force) -- unique pairs a_i < a_j with FP32
// r2 <= 144 that are not 1-2/1-3 exclusions; mode 2: count the pairs the
// nonbonded kernel evaluates (r2 < cutoff^2, r2 >= 0.01, not skipped).
__global__ __launch_bounds__(256) void k_pairs(DevPtrs P, int mode, RefGrid g, int *out_i, int *out_j,
                                               unsigned long long *count, unsigned long long cap) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int b = blockIdx.x * 4 + wave;
    if (b >= P.n_blocks_arena) return;
    const int s = P.block_sys[b];
    if (s < 0 || !P.sys[s].active) return;
    const int ai = P.perm[b * BLOCK + lane];
    float3 xi = make_float3(0, 0, 0);
    if (ai >= 0) xi = ld3(P.x, ai);
    int gi[3] = {0, 0, 0};
    if (mode == 0 && ai >= 0) {
        gi[0] = ref_cell(xi.x, g.ox, g.nx);
        gi[1] = ref_cell(xi.y, g.oy, g.ny);
        gi[2] = ref_cell(xi.z, g.oz, g.nz);
    }
    const float cutoff2 = P.sys[s].cutoff2;
    unsigned long long local = 0;
    const int nrow = P.row_n[b];
    for (int t = 0; t < nrow; ++t) {
        const int jb = P.row_j[(size_t)b * MAX_ROW + t];
        for (int k = 0; k < BLOCK; ++k) {
            const int aj = P.perm[jb * BLOCK + k];
            if (aj < 0 || ai < 0 || ai >= aj) continue;
            const float3 xj = ld3(P.x, aj);
            bool keep;
            if (mode == 2) {
                const float r2 = pair_r2(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z);
                keep = r2 < cutoff2 && r2 >= 0.01f && !in_list(P.skip_off, P.skip_idx, ai, aj);
            } else {
                keep = !(r2_list(xi, xj) > 144.0f) && !in_list(P.ex_off, P.ex_idx, ai, aj);
                if (keep && mode == 0) {
                    keep = abs(ref_cell(xj.x, g.ox, g.nx) - gi[0]) <= 1 && abs(ref_cell(xj.y, g.oy, g.ny) - gi[1]) <= 1 &&
                           abs(ref_cell(xj.z, g.oz, g.nz) - gi[2]) <= 1;
                }
            }
            if (!keep) continue;
            if (mode == 2) {
                ++local;
            } else {
                const unsigned long long pos = atomicAdd(count, 1ull);
                if (pos < cap) {
                    out_i[pos] = ai;
                    out_j[pos] = aj;
                }
            }
        }
    }
    if (mode == 2) {
        for (int o = 32; o > 0; o >>= 1) local += __shfl_xor(local, o);
        if (lane == 0) atomicAdd(count, local);
    }
}

}  // namespace

// Workgroups of a nonbonded kernel variant that the device holds at once
// (occupancy x compute units), queried once per device.
template <bool ENERGY, bool DIRECT, bool TM> static int resident_blocks() {
    static int cache[64];
    int dev = 0;
    (void)hipGetDevice(&dev);
    int &c = cache[dev & 63];
    if (c == 0) {
        int per_cu = 0, cus = 0;
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_nonbonded<ENERGY, DIRECT, TM>, 64 * NB_WAVES, 0);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        c = max(1, per_cu) * max(1, cus);
    }
    return c;
}

// start / stop, when given, are HIP events recorded with the launch itself
// (hipExtLaunchKernelGGL): they bracket the kernel's execution, not the gap
// before its dispatch.
template <bool ENERGY, bool DIRECT, bool TM>
static void launch_nb(const DevPtrs &P, hipStream_t st, hipEvent_t start = nullptr, hipEvent_t stop = nullptr) {
    // persistent grid: every wave resident at once (no second round of
    // workgroups), but no more waves than a typical unit list needs
    const int want = (P.n_blocks_arena * QUARTERS * 12 + NB_WAVES - 1) / NB_WAVES;
    const int res = resident_blocks<ENERGY, DIRECT, TM>();
    int grid = max(1, min(want, res));
    if (res % N_XCD == 0) grid = (grid + N_XCD - 1) / N_XCD * N_XCD;  // whole XCD rounds
    if (start || stop)
        hipExtLaunchKernelGGL((k_nonbonded<ENERGY, DIRECT, TM>), dim3(grid), dim3(64 * NB_WAVES), 0, st, start, stop,
                              0, P);
    else
        hipLaunchKernelGGL((k_nonbonded<ENERGY, DIRECT, TM>), dim3(grid), dim3(64 * NB_WAVES), 0, st, P);
}

void launch_nonbonded(const DevPtrs &P, bool energy, hipStream_t st, hipEvent_t start, hipEvent_t stop, bool boxes) {
    if (boxes) launch_quarter_boxes(P, st);
    if (energy) launch_nb<true, false, false>(P, st, start, stop);  // stores unit partials: no bonded slots read
    else launch_nb<false, false, false>(P, st, start, stop);
}

void launch_nonbonded_direct(const DevPtrs &P, hipStream_t st, hipEvent_t start, hipEvent_t stop) {
    if (P.term_major) launch_nb<false, true, true>(P, st, start, stop);
    else launch_nb<false, true, false>(P, st, start, stop);
}

void launch_bonded(const DevPtrs &P, unsigned term_mask, bool energy, hipStream_t st) {
    const int n = P.n_bonds + P.n_angles + P.n_dih + P.n_exc;
    if (n == 0) return;
    const int grid = (n + 255) / 256;
    if (P.term_major) {
        if (energy) hipLaunchKernelGGL((k_bonded<true, true>), dim3(grid), dim3(256), 0, st, P, term_mask);
        else hipLaunchKernelGGL((k_bonded<false, true>), dim3(grid), dim3(256), 0, st, P, term_mask);
    } else {
        if (energy) hipLaunchKernelGGL((k_bonded<true, false>), dim3(grid), dim3(256), 0, st, P, term_mask);
        else hipLaunchKernelGGL((k_bonded<false, false>), dim3(grid), dim3(256), 0, st, P, term_mask);
    }
}

void launch_gather(const DevPtrs &P, unsigned term_mask, bool energy, hipStream_t st) {
    const int grid = (P.n_atoms_arena / SUB * 64 + 255) / 256;  // one wave per block quarter
    if (P.term_major) {
        if (energy) hipLaunchKernelGGL((k_gather<true, true>), dim3(grid), dim3(256), 0, st, P, term_mask);
        else hipLaunchKernelGGL((k_gather<false, true>), dim3(grid), dim3(256), 0, st, P, term_mask);
    } else {
        if (energy) hipLaunchKernelGGL((k_gather<true, false>), dim3(grid), dim3(256), 0, st, P, term_mask);
        else hipLaunchKernelGGL((k_gather<false, false>), dim3(grid), dim3(256), 0, st, P, term_mask);
    }
}

void launch_energy_reduce(const DevPtrs &P, hipStream_t st) {
    hipLaunchKernelGGL(k_energy_reduce, dim3(P.n_sys), dim3(256), 0, st, P);
}

void launch_zero_forces(const DevPtrs &P, hipStream_t st) {
    const size_t n = (size_t)P.n_atoms_arena * 4;
    hipLaunchKernelGGL(k_zero_forces, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, P);
}

void launch_emit_pairs(const DevPtrs &P, int mode, RefGrid g, int *out_i, int *out_j, unsigned long long *count,
                       unsigned long long cap, hipStream_t st) {
    const int grid = (P.n_blocks_arena + 3) / 4;
    hipLaunchKernelGGL(k_pairs, dim3(grid), dim3(256), 0, st, P, mode, g, out_i, out_j, count, cap);
}

void launch_count_pairs(const DevPtrs &P, unsigned long long *count, hipStream_t st) {
    const int grid = (P.n_blocks_arena + 3) / 4;
    hipLaunchKernelGGL(k_pairs, dim3(grid), dim3(256), 0, st, P, 2, RefGrid{}, (int *)nullptr, (int *)nullptr, count,
                       0ull);
}

}  // namespace prism
