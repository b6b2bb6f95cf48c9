// kernels_integrate.hip -- constrained Langevin OVRVO integrator.
//
// One lane per constraint group (a heavy atom with its <= 3 hydrogens, or a
// lone atom).  Because every constraint couples atoms of one group only, all
// per-atom and per-cluster phases between two force evaluations fuse into one
// kernel; the reference launches each as a separate grid-wide phase
// (fused_engine.rs:6223-6235, nhs_amber_fused.cu:3620-3850, 11128-11186):
//   first half  (phases OU_HALF_N, RATTLE_AFTER_OU_N, KICK_N, RATTLE_PRE_DRIFT,
//                DRIFT_SAVE_XN, SHAKE_POSITION), preceded by the COM removal of
//                steps s > 0, s % com_every == 0 (fused_engine.rs:19023);
//   second half (KICK_NP1, RATTLE_BEFORE_OU_NP1, OU_HALF_NP1, velocity safety
//                clamp, final RATTLE).
// Kicks, Ornstein-Uhlenbeck updates and drifts are FP32 as in the reference;
// SHAKE and RATTLE iterate in FP64 on the group's atoms.
#include "device_common.hpp"

namespace prism {

namespace {

constexpr int IW = 4;  // waves per workgroup
constexpr int HS_GROUP = 128;  // workgroups per first-level counter of the step-end handshake
constexpr unsigned HS_PUBLISHED = 1u << 16;  // count in the low 16 bits, publishers above

// Register-resident group state.  Every loop over the group's atoms or bonds
// below is fully unrolled with a runtime guard, so all indices are
// compile-time constants and nothing spills to scratch memory.
struct GroupState {
    int n;              // atoms in the group (1..4)
    int a[4];
    float m[4];
    float x[4][3], v[4][3];
    long long f[4][3];  // the carried force, raw int64 fixed point
    double w;           // the group's cumulative non-Hamiltonian work, kcal/mol
};

// What the first half reads at its end: each atom's sorted slot and its
// position at the last list build.
struct GroupTail {
    int slot[4];
    float xb[4][3];
};

#define FOR_ATOMS(k) _Pragma("unroll") for (int k = 0; k < 4; ++k) if (k < g.n)
#define FOR_H(h) _Pragma("unroll") for (int h = 1; h < 4; ++h) if (h < g.n)

// Solve the k x k (k <= 3) system M lam = b of one cluster, M symmetric
// positive definite, by cofactors in FP64.  Returns false when M is singular.
// It is solved as 3 x 3 with the unused rows and columns set to the identity
// and their right-hand sides to zero: one code path and one FP64 division for
// every group shape, so a wave that mixes shapes does not run three solvers
// one after another.  M and b are overwritten.
__device__ inline bool solve3(int k, double M[3][3], double b[3], double lam[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (i >= k) {
            for (int j = 0; j < 3; ++j) M[i][j] = M[j][i] = 0.0;
            M[i][i] = 1.0;
            b[i] = 0.0;
        }
    const double c00 = M[1][1] * M[2][2] - M[1][2] * M[2][1];
    const double c01 = M[1][2] * M[2][0] - M[1][0] * M[2][2];
    const double c02 = M[1][0] * M[2][1] - M[1][1] * M[2][0];
    const double det = M[0][0] * c00 + M[0][1] * c01 + M[0][2] * c02;
    if (!(fabs(det) > 1e-300)) return false;
    const double id = 1.0 / det;
    const double c10 = M[0][2] * M[2][1] - M[0][1] * M[2][2];
    const double c11 = M[0][0] * M[2][2] - M[0][2] * M[2][0];
    const double c12 = M[0][1] * M[2][0] - M[0][0] * M[2][1];
    const double c20 = M[0][1] * M[1][2] - M[0][2] * M[1][1];
    const double c21 = M[0][2] * M[1][0] - M[0][0] * M[1][2];
    const double c22 = M[0][0] * M[1][1] - M[0][1] * M[1][0];
    lam[0] = (c00 * b[0] + c10 * b[1] + c20 * b[2]) * id;
    lam[1] = (c01 * b[0] + c11 * b[1] + c21 * b[2]) * id;
    lam[2] = (c02 * b[0] + c12 * b[1] + c22 * b[2]) * id;
    return true;
}

// RATTLE of one cluster (rattle_cluster_velocities, nhs_amber_fused.cu:2175-2360).
// The reference iterates Gauss-Seidel over the X-H bonds with the update
// v_c += m_c^-1 lam r, v_h -= m_h^-1 lam r until r . (v_h - v_c) vanishes;
// those conditions are linear in the multipliers, so here the k x k system
//   sum_j [m_c^-1 (r_h . r_j) + m_h^-1 delta_hj (r_h . r_h)] lam_j = r_h . (v_h - v_c)
// is solved exactly in FP64: the fixed point of that iteration in one step.
// Returns 1, or -1 when the cluster is degenerate or the result misses tol.
__device__ int rattle(GroupState &g, const GroupParam &gp, double tol, int max_iter) {
    (void)max_iter;
    if (g.n < 2) return 0;
    const int k = g.n - 1;
    const double imc = gp.inv_mass_c, imh = gp.inv_mass_h;
    double r[3][3] = {}, u[3][3] = {}, M[3][3] = {}, b[3] = {}, lam[3] = {};
    FOR_H(h) {
        for (int c = 0; c < 3; ++c) {
            r[h - 1][c] = (double)g.x[h][c] - (double)g.x[0][c];
            u[h - 1][c] = (double)g.v[h][c] - (double)g.v[0][c];
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (i < k) {
            b[i] = r[i][0] * u[i][0] + r[i][1] * u[i][1] + r[i][2] * u[i][2];
#pragma unroll
            for (int j = 0; j < 3; ++j)
                if (j < k) {
                    const double rr = r[i][0] * r[j][0] + r[i][1] * r[j][1] + r[i][2] * r[j][2];
                    M[i][j] = imc * rr + (i == j ? imh * rr : 0.0);
                }
        }
    if (!solve3(k, M, b, lam)) return -1;
    double dv[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < 3; ++j)
        if (j < k)
            for (int c = 0; c < 3; ++c) dv[c] += lam[j] * r[j][c];
    double vc[3];
    for (int c = 0; c < 3; ++c) vc[c] = (double)g.v[0][c] + imc * dv[c];
    // residual test |r.u| / (|r| max(|u|, 1)) <= tol, squared so that the
    // lane's dependent chain holds no FP64 division or square root
    const double t = fmax(tol, 1e-12), t2 = t * t;
    bool bad = false, ok = true;
    FOR_H(h) {
        double vh[3];
        for (int c = 0; c < 3; ++c) vh[c] = (double)g.v[h][c] - imh * lam[h - 1] * r[h - 1][c];
        const double ux = vh[0] - vc[0], uy = vh[1] - vc[1], uz = vh[2] - vc[2];
        const double r2 = r[h - 1][0] * r[h - 1][0] + r[h - 1][1] * r[h - 1][1] + r[h - 1][2] * r[h - 1][2];
        const double res = r[h - 1][0] * ux + r[h - 1][1] * uy + r[h - 1][2] * uz;
        const double lhs = res * res, rhs = t2 * r2 * fmax(ux * ux + uy * uy + uz * uz, 1.0);
        bad |= !(r2 >= 1e-12) || !isfinite(lhs) || !isfinite(rhs);
        ok &= lhs <= rhs;
        for (int c = 0; c < 3; ++c) g.v[h][c] = (float)vh[c];
    }
    for (int c = 0; c < 3; ++c) g.v[0][c] = (float)vc[c];
    return (bad || !ok) ? -1 : 1;
}

// SHAKE of one cluster (shake_cluster_positions, nhs_amber_fused.cu:1889-2173).
// The reference moves the atoms along the step-start bond vectors o_h,
//   x_c += m_c^-1 lam_h o_h,  x_h -= m_h^-1 lam_h o_h,  v += dx / dt,
// sweeping Gauss-Seidel until |r_h|^2 = d_h^2.  Here the same k equations in
// the k multipliers are solved by Newton's method in FP64 (Jacobian
// dg_h/dlam_j = -2 (m_h^-1 delta_hj r_h . o_h + m_c^-1 r_h . o_j)), which
// reaches that fixed point in a few iterations.  Returns the iterations used,
// or -1 on failure.
__device__ int shake(GroupState &g, const float xs[4][3], const GroupParam &gp, float dt, double tol, int max_iter) {
    if (g.n < 2) return 0;
    const int k = g.n - 1;
    const double imc = gp.inv_mass_c, imh = gp.inv_mass_h;
    double o[3][3] = {}, r0[3][3] = {}, lam[3] = {0.0, 0.0, 0.0}, d2[3] = {}, o2[3] = {};
    FOR_H(h) {
        for (int c = 0; c < 3; ++c) {
            o[h - 1][c] = (double)xs[h][c] - (double)xs[0][c];
            r0[h - 1][c] = (double)g.x[h][c] - (double)g.x[0][c];
        }
        d2[h - 1] = (double)gp.d[h - 1] * (double)gp.d[h - 1];
        o2[h - 1] = o[h - 1][0] * o[h - 1][0] + o[h - 1][1] * o[h - 1][1] + o[h - 1][2] * o[h - 1][2];
    }
    int used = 0;
    bool converged = false, bad = false;
    while (!converged && !bad && used < max_iter) {
        ++used;
        // current bonds r_h(lam) = r0_h - m_h^-1 lam_h o_h - m_c^-1 sum_j lam_j o_j
        double sc[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int j = 0; j < 3; ++j)
            if (j < k)
                for (int c = 0; c < 3; ++c) sc[c] += lam[j] * o[j][c];
        double r[3][3] = {}, gres[3] = {}, M[3][3] = {}, dl[3] = {}, rr[3] = {};
        bool within = true;  // every |r^2 - d^2| / d^2 <= tol, tested as |r^2 - d^2| <= tol d^2
#pragma unroll
        for (int i = 0; i < 3; ++i)
            if (i < k) {
                for (int c = 0; c < 3; ++c) r[i][c] = r0[i][c] - imh * lam[i] * o[i][c] - imc * sc[c];
                rr[i] = r[i][0] * r[i][0] + r[i][1] * r[i][1] + r[i][2] * r[i][2];
                gres[i] = rr[i] - d2[i];
                within &= fabs(gres[i]) <= tol * d2[i];
                bad |= !isfinite(rr[i]) || !(d2[i] > 0.0);
            }
        if (bad) break;
        if (within) {
            converged = true;
            break;
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
            if (i < k) {
                const double ro = r[i][0] * o[i][0] + r[i][1] * o[i][1] + r[i][2] * o[i][2];
                // the reference's guard against a bond that rotated ~90 degrees,
                // |r.o| <= 1e-6 |r| |o|, squared (no square root)
                const double ss = rr[i] * o2[i];
                bad |= !(ss > 0.0) || ro * ro <= 1e-12 * ss;
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    if (j < k) {
                        const double rj = r[i][0] * o[j][0] + r[i][1] * o[j][1] + r[i][2] * o[j][2];
                        M[i][j] = 2.0 * (imc * rj + (i == j ? imh * ro : 0.0));  // -dg_i/dlam_j
                    }
            }
        if (bad || !solve3(k, M, gres, dl)) {
            bad = true;
            break;
        }
        for (int i = 0; i < 3; ++i) lam[i] += dl[i];  // g + J dl = 0 with J = -M
    }
    if (!converged) return -1;
    const double idt = 1.0 / (double)dt;
    double sc[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < 3; ++j)
        if (j < k)
            for (int c = 0; c < 3; ++c) sc[c] += lam[j] * o[j][c];
    for (int c = 0; c < 3; ++c) {
        g.x[0][c] = (float)((double)g.x[0][c] + imc * sc[c]);
        g.v[0][c] = (float)((double)g.v[0][c] + imc * sc[c] * idt);
    }
    FOR_H(h) {
        for (int c = 0; c < 3; ++c) {
            const double dxh = imh * lam[h - 1] * o[h - 1][c];
            g.x[h][c] = (float)((double)g.x[h][c] - dxh);
            g.v[h][c] = (float)((double)g.v[h][c] - dxh * idt);
        }
    }
    return used;
}

// |v|^2 in FP64 of an FP32 velocity (the work ledger's kinetic energies).
__device__ inline double speed2(const float v[3]) {
    return (double)v[0] * (double)v[0] + (double)v[1] * (double)v[1] + (double)v[2] * (double)v[2];
}
// 0.5 m (|v_after|^2 - |v_before|^2) / 418.4, the kinetic-energy change every
// entry of the reference's work ledger books (nhs_amber_fused.cu:1734-1743,
// :1776-1785, :11178-11185).
__device__ inline double ke_change(double m, double before2, double after2) {
    return 0.5 * m * (after2 - before2) * (1.0 / KCAL_TO_AMU_A2_PS2);
}

// langevin_thermostat with dt/2 (nhs_amber_fused.cu:1713-1744, :3656-3666),
// booking each atom's kinetic-energy change in the work ledger.
__device__ inline void ou_half(GroupState &g, const SysDesc &D, float T, float gamma, unsigned long long step,
                               int purpose, const float *noise) {
    const float c1 = expf(-gamma * (0.5f * D.dt));
    FOR_ATOMS(k) {
        float xi[3];
        if (noise) {
            for (int c = 0; c < 3; ++c) xi[c] = noise[(size_t)g.a[k] * 3 + c];
        } else {
            langevin_normals(D.seed_lo, D.stream, (uint32_t)(g.a[k] - D.atom_begin), step, purpose, xi);
        }
        // hardware reciprocal and square root (~1 ulp): the IEEE-rounded
        // forms cost ~10 dependent instructions each on a latency-bound lane
        const float c2 = __builtin_amdgcn_sqrtf((1.0f - c1 * c1) * (float)KCAL_TO_AMU_A2_PS2 * (float)KB * T *
                                                __builtin_amdgcn_rcpf(g.m[k]));
        const double before = speed2(g.v[k]);
        for (int c = 0; c < 3; ++c) g.v[k][c] = c1 * g.v[k][c] + c2 * xi[c];
        g.w += ke_change(g.m[k], before, speed2(g.v[k]));
    }
}

// Half kick with the 1000 kcal/mol/A clamp of prism_asc_used_force
// (prism_asc_decision.cuh:100-104) when gamma != 0, exact when gamma == 0
// (nhs_amber_fused.cu:1750-1805, :3667-3703).
// consume: zero the accumulators after reading them -- the step's force
// evaluation adds into them (first half only).
__device__ inline void half_kick(GroupState &g, const DevPtrs &P, float dt, float gamma, bool consume) {
    FOR_ATOMS(k) {
        float f[3];
        for (int c = 0; c < 3; ++c) f[c] = from_fixed(g.f[k][c], FORCE_FIXED_BITS);
        if (consume)
            for (int c = 0; c < 3; ++c) P.f[(size_t)g.a[k] * 4 + c] = 0;
        const float kick = 0.5f * dt * (float)KCAL_TO_AMU_A2_PS2 * __builtin_amdgcn_rcpf(g.m[k]);
        if (gamma != 0.0f) {  // |f| > MAX_FORCE tested squared; the rescale is rare
            const float mag2 = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
            if (mag2 > MAX_FORCE * MAX_FORCE) {
                // the clamp's work: used minus unclamped candidate velocity
                // (apply_force_half_kick_with_intervention_work, :1750-1788)
                float raw[3];
                for (int c = 0; c < 3; ++c) raw[c] = g.v[k][c] + kick * f[c];
                const float q = MAX_FORCE * __builtin_amdgcn_rsqf(mag2);
                for (int c = 0; c < 3; ++c) f[c] *= q;
                for (int c = 0; c < 3; ++c) g.v[k][c] += kick * f[c];
                g.w += ke_change(g.m[k], speed2(raw), speed2(g.v[k]));
                continue;
            }
        }
        for (int c = 0; c < 3; ++c) g.v[k][c] += kick * f[c];
    }
}

// Kinetic energy change of a RATTLE projection booked to the work ledger
// with the cluster's masses (rattle_cluster_velocities, :2245-2270): the
// projections that follow a bath update -- after OU_N and the final one after
// OU_NP1 and the velocity clamp -- book there (fused_engine.rs:19273-19277,
// :17735-17740); the pre-drift, SHAKE and pre-OU_NP1 solves go to the
// reference's separate constraint planes and are not part of W.
__device__ inline double group_ke2(const GroupState &g, const GroupParam &gp) {
    double s = 0.0;
    FOR_ATOMS(k) s += speed2(g.v[k]) / (double)(k == 0 ? gp.inv_mass_c : gp.inv_mass_h);
    return s;
}

// Load a constraint group and its atoms' state in two rounds of memory
// traffic: the record and its parameters together (every group index of the
// arena is in bounds, padding included), then everything each atom needs as
// soon as its index is known (a group's atom k exists exactly when k <= n_h).
// The atom loads are unconditional -- an absent atom reads arena atom 0,
// whose values are never used -- so no branch splits them into one wait per
// atom.  The integrator kernels run under one wave per SIMD, so each
// dependent round of loads is exposed latency.  No kernel that calls this
// writes another lane's atoms, so the early loads see the values the later
// uses would.
template <bool TAIL>
__device__ inline bool load_group(const DevPtrs &P, int gidx, GroupState &g, GroupParam &gp, GroupTail &t) {
    const Group G = P.groups[gidx];
    gp = P.gparams[gidx];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int a = G.atom[k];
        const size_t ac = (size_t)max(a, 0);
        g.a[k] = a;
        g.m[k] = P.mass[ac];
        const float4 x = P.x[ac], v = P.v[ac];
        g.x[k][0] = x.x; g.x[k][1] = x.y; g.x[k][2] = x.z;
        g.v[k][0] = v.x; g.v[k][1] = v.y; g.v[k][2] = v.z;
        for (int c = 0; c < 3; ++c) g.f[k][c] = P.f[ac * 4 + c];
        if (TAIL) {
            t.slot[k] = P.slot_of[ac];
            const float4 b = P.x_build[ac];
            t.xb[k][0] = b.x; t.xb[k][1] = b.y; t.xb[k][2] = b.z;
        }
    }
    g.w = P.work_g[gidx];
    if (G.atom[0] < 0) return false;
    g.n = 1 + gp.n_h;
    return true;
}

// RATTLE whose kinetic-energy change goes to the work ledger (the bath-side
// projections, see group_ke2).
__device__ inline int rattle_booked(GroupState &g, const GroupParam &gp, double tol, int max_iter) {
    if (g.n < 2) return 0;
    const double before = group_ke2(g, gp);
    const int r = rattle(g, gp, tol, max_iter);
    g.w += 0.5 * (group_ke2(g, gp) - before) * (1.0 / KCAL_TO_AMU_A2_PS2);
    return r;
}

__device__ inline bool finite3(const float a[3]) { return isfinite(a[0]) && isfinite(a[1]) && isfinite(a[2]); }

// Latch the first failure of a system: status bit plus kind, atom and step.
__device__ inline void record_fault(SysState &S, int kind, int atom, unsigned long long step) {
    atomicOr(&S.status, kind == FAULT_NONFINITE ? ST_NONFINITE : (kind == FAULT_DIVERGED ? ST_DIVERGED : ST_CONSTRAINT));
    if (atomicCAS(&S.fault_atom, -1, atom) == -1) {
        S.fault_kind = kind;
        S.fault_step = step;
    }
}
__device__ inline void note_sweeps(SysState &S, int sweeps, int kind, int atom, unsigned long long step, int &most) {
    if (sweeps < 0) record_fault(S, kind, atom, step);
    else most = max(most, sweeps);
}
// Sweep statistics: the wave's maximum raises the running maximum `prev`
// (read at kernel start) of this wave's own slot; no atomics, called in
// converged control flow.
__device__ inline void publish_sweeps(int *slot, int prev, int most) {
    for (int o = 32; o > 0; o >>= 1) most = max(most, __shfl_xor(most, o));
    if ((threadIdx.x & 63) == 0 && most > prev) *slot = most;
}

__global__ __launch_bounds__(64 * IW) void k_first_half(DevPtrs P, const float *noise) {
    const int lane = threadIdx.x & 63;
    const int gw = blockIdx.x * IW + wave_id();
    if (gw >= P.n_waves_arena) return;
    const int s = P.wave_sys[gw];
    if (s < 0) return;
    const SysDesc &D = P.sys[s];
    if (!D.active) return;
    SysState &S = P.sst[s];
    if (S.halted || parked(D, S)) {  // frozen, but its forces are still re-evaluated into zeroed accumulators
        const Group G = P.groups[gw * 64 + lane];
        for (int k = 0; k < 4; ++k)
            if (G.atom[k] >= 0)
                for (int c = 0; c < 3; ++c) P.f[(size_t)G.atom[k] * 4 + c] = 0;
        return;
    }
    const unsigned long long step = S.step;
    float T, gamma;
    schedule(D, step + 1, T, gamma);
    const bool com = D.com_every > 0 && step > 0 && step % (unsigned long long)D.com_every == 0;
    const int prev_sh = P.sweeps[2 * gw], prev_ra = P.sweeps[2 * gw + 1];
    GroupState g;
    GroupParam gp;
    GroupTail tail;
    bool moved = false;
    int most_sh = 0, most_ra = 0;
    if (load_group<true>(P, gw * 64 + lane, g, gp, tail)) {
        if (com) {
            FOR_ATOMS(k) {
                for (int c = 0; c < 3; ++c) g.v[k][c] -= (float)S.com_v[c];
            }
        }
        if (gamma != 0.0f) ou_half(g, D, T, gamma, step, 1, noise);
        note_sweeps(S, rattle_booked(g, gp, D.rattle_tol, D.constraint_max_iter), FAULT_RATTLE_1, g.a[0], step, most_ra);
        half_kick(g, P, D.dt, gamma, true);
        note_sweeps(S, rattle(g, gp, D.rattle_tol, D.constraint_max_iter), FAULT_RATTLE_2, g.a[0], step, most_ra);
        float xs[4][3];
        FOR_ATOMS(k) {
            for (int c = 0; c < 3; ++c) {
                xs[k][c] = g.x[k][c];
                g.x[k][c] += D.dt * g.v[k][c];
            }
        }
        note_sweeps(S, shake(g, xs, gp, D.dt, D.shake_tol, D.constraint_max_iter), FAULT_SHAKE, g.a[0], step, most_sh);
        FOR_ATOMS(k) {
            const int a = g.a[k];
            if (!(finite3(g.x[k]) && finite3(g.v[k]))) record_fault(S, FAULT_NONFINITE, a, step);
            // heartbeat_check (housekeeping.cu:27-45): a coordinate beyond
            // 1000 A means the system has diverged
            else if (fmaxf(fmaxf(fabsf(g.x[k][0]), fabsf(g.x[k][1])), fabsf(g.x[k][2])) > MAX_COORD)
                record_fault(S, FAULT_DIVERGED, a, step);
            P.x[a] = make_float4(g.x[k][0], g.x[k][1], g.x[k][2], 0.0f);
            P.v[a] = make_float4(g.v[k][0], g.v[k][1], g.v[k][2], 0.0f);
            float *q = (float *)&P.xq[tail.slot[k]];
            q[0] = g.x[k][0]; q[1] = g.x[k][1]; q[2] = g.x[k][2];
            const float dx = g.x[k][0] - tail.xb[k][0], dy = g.x[k][1] - tail.xb[k][1], dz = g.x[k][2] - tail.xb[k][2];
            moved |= dx * dx + dy * dy + dz * dz > D.half_skin2;
        }
        P.work_g[gw * 64 + lane] = g.w;
    }
    publish_sweeps(P.sweeps + 2 * gw, prev_sh, most_sh);
    publish_sweeps(P.sweeps + 2 * gw + 1, prev_ra, most_ra);
    const bool due = (long long)step - S.rebuild_step >= (long long)D.rebuild_every;
    if (__ballot(moved) != 0ull || due) {
        if (lane == 0) {
            S.rebuild_flag = 1;
            P.counters->any_rebuild = 1;
        }
    }
}

__global__ __launch_bounds__(64 * IW) void k_second_half(DevPtrs P, const float *noise) {
    const int lane = threadIdx.x & 63;
    const int gw = blockIdx.x * IW + wave_id();
    __shared__ int last;
    bool publish = false;  // this thread wrote something the last workgroup reads
    const int s = gw < P.n_waves_arena ? P.wave_sys[gw] : -1;
    if (s >= 0 && P.sys[s].active && !P.sst[s].halted && !parked(P.sys[s], P.sst[s])) {
        const SysDesc &D = P.sys[s];
        SysState &S = P.sst[s];
        const unsigned long long step = S.step;
        float T, gamma;
        schedule(D, step + 1, T, gamma);
        const bool next_com = D.com_every > 0 && (step + 1) % (unsigned long long)D.com_every == 0;
        const int prev_ra = P.sweeps[2 * gw + 1];
        GroupState g;
        GroupParam gp;
        GroupTail unused;
        double mv[3] = {0.0, 0.0, 0.0};
        int most_ra = 0;
        if (load_group<false>(P, gw * 64 + lane, g, gp, unused)) {
            half_kick(g, P, D.dt, gamma, false);
            int sw = rattle(g, gp, D.rattle_tol, D.constraint_max_iter);
            publish |= sw < 0;
            note_sweeps(S, sw, FAULT_RATTLE_3, g.a[0], step, most_ra);
            if (gamma != 0.0f) ou_half(g, D, T, gamma, step, 2, noise);
            FOR_ATOMS(k) {  // nhs_velocity_safety_clamp (nhs_amber_fused.cu:11148-11186)
                // |v| > max(20 sqrt(418.4 kB T / max(m, 1)), 2) tested squared
                const float vmax2 = fmaxf(400.0f * (float)KCAL_TO_AMU_A2_PS2 * (float)KB * T *
                                              __builtin_amdgcn_rcpf(fmaxf(g.m[k], 1.0f)), 4.0f);
                const float sp2 = g.v[k][0] * g.v[k][0] + g.v[k][1] * g.v[k][1] + g.v[k][2] * g.v[k][2];
                if (sp2 > vmax2) {
                    const double before = speed2(g.v[k]);
                    const float sc = __builtin_amdgcn_sqrtf(vmax2) * __builtin_amdgcn_rsqf(sp2);
                    for (int c = 0; c < 3; ++c) g.v[k][c] *= sc;
                    g.w += ke_change(g.m[k], before, speed2(g.v[k]));
                }
            }
            sw = rattle_booked(g, gp, D.rattle_tol, D.constraint_max_iter);
            publish |= sw < 0;
            note_sweeps(S, sw, FAULT_RATTLE_4, g.a[0], step, most_ra);
            FOR_ATOMS(k) {
                if (!finite3(g.v[k])) {
                    record_fault(S, FAULT_NONFINITE, g.a[k], step);
                    publish = true;
                }
                P.v[g.a[k]] = make_float4(g.v[k][0], g.v[k][1], g.v[k][2], 0.0f);
                for (int c = 0; c < 3; ++c) mv[c] += (double)g.m[k] * g.v[k][c];
            }
            P.work_g[gw * 64 + lane] = g.w;
        }
        publish_sweeps(P.sweeps + 2 * gw + 1, prev_ra, most_ra);
        if (next_com) {
            for (int c = 0; c < 3; ++c) mv[c] = wave_sum(mv[c]);
            if (lane == 0) {
                for (int c = 0; c < 3; ++c) P.com_partial[(size_t)gw * 3 + c] = mv[c];
                publish = true;
            }
        }
    }
    // Last-block handshake (no waiting): the workgroup that retires last sums
    // the COM partials in wave order and advances every system's step counter.
    // A workgroup releases its writes (agent-scope fence) only when the last
    // one reads any of them -- COM partials, a latched fault: the fence writes
    // back this XCD's L2 and costs microseconds per workgroup, so on the other
    // steps the counter increment goes out unfenced.  The count has two
    // levels, HS_GROUP workgroups per group counter and one increment per
    // group on the global one: atomics on one address serialise (~12 ns
    // each), which a flat count over the thousands of workgroups of a large
    // arena would pay on every step.  A workgroup that published adds
    // HS_PUBLISHED, so the last of its group knows to pass the writes on.
    const int pub = __syncthreads_or(publish);
    if (pub) __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned g = blockIdx.x / HS_GROUP;
        const unsigned gsize = min((unsigned)HS_GROUP, gridDim.x - g * HS_GROUP);
        const unsigned n_groups = (gridDim.x + HS_GROUP - 1) / HS_GROUP;
        const unsigned old = atomicAdd(&P.hs_group[g], pub ? HS_PUBLISHED + 1u : 1u);
        int is_last = 0;
        if ((old & (HS_PUBLISHED - 1u)) == gsize - 1) {  // the group's last workgroup
            P.hs_group[g] = 0u;  // for the next launch; no other workgroup touches it now
            if (n_groups == 1) {  // a small arena: one level
                is_last = 1;
            } else {
                const bool gpub = pub || old >= HS_PUBLISHED;
                if (gpub) __threadfence();
                const unsigned o2 = atomicAdd(&P.counters->done_groups, gpub ? HS_PUBLISHED + 1u : 1u);
                is_last = (o2 & (HS_PUBLISHED - 1u)) == n_groups - 1;
            }
        }
        last = is_last;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    for (int q = threadIdx.x; q < P.n_sys; q += blockDim.x) {
        const SysDesc &D = P.sys[q];
        if (!D.active) continue;
        SysState &S = P.sst[q];
        if (S.halted || parked(D, S)) continue;
        const unsigned long long step = S.step;
        if (D.com_every > 0 && (step + 1) % (unsigned long long)D.com_every == 0) {
            double acc[3] = {0.0, 0.0, 0.0};
            for (int w = D.wave_begin; w < D.wave_begin + D.n_waves; ++w)
                for (int c = 0; c < 3; ++c) acc[c] += P.com_partial[(size_t)w * 3 + c];
            for (int c = 0; c < 3; ++c) S.com_v[c] = acc[c] * D.inv_total_mass;
        }
        if (S.rebuild_flag) {  // this step's rebuild is done: consume its flags
            S.rebuild_flag = 0;
            S.resort_flag = 0;
            S.rebuild_step = (long long)step;
            S.n_rebuilds += 1;
        }
        S.step = step + 1;
        // a latched fault stops this system here; the others keep running
        if (S.status & ST_HALT) S.halted = 1;
    }
    if (threadIdx.x == 0) {
        P.counters->done_groups = 0u;
        P.counters->any_dirty = 0;
        P.counters->any_rebuild = 0;
        // the step's end on the device's constant-rate clock: the host turns
        // consecutive stamps into per-step latencies that include every gap
        // between steps (host work, admissions, other launches)
        const unsigned long long k = P.counters->step_ends;
        P.step_ts[k % TS_RING] = (unsigned long long)wall_clock64();
        P.counters->step_ends = k + 1;
    }
}

// Per-system reductions in FP64 (one workgroup per system, fixed order):
// kinetic energy, centre-of-mass velocity and the cumulative work ledger (the
// role of nhs_reduce_external_work_components, nhs_amber_fused.cu:11097-11126,
// summed here over constraint groups instead of atoms).  The COM velocity
// carried to the next COM-removal step is overwritten only for com_sys (a
// system index, -2 for all, -1 for none), so reading energies never perturbs
// a trajectory.
__global__ __launch_bounds__(256) void k_sys_reduce(DevPtrs P, double *out /*[S][RED_STRIDE]*/, int com_sys) {
    const int s = blockIdx.x;
    if (s >= P.n_sys) return;
    const SysDesc &D = P.sys[s];
    __shared__ double red[4][5];
    double ke = 0.0, p[3] = {0.0, 0.0, 0.0}, w = 0.0;
    for (int i = threadIdx.x; i < D.n_atoms; i += blockDim.x) {
        const int a = D.atom_begin + i;
        const float4 v = P.v[a];
        const double m = P.mass[a];
        ke += 0.5 * m * ((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z);
        p[0] += m * v.x; p[1] += m * v.y; p[2] += m * v.z;
    }
    for (int i = threadIdx.x; i < D.n_waves * 64; i += blockDim.x) w += P.work_g[(size_t)D.wave_begin * 64 + i];
    ke = wave_sum(ke);
    for (int c = 0; c < 3; ++c) p[c] = wave_sum(p[c]);
    w = wave_sum(w);
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) { red[wv][0] = ke; red[wv][1] = p[0]; red[wv][2] = p[1]; red[wv][3] = p[2]; red[wv][4] = w; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r[5] = {0, 0, 0, 0, 0};
        for (int k = 0; k < 4; ++k)
            for (int c = 0; c < 5; ++c) r[c] += red[k][c];
        out[s * RED_STRIDE + 0] = r[0] / KCAL_TO_AMU_A2_PS2;
        for (int c = 0; c < 3; ++c) {
            out[s * RED_STRIDE + 1 + c] = r[1 + c] * D.inv_total_mass;
            if (com_sys == -2 || com_sys == s) P.sst[s].com_v[c] = r[1 + c] * D.inv_total_mass;
        }
        out[s * RED_STRIDE + 4] = r[4];
    }
}

}  // namespace

void launch_first_half(const DevPtrs &P, const float *noise, hipStream_t st) {
    const int grid = (P.n_waves_arena + IW - 1) / IW;
    hipLaunchKernelGGL(k_first_half, dim3(grid), dim3(64 * IW), 0, st, P, noise);
}

void launch_second_half(const DevPtrs &P, const float *noise, hipStream_t st) {
    const int grid = (P.n_waves_arena + IW - 1) / IW;
    hipLaunchKernelGGL(k_second_half, dim3(grid), dim3(64 * IW), 0, st, P, noise);
}

void launch_kinetic(const DevPtrs &P, double *out, int com_sys, hipStream_t st) {
    hipLaunchKernelGGL(k_sys_reduce, dim3(P.n_sys), dim3(256), 0, st, P, out, com_sys);
}

}  // namespace prism
