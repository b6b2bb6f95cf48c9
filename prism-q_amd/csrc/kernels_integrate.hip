// kernels_integrate.hip -- constrained Langevin OVRVO integrator.
//
// Work is laid out by constraint group (a heavy atom with its <= 3 hydrogens,
// or a lone atom), a lane or a quad of lanes per group (the two layouts are
// described above k_first_half1).  Because every constraint couples atoms of
// one group only, all per-atom and per-cluster phases between two force
// evaluations fuse into one kernel; the reference launches each as a
// separate grid-wide phase
// (fused_engine.rs:6223-6235, nhs_amber_fused.cu:3620-3850, 11128-11186):
//   first half  (phases OU_HALF_N, RATTLE_AFTER_OU_N, KICK_N, RATTLE_PRE_DRIFT,
//                DRIFT_SAVE_XN, SHAKE_POSITION), preceded by the COM removal of
//                steps s > 0, s % com_every == 0 (fused_engine.rs:19023);
//   second half (KICK_NP1, RATTLE_BEFORE_OU_NP1, OU_HALF_NP1, velocity safety
//                clamp, final RATTLE).
// Kicks, Ornstein-Uhlenbeck updates and drifts are FP32 as in the reference;
// SHAKE and RATTLE iterate in FP64 on the group's atoms.
// Contraction into FMAs per source expression only (the C++ standard's rule),
// here and in the shared device helpers: the two group layouts inline the
// same phase functions into different kernels, and the default (fusing across
// statements wherever the instruction selector finds a multiply feeding an
// add) could round them differently.
#pragma clang fp contract(on)
#include "device_common.hpp"

namespace prism {

namespace {

constexpr int IW = 4;  // waves per workgroup (the four-lane layout needs exactly 4: a workgroup per group wave)
constexpr int QUAD_LAYOUT_MAX_WAVES = 512;
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
//
// The matrix depends on the positions only, and the two RATTLE projections
// of each half step see the same positions, so the solve is split: a factor
// (the bond vectors and the cofactors of solve3's padded 3 x 3 system) once
// per half step, then one application per projection -- the same operations
// on the same operands as solve3, so bitwise the same multipliers.
struct RattleFactor {
    double r[3][3];  // bond vectors x_h - x_c
    double c[3][3];  // c[i][j]: cofactor (i, j) of the padded matrix
    double id;       // 1 / det
    bool ok;         // det nonzero
};

__device__ inline void rattle_factor(const GroupState &g, const GroupParam &gp, RattleFactor &F) {
    const int k = g.n - 1;
    const double imc = gp.inv_mass_c, imh = gp.inv_mass_h;
    double M[3][3] = {};
#pragma unroll
    for (int i = 0; i < 3; ++i)
        for (int c = 0; c < 3; ++c) F.r[i][c] = 0.0;
    FOR_H(h) {
        for (int c = 0; c < 3; ++c) F.r[h - 1][c] = (double)g.x[h][c] - (double)g.x[0][c];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (i < k) {
#pragma unroll
            for (int j = 0; j < 3; ++j)
                if (j < k) {
                    const double rr = F.r[i][0] * F.r[j][0] + F.r[i][1] * F.r[j][1] + F.r[i][2] * F.r[j][2];
                    M[i][j] = imc * rr + (i == j ? imh * rr : 0.0);
                }
        }
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (i >= k) {
            for (int j = 0; j < 3; ++j) M[i][j] = M[j][i] = 0.0;
            M[i][i] = 1.0;
        }
    const double c00 = M[1][1] * M[2][2] - M[1][2] * M[2][1];
    const double c01 = M[1][2] * M[2][0] - M[1][0] * M[2][2];
    const double c02 = M[1][0] * M[2][1] - M[1][1] * M[2][0];
    const double det = M[0][0] * c00 + M[0][1] * c01 + M[0][2] * c02;
    F.ok = fabs(det) > 1e-300;
    F.id = 1.0 / det;
    F.c[0][0] = c00; F.c[0][1] = c01; F.c[0][2] = c02;
    F.c[1][0] = M[0][2] * M[2][1] - M[0][1] * M[2][2];
    F.c[1][1] = M[0][0] * M[2][2] - M[0][2] * M[2][0];
    F.c[1][2] = M[0][1] * M[2][0] - M[0][0] * M[2][1];
    F.c[2][0] = M[0][1] * M[1][2] - M[0][2] * M[1][1];
    F.c[2][1] = M[0][2] * M[1][0] - M[0][0] * M[1][2];
    F.c[2][2] = M[0][0] * M[1][1] - M[0][1] * M[1][0];
}

// Returns 1, or -1 when the cluster is degenerate or the result misses tol.
__device__ int rattle_apply(GroupState &g, const GroupParam &gp, const RattleFactor &F, double tol) {
    if (g.n < 2) return 0;
    const int k = g.n - 1;
    const double imc = gp.inv_mass_c, imh = gp.inv_mass_h;
    double u[3][3] = {}, b[3] = {}, lam[3] = {};
    FOR_H(h) {
        for (int c = 0; c < 3; ++c) u[h - 1][c] = (double)g.v[h][c] - (double)g.v[0][c];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (i < k) b[i] = F.r[i][0] * u[i][0] + F.r[i][1] * u[i][1] + F.r[i][2] * u[i][2];
    if (!F.ok) return -1;
    lam[0] = (F.c[0][0] * b[0] + F.c[1][0] * b[1] + F.c[2][0] * b[2]) * F.id;
    lam[1] = (F.c[0][1] * b[0] + F.c[1][1] * b[1] + F.c[2][1] * b[2]) * F.id;
    lam[2] = (F.c[0][2] * b[0] + F.c[1][2] * b[1] + F.c[2][2] * b[2]) * F.id;
    double dv[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < 3; ++j)
        if (j < k)
            for (int c = 0; c < 3; ++c) dv[c] += lam[j] * F.r[j][c];
    double vc[3];
    for (int c = 0; c < 3; ++c) vc[c] = (double)g.v[0][c] + imc * dv[c];
    // residual test |r.u| / (|r| max(|u|, 1)) <= tol, squared so that the
    // lane's dependent chain holds no FP64 division or square root
    const double t = fmax(tol, 1e-12), t2 = t * t;
    bool bad = false, ok = true;
    FOR_H(h) {
        double vh[3];
        for (int c = 0; c < 3; ++c) vh[c] = (double)g.v[h][c] - imh * lam[h - 1] * F.r[h - 1][c];
        const double ux = vh[0] - vc[0], uy = vh[1] - vc[1], uz = vh[2] - vc[2];
        const double *r = F.r[h - 1];
        const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        const double res = r[0] * ux + r[1] * uy + r[2] * uz;
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

// ---- per-atom phases: one atom's state in, its update out.  Both group
// layouts below call these, so they give bitwise the same trajectory.

// langevin_thermostat with dt/2 (nhs_amber_fused.cu:1713-1744, :3656-3666)
// of atom a, in two parts: the normals, which depend on the atom, the step
// and the phase only (so a kernel can draw them while its loads are in
// flight), then the update; returns its kinetic-energy change for the work
// ledger.  `a` may be -1 (no atom: the draw is then never used).
__device__ inline void ou_noise(float xi[3], int a, const SysDesc &D, unsigned long long step, int purpose,
                                const float *noise) {
    if (noise) {
        for (int c = 0; c < 3; ++c) xi[c] = noise[(size_t)max(a, 0) * 3 + c];
    } else {
        langevin_normals(D.seed_lo, D.stream, (uint32_t)(a - D.atom_begin), step, purpose, xi);
    }
}
__device__ inline double ou_apply(float v[3], float m, float c1, float T, const float xi[3]) {
    // hardware reciprocal and square root (~1 ulp): the IEEE-rounded forms
    // cost ~10 dependent instructions each on a latency-bound lane
    const float c2 = __builtin_amdgcn_sqrtf((1.0f - c1 * c1) * (float)KCAL_TO_AMU_A2_PS2 * (float)KB * T *
                                            __builtin_amdgcn_rcpf(m));
    const double before = speed2(v);
    for (int c = 0; c < 3; ++c) v[c] = c1 * v[c] + c2 * xi[c];
    return ke_change(m, before, speed2(v));
}
__device__ inline double ou_atom(float v[3], float m, int a, const SysDesc &D, float c1, float T,
                                 unsigned long long step, int purpose, const float *noise) {
    float xi[3];
    ou_noise(xi, a, D, step, purpose, noise);
    return ou_apply(v, m, c1, T, xi);
}

// Half kick of one atom with the 1000 kcal/mol/A clamp of
// prism_asc_used_force (prism_asc_decision.cuh:100-104) when gamma != 0,
// exact when gamma == 0 (nhs_amber_fused.cu:1750-1805, :3667-3703).  Returns
// true when the clamp acted, with its work -- used minus unclamped candidate
// velocity (apply_force_half_kick_with_intervention_work, :1750-1788) -- in dw.
__device__ inline bool kick_atom(float v[3], const long long fr[3], float m, float dt, float gamma, double &dw) {
    float f[3];
    for (int c = 0; c < 3; ++c) f[c] = from_fixed(fr[c], FORCE_FIXED_BITS);
    const float kick = 0.5f * dt * (float)KCAL_TO_AMU_A2_PS2 * __builtin_amdgcn_rcpf(m);
    if (gamma != 0.0f) {  // |f| > MAX_FORCE tested squared; the rescale is rare
        const float mag2 = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
        if (mag2 > MAX_FORCE * MAX_FORCE) {
            float raw[3];
            for (int c = 0; c < 3; ++c) raw[c] = v[c] + kick * f[c];
            const float q = MAX_FORCE * __builtin_amdgcn_rsqf(mag2);
            for (int c = 0; c < 3; ++c) f[c] *= q;
            for (int c = 0; c < 3; ++c) v[c] += kick * f[c];
            dw = ke_change(m, speed2(raw), speed2(v));
            return true;
        }
    }
    for (int c = 0; c < 3; ++c) v[c] += kick * f[c];
    return false;
}

// nhs_velocity_safety_clamp (nhs_amber_fused.cu:11148-11186) of one atom:
// |v| > max(20 sqrt(418.4 kB T / max(m, 1)), 2), tested squared.  Returns
// true when it acted, with the kinetic-energy change in dw.
__device__ inline bool vclamp_atom(float v[3], float m, float T, double &dw) {
    const float vmax2 = fmaxf(400.0f * (float)KCAL_TO_AMU_A2_PS2 * (float)KB * T *
                                  __builtin_amdgcn_rcpf(fmaxf(m, 1.0f)), 4.0f);
    const float sp2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (sp2 > vmax2) {
        const double before = speed2(v);
        const float sc = __builtin_amdgcn_sqrtf(vmax2) * __builtin_amdgcn_rsqf(sp2);
        for (int c = 0; c < 3; ++c) v[c] *= sc;
        dw = ke_change(m, before, speed2(v));
        return true;
    }
    return false;
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
__device__ inline int rattle_booked(GroupState &g, const GroupParam &gp, const RattleFactor &F, double tol) {
    if (g.n < 2) return 0;
    const double before = group_ke2(g, gp);
    const int r = rattle_apply(g, gp, F, tol);
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
// Sweep statistics: a lane whose count exceeds the running maximum `prev`
// (read at kernel start) of its group wave's slot raises it with an atomic
// that nothing waits for -- rare once the counts have settled, and cheaper on
// the lane's dependent chain than a reduction over the wave first.
__device__ inline void publish_sweeps(int *slot, int prev, int most) {
    if (most > prev) atomicMax(slot, most);
}

// A coordinate beyond 1000 A means the system has diverged (heartbeat_check,
// housekeeping.cu:27-45); a non-finite one is a fault of its own.
__device__ inline void check_atom(SysState &S, const float x[3], const float v[3], int a, unsigned long long step) {
    if (!(finite3(x) && finite3(v))) record_fault(S, FAULT_NONFINITE, a, step);
    else if (fmaxf(fmaxf(fabsf(x[0]), fabsf(x[1])), fabsf(x[2])) > MAX_COORD) record_fault(S, FAULT_DIVERGED, a, step);
}

// ---- group layouts
//
// Layout 1 (a lane per group): a lane holds a whole group, up to 4 atoms, and
// runs every phase for its atoms in turn.
// Layout 4 (a quad of lanes per group): lane k of the quad holds atom k.  The
// per-atom phases (the Ornstein-Uhlenbeck update with its Philox draws, the
// kick, the clamps, the loads and stores) run once per lane instead of four
// times in a row; every lane of the quad also holds a copy of the group's
// positions and velocities, exchanged with quad broadcasts (DPP), on which it
// runs the cluster's RATTLE/SHAKE -- the same FP64 operations on the same
// inputs as layout 1.  The work ledger gathers the lanes' contributions in
// atom order, and the centre-of-mass partials are summed in one tree in both
// layouts, so the two give bitwise the same positions, velocities and work.
// A small arena -- fewer group waves than the GPU has SIMDs -- leaves most of
// the chip idle and is bound by the length of one lane's dependent chain,
// which layout 4 shortens; a large arena is bound by throughput, where
// layout 1 does no work four times over.

// The value `v` of lane J of this lane's quad (DPP quad_perm [J,J,J,J]);
// every lane of the quad must be active.
template <int J>
__device__ inline int quad_lane(int v) {
    return __builtin_amdgcn_update_dpp(0, v, J * 0x55, 0xF, 0xF, false);
}
template <int J>
__device__ inline float quad_lane(float v) {
    return __int_as_float(quad_lane<J>(__float_as_int(v)));
}
template <int J>
__device__ inline double quad_lane(double v) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)quad_lane<J>((int)(unsigned)(b & 0xffffffffll));
    const unsigned hi = (unsigned)quad_lane<J>((int)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The quad's positions or velocities into every lane's group copy.
__device__ inline void quad_gather_x(GroupState &g, const float x[3]) {
#define PRISM_QG(J) \
    for (int c = 0; c < 3; ++c) g.x[J][c] = quad_lane<J>(x[c]);
    PRISM_QG(0) PRISM_QG(1) PRISM_QG(2) PRISM_QG(3)
#undef PRISM_QG
}
__device__ inline void quad_gather_v(GroupState &g, const float v[3]) {
#define PRISM_QG(J) \
    for (int c = 0; c < 3; ++c) g.v[J][c] = quad_lane<J>(v[c]);
    PRISM_QG(0) PRISM_QG(1) PRISM_QG(2) PRISM_QG(3)
#undef PRISM_QG
}
// Atom k's row of a group array, selected among values: handed the array
// itself, the compiler folds the selects into one dynamically indexed load,
// which moves the whole group state to scratch memory.
__device__ inline float pick4(int k, float a, float b, float c, float d) {
    return k == 0 ? a : (k == 1 ? b : (k == 2 ? c : d));
}
#define PRISM_OWN_ROW(src, k, out) \
    for (int c_ = 0; c_ < 3; ++c_) (out)[c_] = pick4((k), (src)[0][c_], (src)[1][c_], (src)[2][c_], (src)[3][c_])
// w += the contributions `c` of the quad's lanes whose `has` is set, in lane
// (atom) order: the order in which layout 1 adds its atoms' contributions.
// Every lane of the quad ends with the same w.
__device__ inline void quad_book(double &w, double c, bool has) {
    const int h = has ? 1 : 0;
#define PRISM_QB(J)                        \
    {                                      \
        const double cj = quad_lane<J>(c); \
        if (quad_lane<J>(h)) w += cj;      \
    }
    PRISM_QB(0) PRISM_QB(1) PRISM_QB(2) PRISM_QB(3)
#undef PRISM_QB
}

// Mass-weighted velocity sum of a group over its atoms in order.
__device__ inline void group_momentum(const GroupState &g, double mv[3]) {
    FOR_ATOMS(k) {
        for (int c = 0; c < 3; ++c) mv[c] += (double)g.m[k] * g.v[k][c];
    }
}
// Sum of the groups of a 16-group sub-block of a group wave: a butterfly over
// the groups' lanes at distances `first`, 2 first, 4 first, 8 first (groups 1,
// 2, 4, 8 apart in both layouts), so both layouts form the same tree.
__device__ inline double sum16(double v, int first) {
    for (int o = first; o < 16 * first; o <<= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ inline void freeze_atom(const DevPtrs &P, int a) {
    if (a >= 0)
        for (int c = 0; c < 3; ++c) P.f[(size_t)a * 4 + c] = 0;
}

// ---- first half: COM removal, OU_HALF_N, RATTLE_AFTER_OU_N (booked),
// KICK_N, RATTLE_PRE_DRIFT, DRIFT_SAVE_XN, SHAKE_POSITION, then the
// divergence checks, the stores and the displacement trigger of the pair list.

__global__ __launch_bounds__(64 * IW) void k_first_half1(DevPtrs P, const float *noise) {
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
        for (int k = 0; k < 4; ++k) freeze_atom(P, G.atom[k]);
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
    if (load_group<false>(P, gw * 64 + lane, g, gp, tail)) {
        if (com) {
            FOR_ATOMS(k) {
                for (int c = 0; c < 3; ++c) g.v[k][c] -= (float)S.com_v[c];
            }
        }
        if (gamma != 0.0f) {
            const float c1 = expf(-gamma * (0.5f * D.dt));
            FOR_ATOMS(k) g.w += ou_atom(g.v[k], g.m[k], g.a[k], D, c1, T, step, 1, noise);
        }
        RattleFactor F;
        rattle_factor(g, gp, F);
        note_sweeps(S, rattle_booked(g, gp, F, D.rattle_tol), FAULT_RATTLE_1, g.a[0], step, most_ra);
        FOR_ATOMS(k) {
            freeze_atom(P, g.a[k]);  // the step's force evaluation adds into the zeroed accumulator
            double dw;
            if (kick_atom(g.v[k], g.f[k], g.m[k], D.dt, gamma, dw)) g.w += dw;
        }
        note_sweeps(S, rattle_apply(g, gp, F, D.rattle_tol), FAULT_RATTLE_2, g.a[0], step, most_ra);
        float xs[4][3];
        FOR_ATOMS(k) {
            for (int c = 0; c < 3; ++c) {
                xs[k][c] = g.x[k][c];
                g.x[k][c] += D.dt * g.v[k][c];
            }
        }
        note_sweeps(S, shake(g, xs, gp, D.dt, D.shake_tol, D.constraint_max_iter), FAULT_SHAKE, g.a[0], step, most_sh);
        // each atom's sorted slot and build position, loaded only now: held
        // through the constraint solves they would cost the kernel a wave
        // per SIMD (188 VGPRs, two waves, instead of 166, three)
        FOR_ATOMS(k) {
            tail.slot[k] = P.slot_of[g.a[k]];
            const float4 bb = P.x_build[g.a[k]];
            tail.xb[k][0] = bb.x; tail.xb[k][1] = bb.y; tail.xb[k][2] = bb.z;
        }
        FOR_ATOMS(k) {
            const int a = g.a[k];
            check_atom(S, g.x[k], g.v[k], a, step);
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

// Layout 4: kernel wave kw covers groups 16 (kw & 3) .. 16 (kw & 3) + 15 of
// group wave kw >> 2; a workgroup is one group wave.  The phases are
// functions of what one lane holds (QuadAtom) so that the separate halves
// and the fused step seam (k_seam4) run the same code.

// One lane's atom as both halves read it: the group record and parameters,
// the atom's state and carried force, its sorted slot and build position, the
// group's cumulative work (quad-uniform) and the group wave's sweep maxima.
struct QuadAtom {
    int a;  // this lane's atom, -1 if the group has none here
    GroupParam gp;
    float m;
    float x[3], v[3];
    long long fr[3];
    int slot;
    float xb[3];
    double w;
    int prev_sh, prev_ra;
};

// Every load of a lane in one round after the group record: an absent atom
// reads arena atom 0, whose values are never used.
template <bool TAIL>
__device__ inline void load_quad(const DevPtrs &P, int gw, int gidx, int k, QuadAtom &q) {
    q.a = P.groups[gidx].atom[k];
    q.gp = P.gparams[gidx];
    const size_t ac = (size_t)max(q.a, 0);
    q.m = P.mass[ac];
    const float4 xa = P.x[ac], va = P.v[ac];
    q.x[0] = xa.x; q.x[1] = xa.y; q.x[2] = xa.z;
    q.v[0] = va.x; q.v[1] = va.y; q.v[2] = va.z;
    for (int c = 0; c < 3; ++c) q.fr[c] = P.f[ac * 4 + c];
    if (TAIL) {
        q.slot = P.slot_of[ac];
        const float4 b = P.x_build[ac];
        q.xb[0] = b.x; q.xb[1] = b.y; q.xb[2] = b.z;
    }
    q.w = P.work_g[gidx];
    q.prev_sh = P.sweeps[2 * gw];
    q.prev_ra = P.sweeps[2 * gw + 1];
}

// The fields of a system's step state that the first half reads.
struct StepView {
    unsigned long long step;
    long long rebuild_step;
    int halted;
    double com_v[3];
};
__device__ inline StepView step_view(const SysState &S) {
    StepView V;
    V.step = S.step;
    V.rebuild_step = S.rebuild_step;
    V.halted = S.halted;
    for (int c = 0; c < 3; ++c) V.com_v[c] = S.com_v[c];
    return V;
}

// First half of one lane's atom (the step's opening phases, see above).
__device__ inline void first_half4(const DevPtrs &P, const float *noise, int gw, int gidx, int lane, const SysDesc &D,
                                   SysState &S, const StepView &V, const QuadAtom &q) {
    const int k = lane & 3, a = q.a;
    if (V.halted || (D.end_step != 0 && V.step >= D.end_step)) {  // halted or parked (see parked())
        freeze_atom(P, a);
        return;
    }
    const unsigned long long step = V.step;
    float T, gamma;
    schedule(D, step + 1, T, gamma);
    const bool com = D.com_every > 0 && step > 0 && step % (unsigned long long)D.com_every == 0;
    GroupState g;
    g.w = q.w;
    bool moved = false;
    int most_sh = 0, most_ra = 0;
    if (quad_lane<0>(a) >= 0) {  // quad-uniform: not a padding group
        g.n = 1 + q.gp.n_h;
        const bool mine = k < g.n;
        float x[3] = {q.x[0], q.x[1], q.x[2]}, v[3] = {q.v[0], q.v[1], q.v[2]};
        float xi[3] = {0.0f, 0.0f, 0.0f};
        if (gamma != 0.0f) ou_noise(xi, a, D, step, 1, noise);
        quad_gather_x(g, x);
        RattleFactor F;  // both projections below see these positions
        rattle_factor(g, q.gp, F);
        if (com)
            for (int c = 0; c < 3; ++c) v[c] -= (float)V.com_v[c];
        if (gamma != 0.0f) {
            const float c1 = expf(-gamma * (0.5f * D.dt));
            double dw = 0.0;
            if (mine) dw = ou_apply(v, q.m, c1, T, xi);
            quad_book(g.w, dw, mine);
        }
        quad_gather_v(g, v);
        int r = rattle_booked(g, q.gp, F, D.rattle_tol);
        if (k == 0) note_sweeps(S, r, FAULT_RATTLE_1, a, step, most_ra);
        PRISM_OWN_ROW(g.v, k, v);
        {
            double dw = 0.0;
            bool hit = false;
            if (mine) {
                freeze_atom(P, a);
                hit = kick_atom(v, q.fr, q.m, D.dt, gamma, dw);
            }
            quad_book(g.w, dw, hit);
        }
        quad_gather_v(g, v);
        r = rattle_apply(g, q.gp, F, D.rattle_tol);
        if (k == 0) note_sweeps(S, r, FAULT_RATTLE_2, a, step, most_ra);
        float xs[4][3];
        FOR_ATOMS(j) {
            for (int c = 0; c < 3; ++c) {
                xs[j][c] = g.x[j][c];
                g.x[j][c] += D.dt * g.v[j][c];
            }
        }
        r = shake(g, xs, q.gp, D.dt, D.shake_tol, D.constraint_max_iter);
        if (k == 0) note_sweeps(S, r, FAULT_SHAKE, a, step, most_sh);
        PRISM_OWN_ROW(g.x, k, x);
        PRISM_OWN_ROW(g.v, k, v);
        if (mine) {
            check_atom(S, x, v, a, step);
            P.x[a] = make_float4(x[0], x[1], x[2], 0.0f);
            P.v[a] = make_float4(v[0], v[1], v[2], 0.0f);
            float *xq = (float *)&P.xq[q.slot];
            xq[0] = x[0]; xq[1] = x[1]; xq[2] = x[2];
            const float dx = x[0] - q.xb[0], dy = x[1] - q.xb[1], dz = x[2] - q.xb[2];
            moved = dx * dx + dy * dy + dz * dz > D.half_skin2;
        }
        if (k == 0) P.work_g[gidx] = g.w;
    }
    publish_sweeps(P.sweeps + 2 * gw, q.prev_sh, most_sh);
    publish_sweeps(P.sweeps + 2 * gw + 1, q.prev_ra, most_ra);
    const bool due = (long long)step - V.rebuild_step >= (long long)D.rebuild_every;
    if (__ballot(moved) != 0ull || due) {
        if (lane == 0) {
            S.rebuild_flag = 1;
            P.counters->any_rebuild = 1;
        }
    }
}

__global__ __launch_bounds__(64 * IW) void k_first_half4(DevPtrs P, const float *noise) {
    const int lane = threadIdx.x & 63;
    const int kw = blockIdx.x * IW + wave_id();
    const int gw = kw >> 2;
    if (gw >= P.n_waves_arena) return;
    const int s = P.wave_sys[gw];
    if (s < 0) return;
    const SysDesc &D = P.sys[s];
    if (!D.active) return;
    const int gidx = gw * 64 + (kw & 3) * 16 + (lane >> 2);
    QuadAtom q;
    load_quad<true>(P, gw, gidx, lane & 3, q);
    first_half4(P, noise, gw, gidx, lane, D, P.sst[s], step_view(P.sst[s]), q);
}

// ---- second half: KICK_NP1, RATTLE_BEFORE_OU_NP1, OU_HALF_NP1, the velocity
// safety clamp and the final RATTLE (booked); then the step-end handshake.

// Last-block handshake (no waiting): the workgroup that retires last sums the
// COM partials in a fixed order and advances every system's step counter.  A
// workgroup releases its writes (agent-scope fence) only when the last one
// reads any of them -- COM partials, a latched fault: the fence writes back
// this XCD's L2 and costs microseconds per workgroup, so on the other steps
// the counter increment goes out unfenced.  The count has two levels,
// HS_GROUP workgroups per group counter and one increment per group on the
// global one: atomics on one address serialise (~12 ns each), which a flat
// count over the thousands of workgroups of a large arena would pay on every
// step.  A workgroup that published adds HS_PUBLISHED, so the last of its
// group knows to pass the writes on.
// The count: in every thread of the workgroup, 0 when another workgroup
// retires last; in the last one 1, or 3 when some workgroup of the launch
// published.  `after` is a value the caller loaded that must be in hand
// before this workgroup is counted.
__device__ inline int step_end_arrive(const DevPtrs &P, bool publish, unsigned after = 0u) {
    __shared__ int last;
    const int pub = __syncthreads_or(publish);
    if (pub) __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        asm volatile("" ::"v"(after) : "memory");
        const unsigned g = blockIdx.x / HS_GROUP;
        const unsigned gsize = min((unsigned)HS_GROUP, gridDim.x - g * HS_GROUP);
        const unsigned n_groups = (gridDim.x + HS_GROUP - 1) / HS_GROUP;
        const unsigned old = atomicAdd(&P.hs_group[g], pub ? HS_PUBLISHED + 1u : 1u);
        int is_last = 0;
        if ((old & (HS_PUBLISHED - 1u)) == gsize - 1) {  // the group's last workgroup
            P.hs_group[g] = 0u;  // for the next launch; no other workgroup touches it now
            const bool gpub = pub || old >= HS_PUBLISHED;
            if (n_groups == 1) {  // a small arena: one level
                is_last = gpub ? 3 : 1;
            } else {
                if (gpub) __threadfence();
                const unsigned o2 = atomicAdd(&P.counters->done_groups, gpub ? HS_PUBLISHED + 1u : 1u);
                if ((o2 & (HS_PUBLISHED - 1u)) == n_groups - 1) is_last = (gpub || o2 >= HS_PUBLISHED) ? 3 : 1;
            }
        }
        last = is_last;
    }
    __syncthreads();
    return last;
}

// The last workgroup's bookkeeping: COM sums, rebuild flags, step counters,
// halting, the step-end stamp.
__device__ inline void step_end_book(const DevPtrs &P) {
    __threadfence();
    for (int q = threadIdx.x; q < P.n_sys; q += blockDim.x) {
        const SysDesc &D = P.sys[q];
        if (!D.active) continue;
        SysState &S = P.sst[q];
        if (S.halted || parked(D, S)) continue;
        const unsigned long long step = S.step;
        if (D.com_every > 0 && (step + 1) % (unsigned long long)D.com_every == 0) {
            // each group wave's four 16-group sums, (s0 + s1) + (s2 + s3),
            // then the group waves in order
            double acc[3] = {0.0, 0.0, 0.0};
            for (int w = D.wave_begin; w < D.wave_begin + D.n_waves; ++w) {
                const double *p = P.com_partial + (size_t)w * 12;
                for (int c = 0; c < 3; ++c) acc[c] += (p[c] + p[3 + c]) + (p[6 + c] + p[9 + c]);
            }
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
        P.counters->nb_arrive = 0u;  // the nonbonded rows barrier, for the next rebuild step
        P.counters->nb_go = 0u;
        P.counters->any_rebuild = 0;
        // the step's end on the device's constant-rate clock: the host turns
        // consecutive stamps into per-step latencies that include every gap
        // between steps (host work, admissions, other launches)
        const unsigned long long k = P.counters->step_ends;
        P.step_ts[k % TS_RING] = (unsigned long long)wall_clock64();
        P.counters->step_ends = k + 1;
    }
}

__device__ inline void step_end(const DevPtrs &P, bool publish) {
    if (step_end_arrive(P, publish)) step_end_book(P);
}

__global__ __launch_bounds__(64 * IW) void k_second_half1(DevPtrs P, const float *noise) {
    const int lane = threadIdx.x & 63;
    const int gw = blockIdx.x * IW + wave_id();
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
            FOR_ATOMS(k) {
                double dw;
                if (kick_atom(g.v[k], g.f[k], g.m[k], D.dt, gamma, dw)) g.w += dw;
            }
            RattleFactor F;
            rattle_factor(g, gp, F);
            int sw = rattle_apply(g, gp, F, D.rattle_tol);
            publish |= sw < 0;
            note_sweeps(S, sw, FAULT_RATTLE_3, g.a[0], step, most_ra);
            if (gamma != 0.0f) {
                const float c1 = expf(-gamma * (0.5f * D.dt));
                FOR_ATOMS(k) g.w += ou_atom(g.v[k], g.m[k], g.a[k], D, c1, T, step, 2, noise);
            }
            FOR_ATOMS(k) {
                double dw;
                if (vclamp_atom(g.v[k], g.m[k], T, dw)) g.w += dw;
            }
            sw = rattle_booked(g, gp, F, D.rattle_tol);
            publish |= sw < 0;
            note_sweeps(S, sw, FAULT_RATTLE_4, g.a[0], step, most_ra);
            FOR_ATOMS(k) {
                if (!finite3(g.v[k])) {
                    record_fault(S, FAULT_NONFINITE, g.a[k], step);
                    publish = true;
                }
                P.v[g.a[k]] = make_float4(g.v[k][0], g.v[k][1], g.v[k][2], 0.0f);
            }
            group_momentum(g, mv);
            P.work_g[gw * 64 + lane] = g.w;
        }
        publish_sweeps(P.sweeps + 2 * gw + 1, prev_ra, most_ra);
        if (next_com) {
            for (int c = 0; c < 3; ++c) mv[c] = sum16(mv[c], 1);
            if ((lane & 15) == 0) {
                for (int c = 0; c < 3; ++c) P.com_partial[((size_t)gw * 4 + (lane >> 4)) * 3 + c] = mv[c];
                publish = true;
            }
        }
    }
    step_end(P, publish);
}

// Second half of one lane's atom (the step's closing phases, see above); q.v
// and q.w leave holding what was stored.  Returns whether the lane wrote
// something the step end reads (COM partials, a fault).
__device__ inline bool second_half4(const DevPtrs &P, const float *noise, int gw, int gidx, int kw, int lane,
                                    const SysDesc &D, SysState &S, QuadAtom &q) {
    const int k = lane & 3, a = q.a;
    bool publish = false;
    const unsigned long long step = S.step;
    float T, gamma;
    schedule(D, step + 1, T, gamma);
    const bool next_com = D.com_every > 0 && (step + 1) % (unsigned long long)D.com_every == 0;
    GroupState g;
    g.w = q.w;
    double mv[3] = {0.0, 0.0, 0.0};
    int most_ra = 0;
    if (quad_lane<0>(a) >= 0) {
        g.n = 1 + q.gp.n_h;
        const bool mine = k < g.n;
        float x[3] = {q.x[0], q.x[1], q.x[2]}, v[3] = {q.v[0], q.v[1], q.v[2]};
        float xi[3] = {0.0f, 0.0f, 0.0f};
        if (gamma != 0.0f) ou_noise(xi, a, D, step, 2, noise);
        quad_gather_x(g, x);
        RattleFactor F;  // both projections below see these positions
        rattle_factor(g, q.gp, F);
        {
            double dw = 0.0;
            bool hit = false;
            if (mine) hit = kick_atom(v, q.fr, q.m, D.dt, gamma, dw);
            quad_book(g.w, dw, hit);
        }
        quad_gather_v(g, v);
        int sw = rattle_apply(g, q.gp, F, D.rattle_tol);
        publish |= sw < 0;
        if (k == 0) note_sweeps(S, sw, FAULT_RATTLE_3, a, step, most_ra);
        PRISM_OWN_ROW(g.v, k, v);
        double d_ou = 0.0, d_vc = 0.0;
        bool h_vc = false;
        if (gamma != 0.0f) {
            const float c1 = expf(-gamma * (0.5f * D.dt));
            if (mine) d_ou = ou_apply(v, q.m, c1, T, xi);
            quad_book(g.w, d_ou, mine);
        }
        if (mine) h_vc = vclamp_atom(v, q.m, T, d_vc);
        quad_book(g.w, d_vc, h_vc);
        quad_gather_v(g, v);
        sw = rattle_booked(g, q.gp, F, D.rattle_tol);
        publish |= sw < 0;
        if (k == 0) note_sweeps(S, sw, FAULT_RATTLE_4, a, step, most_ra);
        PRISM_OWN_ROW(g.v, k, v);
        if (mine) {
            if (!finite3(v)) {
                record_fault(S, FAULT_NONFINITE, a, step);
                publish = true;
            }
            P.v[a] = make_float4(v[0], v[1], v[2], 0.0f);
        }
        for (int c = 0; c < 3; ++c) q.v[c] = v[c];
        g.m[0] = quad_lane<0>(q.m); g.m[1] = quad_lane<1>(q.m); g.m[2] = quad_lane<2>(q.m); g.m[3] = quad_lane<3>(q.m);
        group_momentum(g, mv);
        if (k == 0) P.work_g[gidx] = g.w;
        q.w = g.w;
    }
    publish_sweeps(P.sweeps + 2 * gw + 1, q.prev_ra, most_ra);
    if (next_com) {
        for (int c = 0; c < 3; ++c) mv[c] = sum16(mv[c], 4);
        if (lane == 0) {
            for (int c = 0; c < 3; ++c) P.com_partial[((size_t)gw * 4 + (kw & 3)) * 3 + c] = mv[c];
            publish = true;
        }
    }
    return publish;
}

__global__ __launch_bounds__(64 * IW) void k_second_half4(DevPtrs P, const float *noise) {
    const int lane = threadIdx.x & 63;
    const int kw = blockIdx.x * IW + wave_id();
    const int gw = kw >> 2;
    bool publish = false;
    const int s = gw < P.n_waves_arena ? P.wave_sys[gw] : -1;
    if (s >= 0 && P.sys[s].active && !P.sst[s].halted && !parked(P.sys[s], P.sst[s])) {
        const int gidx = gw * 64 + (kw & 3) * 16 + (lane >> 2);
        QuadAtom q;
        load_quad<false>(P, gw, gidx, lane & 3, q);
        publish = second_half4(P, noise, gw, gidx, kw, lane, P.sys[s], P.sst[s], q);
    }
    step_end(P, publish);
}

// ---- the step seam: the second half of step n and the first half of step
// n + 1 in one launch (no explicit noise), joined by a grid barrier at the
// point where the separate kernels would end and start.  The lanes keep their
// atoms' records, positions, forces and velocities in registers across the
// barrier: one round of loads and one launch per step fewer, bitwise the same
// trajectory.
//
// The last workgroup to arrive bumps counters->seam_gen, which the others
// wait for.  On a step on which some workgroup published (COM partials, a
// fault latched in this launch) it first does the whole step-end bookkeeping
// (step_end_book) and writes it back, and the bump says so (odd): the others
// then read their system's new step state from memory at agent scope.  On the
// other steps the new state follows from the state at kernel start
// (view_after), so every workgroup derives it itself: the last one clears the
// consumed rebuild flags with atomics (which complete before the bump, so no
// first half's new flag can be overwritten), bumps, and only then writes the
// counters from the values it read before the bump.
//
// The barrier needs every workgroup resident at once; the host launches the
// seam only when the kernel's occupancy times the compute units covers the
// grid (seam_fits).  The wait is bounded all the same: a workgroup that has
// waited SEAM_WAIT_TICKS of the constant-rate clock latches ST_BARRIER on its
// system and skips the first half, so the launch always ends.
constexpr unsigned long long SEAM_WAIT_TICKS = 1ull << 26;  // ~0.67 s at 100 MHz

__device__ inline unsigned long long agent_load(const void *p) {
    return __hip_atomic_load((unsigned long long *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A system's step-end fields at kernel start.
struct StartState {
    unsigned long long step;
    long long rebuild_step;
    int rebuild_flag, halted, status, n_rebuilds;
};
__device__ inline StartState start_state(const SysState &S) {
    return StartState{S.step, S.rebuild_step, S.rebuild_flag, S.halted, S.status, S.n_rebuilds};
}
// What step_end_book leaves a system with when nobody published: no COM sum
// (only COM steps publish) and no fault beyond those latched before the
// launch.
__device__ inline StepView view_after(const SysDesc &D, const StartState &s0) {
    StepView V;
    V.step = s0.step;
    V.rebuild_step = s0.rebuild_step;
    V.halted = s0.halted;
    for (int c = 0; c < 3; ++c) V.com_v[c] = 0.0;  // read on COM steps only, which publish
    if (D.active && !s0.halted && !(D.end_step != 0 && s0.step >= D.end_step)) {
        if (s0.rebuild_flag) V.rebuild_step = (long long)s0.step;
        V.step = s0.step + 1;
        if (s0.status & ST_HALT) V.halted = 1;
    }
    return V;
}

// The last workgroup's bookkeeping on a step nobody published: the flags
// cleared before the bump, everything else after it from the values read
// before it.  Returns after the bump.
__device__ inline void seam_book_quiet(const DevPtrs &P, unsigned gen0) {
    const int q = threadIdx.x;
    const bool mine = q < P.n_sys;
    StartState s0{};
    bool adv = false;
    int cleared = 0;
    if (mine) {
        const SysDesc &D = P.sys[q];
        SysState &S = P.sst[q];
        s0 = start_state(S);
        adv = D.active && !s0.halted && !(D.end_step != 0 && s0.step >= D.end_step);
        if (adv && s0.rebuild_flag) cleared = 1 | atomicExch(&S.rebuild_flag, 0) | atomicExch(&S.resort_flag, 0);
    }
    if (threadIdx.x == 0 && P.counters->any_rebuild) cleared |= 1 | atomicExch(&P.counters->any_rebuild, 0);
    // every clear complete (memory counter zero) before the bump
    if (cleared) __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0)
        __hip_atomic_store(&P.counters->seam_gen, (gen0 & ~1u) + 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (mine && adv) {
        SysState &S = P.sst[q];
        if (s0.rebuild_flag) {
            S.rebuild_step = (long long)s0.step;
            S.n_rebuilds = s0.n_rebuilds + 1;
        }
        S.step = s0.step + 1;
        if (s0.status & ST_HALT) S.halted = 1;
    }
    if (threadIdx.x == 0) {
        P.counters->done_groups = 0u;
        P.counters->any_dirty = 0;
        P.counters->nb_arrive = 0u;  // the nonbonded rows barrier, for the next rebuild step
        P.counters->nb_go = 0u;
        const unsigned long long k = P.counters->step_ends;
        P.step_ts[k % TS_RING] = (unsigned long long)wall_clock64();
        P.counters->step_ends = k + 1;
    }
}

// 0: the wait timed out; 1: the new step state follows from the state at
// kernel start (view_after); 2: read it from memory.
__device__ inline int seam_barrier(const DevPtrs &P, bool publish, unsigned gen0, int s) {
    __shared__ int go;
    const int arrived = step_end_arrive(P, publish, gen0);
    if (arrived == 3) {
        step_end_book(P);
        // the bookkeeping's writers (one thread per system, thread 0 for
        // the counters) write it back before the bump
        if ((int)threadIdx.x < P.n_sys || threadIdx.x == 0) __threadfence();
        __syncthreads();
        if (threadIdx.x == 0) {
            __hip_atomic_store(&P.counters->seam_gen, (gen0 & ~1u) + 3u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            go = 2;
        }
    } else if (arrived == 1) {
        seam_book_quiet(P, gen0);
        if (threadIdx.x == 0) go = 1;
    } else if (threadIdx.x == 0) {
        const unsigned long long t0 = wall_clock64();
        unsigned gen = gen0;
        int ok = 1;
        while ((gen = __hip_atomic_load(&P.counters->seam_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == gen0) {
            if (wall_clock64() - t0 > SEAM_WAIT_TICKS) {
                ok = 0;
                atomicOr(&P.sst[max(s, 0)].status, ST_BARRIER);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        go = ok ? 1 + (int)(gen & 1u) : 0;
    }
    __syncthreads();
    return go;
}

__global__ __launch_bounds__(64 * IW) void k_seam4(DevPtrs P) {
    __shared__ StepView view;
    const int lane = threadIdx.x & 63;
    const int kw = blockIdx.x * IW + wave_id();
    const int gw = kw >> 2;  // the workgroup's group wave
    // the generation before this launch's bump: the arrival waits for it
    // (step_end_arrive's `after`)
    unsigned gen0 = 0u;
    if (threadIdx.x == 0) gen0 = __hip_atomic_load(&P.counters->seam_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int s = gw < P.n_waves_arena ? P.wave_sys[gw] : -1;
    const bool live = s >= 0 && P.sys[s].active;
    StartState s0{};
    if (live && threadIdx.x == 0) s0 = start_state(P.sst[s]);
    const int gidx = gw * 64 + (kw & 3) * 16 + (lane >> 2);
    QuadAtom q;
    bool publish = false;
    if (live) {
        load_quad<true>(P, gw, gidx, lane & 3, q);
        if (!P.sst[s].halted && !parked(P.sys[s], P.sst[s]))
            publish = second_half4(P, nullptr, gw, gidx, kw, lane, P.sys[s], P.sst[s], q);
    }
    const int go = seam_barrier(P, publish, gen0, s);
    if (!go || !live) return;
    if (threadIdx.x == 0) {
        if (go == 1) {
            view = view_after(P.sys[s], s0);
        } else {
            const SysState &S = P.sst[s];
            view.step = agent_load(&S.step);
            view.rebuild_step = (long long)agent_load(&S.rebuild_step);
            view.halted = __hip_atomic_load((int *)&S.halted, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (int c = 0; c < 3; ++c) view.com_v[c] = __longlong_as_double((long long)agent_load(&S.com_v[c]));
        }
    }
    __syncthreads();
    const StepView V = view;
    first_half4(P, nullptr, gw, gidx, lane, P.sys[s], P.sst[s], V, q);
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

// The four-lane layout when the arena's group waves are few enough that the
// integrator is bound by one lane's chain (each group wave becomes four
// kernel waves: at this bound, up to 2,048 of them on the GPU's 1,024 SIMDs).
static bool quad_layout(const DevPtrs &P) { return P.n_waves_arena <= QUAD_LAYOUT_MAX_WAVES; }

void launch_first_half(const DevPtrs &P, const float *noise, hipStream_t st) {
    if (quad_layout(P))
        hipLaunchKernelGGL(k_first_half4, dim3(P.n_waves_arena), dim3(64 * IW), 0, st, P, noise);
    else
        hipLaunchKernelGGL(k_first_half1, dim3((P.n_waves_arena + IW - 1) / IW), dim3(64 * IW), 0, st, P, noise);
}

void launch_second_half(const DevPtrs &P, const float *noise, hipStream_t st) {
    if (quad_layout(P))
        hipLaunchKernelGGL(k_second_half4, dim3(P.n_waves_arena), dim3(64 * IW), 0, st, P, noise);
    else
        hipLaunchKernelGGL(k_second_half1, dim3((P.n_waves_arena + IW - 1) / IW), dim3(64 * IW), 0, st, P, noise);
}

bool seam_fits(const DevPtrs &P) {
    if (!quad_layout(P)) return false;
    int dev = 0, per_cu = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_seam4, 64 * IW, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return false;
    return (long long)per_cu * cus >= (long long)P.n_waves_arena;
}

void launch_seam(const DevPtrs &P, hipStream_t st) {
    hipLaunchKernelGGL(k_seam4, dim3(P.n_waves_arena), dim3(64 * IW), 0, st, P);
}

void launch_kinetic(const DevPtrs &P, double *out, int com_sys, hipStream_t st) {
    hipLaunchKernelGGL(k_sys_reduce, dim3(P.n_sys), dim3(256), 0, st, P, out, com_sys);
}

}  // namespace prism
