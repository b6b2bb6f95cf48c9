// bonded.hpp -- bonded terms and exact 1-4 exceptions of one force
// evaluation, one term per thread (compute_bond_force, compute_angle_force,
// compute_dihedral_force of nhs_amber_fused.cu:354-625 and the exception
// branch :3316-3470), shared by the stand-alone bonded kernel and the kernels
// that carry the bonded terms in their idle waves.
#pragma once
#include "device_common.hpp"

namespace prism {

__device__ inline float3 ld3(const float4 *x, int a) {
    const float4 p = x[a];
    return make_float3(p.x, p.y, p.z);
}
__device__ inline double3 ld3d(const float4 *x, int a) {
    const float4 p = x[a];
    return make_double3(p.x, p.y, p.z);
}
__device__ inline double3 sub3(double3 a, double3 b) { return make_double3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline double dot3(double3 a, double3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline double3 cross3(double3 a, double3 b) {
    return make_double3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// 1/sqrt(x) and 1/x in FP64 from the hardware estimates (v_rsq_f64,
// v_rcp_f64: about 2^-23 relative) and two Newton steps each (about 1e-16),
// a fraction of the instructions of an IEEE square root and division.
__device__ inline double rsqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    for (int k = 0; k < 2; ++k) {
        const double e = fma(-x * y, y, 1.0);  // 1 - x y^2
        y = fma(0.5 * y, e, y);
    }
    return y;
}

// Slot of a term's position q in the arena's contribution layout (TM:
// term-major, the layout is a kernel template parameter).
template <bool TM>
struct Slots {
    const int *atom_major;  // dst_* row of the term
    int term_first;         // the term's first slot in the term-major layout
    __device__ int operator[](int q) const { return TM ? term_first + q : atom_major[q]; }
};

__device__ inline void put(const DevPtrs &P, int slot, double x, double y, double z) {
    float *c = P.contrib + (size_t)slot * 3;
    c[0] = (float)x;
    c[1] = (float)y;
    c[2] = (float)z;
}

// One thread per term: bonds, angles, dihedrals, exceptions in that order.
// Every term writes the force on each of its atoms into that position's
// contribution slot (zeros when the term is masked out or degenerate); the
// gather adds an atom's slots as int64 fixed point, so there are no atomics.  The
// bonded terms and exceptions evaluate in FP64 from the FP32 state: on CDNA4
// the FP64 vector rate equals the unpacked FP32 rate, and it removes the
// cancellation of r - r0 and theta - theta0 near equilibrium that limits an
// FP32 evaluation to ~1e-5 relative.
template <bool ENERGY, bool TM>
__device__ inline void bonded_term(const DevPtrs &P, unsigned mask, int tid) {
    if (tid >= P.n_bonds + P.n_angles + P.n_dih + P.n_exc) return;
    // each branch loads its term record and the chunk's live flag together,
    // so a live term waits for memory once before its atoms' positions (the
    // record's first index, never below -1, enters the test so that the
    // compiler cannot move the record load after it)
    const unsigned char *live = P.term_live + tid / TERM_CHUNK;
    if (P.terms_sparse && !*live) return;  // a pool's free chunks skip their records
    int idx = tid;
    double energy = 0.0;
    if (idx < P.n_bonds) {
        const Bond B = P.bonds[idx];
        if ((*live == 0) | (B.i < -1)) return;
        const Slots<TM> dst{P.dst_bond + 2 * idx, 2 * idx};
        double fm = 0.0;
        double3 r = make_double3(0, 0, 0);
        if ((mask & PRISM_MD_TERM_BOND) && B.i >= 0) {
            r = sub3(ld3d(P.x, B.j), ld3d(P.x, B.i));
            const double r2 = dot3(r, r);
            if (r2 >= 1e-16) {  // |r| >= 1e-8
                const double inv_len = rsqrt_nr(r2);
                const double dr = r2 * inv_len - (double)B.r0;
                fm = -2.0 * (double)B.k * dr * inv_len;
                energy = (double)B.k * dr * dr;
            }
        }
        if (B.i >= 0) {
            put(P, dst[0], -fm * r.x, -fm * r.y, -fm * r.z);
            put(P, dst[1], fm * r.x, fm * r.y, fm * r.z);
        }
    } else if ((idx -= P.n_bonds) < P.n_angles) {
        const Angle A = P.angles[idx];
        if ((*live == 0) | (A.i < -1)) return;
        const Slots<TM> dst{P.dst_angle + 3 * idx, 2 * P.n_bonds + 3 * idx};
        double fm = 0.0;
        double3 di = make_double3(0, 0, 0), dk = make_double3(0, 0, 0);
        if ((mask & PRISM_MD_TERM_ANGLE) && A.i >= 0) {
            const double3 pj = ld3d(P.x, A.j);
            const double3 a = sub3(ld3d(P.x, A.i), pj), b = sub3(ld3d(P.x, A.k), pj);
            const double a2 = dot3(a, a), b2 = dot3(b, b);
            if (a2 >= 1e-16 && b2 >= 1e-16) {  // |a|, |b| >= 1e-8
                const double ia = rsqrt_nr(a2), ib = rsqrt_nr(b2);
                double cs = dot3(a, b) * ia * ib;
                cs = fmax(-1.0, fmin(1.0, cs));
                const double dth = acos(cs) - (double)A.theta0;
                const double s2 = 1.0 - cs * cs;  // sin^2 theta, sin theta held at >= 1e-8
                const double inv_sn = s2 < 1e-16 ? 1e8 : rsqrt_nr(s2);
                fm = -2.0 * (double)A.kf * dth * inv_sn;
                di = make_double3((b.x * ib - cs * a.x * ia) * ia, (b.y * ib - cs * a.y * ia) * ia, (b.z * ib - cs * a.z * ia) * ia);
                dk = make_double3((a.x * ia - cs * b.x * ib) * ib, (a.y * ia - cs * b.y * ib) * ib, (a.z * ia - cs * b.z * ib) * ib);
                energy = (double)A.kf * dth * dth;
            }
        }
        if (A.i >= 0) {
            put(P, dst[0], -fm * di.x, -fm * di.y, -fm * di.z);
            put(P, dst[1], fm * (di.x + dk.x), fm * (di.y + dk.y), fm * (di.z + dk.z));
            put(P, dst[2], -fm * dk.x, -fm * dk.y, -fm * dk.z);
        }
    } else if ((idx -= P.n_angles) < P.n_dih) {
        const Dihedral Q = P.dih[idx];
        if ((*live == 0) | (Q.i < -1)) return;
        const Slots<TM> dst{P.dst_dih + 4 * idx, 2 * P.n_bonds + 3 * P.n_angles + 4 * idx};
        double3 f1 = make_double3(0, 0, 0), f4 = make_double3(0, 0, 0), sv = make_double3(0, 0, 0);
        if ((mask & PRISM_MD_TERM_DIHEDRAL) && Q.i >= 0) {
            const double3 pj = ld3d(P.x, Q.j), pk = ld3d(P.x, Q.k);
            const double3 v0 = sub3(ld3d(P.x, Q.i), pj), v1 = sub3(pk, pj), v2 = sub3(pk, ld3d(P.x, Q.l));
            const double3 c0 = cross3(v0, v1), c1 = cross3(v1, v2);
            const double n1 = dot3(c0, c0), n2 = dot3(c1, c1), nb = dot3(v1, v1);
            if (n1 >= 1e-16 && n2 >= 1e-16 && nb >= 1e-16) {
                const double inv_lb = rsqrt_nr(nb), lb = nb * inv_lb, inv_nb = inv_lb * inv_lb;
                // c0 x c1 = v1 (v0 . c1), so phi = atan2(|v1| v0.c1, c0.c1): the
                // reference's acos / asin branches (nhs_amber_fused.cu:562-580) in
                // exact arithmetic, here as cos phi and sin phi over
                // R = |c0| |c1| = sqrt(n1 n2)
                const double inv_r = rsqrt_nr(n1 * n2), inv_r2 = inv_r * inv_r;
                const double cp = dot3(c0, c1) * inv_r, sp = lb * dot3(v0, c1) * inv_r;
                // cos(n phi), sin(n phi) by angle addition, then shifted by the
                // phase: no transcendental on the term's chain
                const int m = Q.n < 0 ? -Q.n : Q.n;
                double cn = 1.0, sn = 0.0;
                for (int t = 0; t < m; ++t) {
                    const double c = cn * cp - sn * sp;
                    sn = sn * cp + cn * sp;
                    cn = c;
                }
                if (Q.n < 0) sn = -sn;
                const double cd = Q.cos_phase, sd = Q.sin_phase;
                const double sa = sn * cd - cn * sd;  // sin(n phi - phase)
                const double ca = cn * cd + sn * sd;  // cos(n phi - phase)
                const double nn = (double)Q.n;
                const double dE = -(double)Q.kf * nn * sa;
                const double ff0 = -dE * lb * (n2 * inv_r2), ff3 = dE * lb * (n1 * inv_r2);  // / n1, / n2
                const double ff1 = dot3(v0, v1) * inv_nb, ff2 = dot3(v2, v1) * inv_nb;
                f1 = make_double3(ff0 * c0.x, ff0 * c0.y, ff0 * c0.z);
                f4 = make_double3(ff3 * c1.x, ff3 * c1.y, ff3 * c1.z);
                sv = make_double3(ff1 * f1.x - ff2 * f4.x, ff1 * f1.y - ff2 * f4.y, ff1 * f1.z - ff2 * f4.z);
                energy = (double)Q.kf * (1.0 + ca);
            }
        }
        if (Q.i >= 0) {
            put(P, dst[0], f1.x, f1.y, f1.z);
            put(P, dst[1], sv.x - f1.x, sv.y - f1.y, sv.z - f1.z);
            put(P, dst[2], -sv.x - f4.x, -sv.y - f4.y, -sv.z - f4.z);
            put(P, dst[3], f4.x, f4.y, f4.z);
        }
    } else if ((idx -= P.n_dih) < P.n_exc) {
        const Exception14 X = P.exc[idx];
        if ((*live == 0) | (X.i < -1)) return;
        const Slots<TM> dst{P.dst_exc + 2 * idx, 2 * P.n_bonds + 3 * P.n_angles + 4 * P.n_dih + 2 * idx};
        double f = 0.0;
        double3 d = make_double3(0, 0, 0);
        if ((mask & PRISM_MD_TERM_EXCEPTION) && X.i >= 0) {
            d = sub3(ld3d(P.x, X.i), ld3d(P.x, X.j));
            const double r2 = dot3(d, d);
            if (r2 < 1e-8) {  // coincident exception atoms (nhs_amber_fused.cu:3338-3343)
                atomicOr(&P.sst[P.atom_sys[X.i]].status, ST_EXCEPTION_COINCIDENT);
            } else {
                const double inv_r = rsqrt_nr(r2), inv_r2 = inv_r * inv_r;
                const double s2 = (double)X.sigma * (double)X.sigma * inv_r2, s6 = s2 * s2 * s2;
                const double qq = COULOMB * (double)X.qq;
                const double eps = X.eps;
                f = (24.0 * eps * s6 * (2.0 * s6 - 1.0) + qq * inv_r) * inv_r2;
                energy = 4.0 * eps * (s6 * s6 - s6) + qq * inv_r;
            }
        }
        if (X.i >= 0) {
            put(P, dst[0], f * d.x, f * d.y, f * d.z);
            put(P, dst[1], -f * d.x, -f * d.y, -f * d.z);
        }
    } else {
        return;
    }
    if (ENERGY) P.term_e[tid] = energy;
}

}  // namespace prism
