// oracle.cpp -- FP64 CPU restatement of the reference classical MD step.
//
// TEST INFRASTRUCTURE ONLY: the parity checker.  Loaded by tests/, the
// smoke check and the benchmark's CPU-baseline leg; never by product code.
//
// What pins each function (beyond being a term-by-term restatement); the
// checks live in tests/test_oracle_*.py:
//   oracle_forces  bond/angle/dihedral/nonbonded/exception: central finite
//                  differences of the oracle's own energies, closed forms
//                  (2-atom bond, angle at and off theta0, torsion at phi = 0,
//                  pi/2, pi, LJ minimum at 2^(1/6) sigma, Coulomb pair,
//                  V(rc) = 0 after the shift, exceptions without cutoff),
//                  sum F = 0 and sum r x F = 0 per term, and a planted sign
//                  error that the finite-difference check must reject.
//   oracle_pairs   an independent numpy brute force and cell scheme (float32,
//                  one rounding per operation), the known stencil defect of
//                  10 A cells (only 10-12 A pairs missed,
//                  scripts/physics_refs/neighbor_list.py:1-11), agreement of
//                  both modes on every pair within 10 A, and the reference's
//                  capacity errors.
//   oracle_run     closed forms: OU half-steps with explicit noise, the OU
//                  stationary variance 418.4 kB T/m, free-particle drift,
//                  velocity Verlet with the 0.5*dt*418.4 half-kick constant
//                  (fused_engine.rs:4723-4731), X-H lengths after SHAKE, zero
//                  bond-parallel relative velocity after RATTLE, momentum
//                  conservation of SHAKE, second-order NVE energy error.
//   oracle_run work the energy balance: E(t) - E(0) - W is a second-order
//                  shadow error while W is hundreds of kcal/mol, so a dropped,
//                  extra or sign-flipped ledger entry fails; W == 0 without a
//                  bath or clamps.
//   oracle_philox  Random123 known-answer vectors quoted at
//                  scripts/physics_refs/rng_streams.py:82-88, and the keyed
//                  counter layout of rng_streams.py:75-79.
//   oracle_schedule the director table of
//                  crates/prism-nhs/tests/protocol_director_thermal_contract.cpp:13-66
//                  and the friction ramp of fused_engine.rs:15370-15378.
// Parity unpinned: the reference's own Langevin noise (cuRAND XORWOW, a
// library outside the reference tree) is not reproduced; integrator parity
// is checked with the noise supplied as an explicit array.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr double kCoulomb = 332.0636;        // nhs_amber_fused.cu:66
constexpr double kKcalToAmu = 418.4;         // nhs_amber_fused.cu:69
constexpr double kBoltzmann = 0.001987204;   // nhs_amber_fused.cu:70
constexpr double kMaxForce = 1000.0;         // nhs_amber_fused.cu:71
constexpr double kPi = 3.14159265358979323846;

struct Vec { double x, y, z; };
inline Vec vsub(const double *a, const double *b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline double vdot(Vec a, Vec b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec vcross(Vec a, Vec b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline void vadd(double *f, Vec d, double s) { f[0] += s * d.x; f[1] += s * d.y; f[2] += s * d.z; }

// FP32 squared distance exactly as the reference's list/cutoff test rounds it:
// dx = xj - xi in float, r2 = (dx*dx + dy*dy) + dz*dz, no contraction (this
// file is compiled with -ffp-contract=off).  nhs_neighbor_warp.cuh:60-63,
// nhs_amber_fused.cu:3148-3159.
inline float r2_f32(const float *xi, const float *xj) {
    const float dx = xj[0] - xi[0], dy = xj[1] - xi[1], dz = xj[2] - xi[2];
    const float a = dx * dx;
    const float b = dy * dy;
    const float c = dz * dz;
    return (a + b) + c;
}

// FP32 squared distance of the nonbonded cutoff test: d = xi - xj in float,
// r2 = fma(dz, dz, fma(dy, dy, dx*dx)) -- the contracted form a CUDA/HIP
// compiler emits for the reference's dx*dx + dy*dy + dz*dz at
// nhs_amber_fused.cu:3156-3159, fixed to this order so that membership of
// every pair in the cutoff sphere is decided identically here and on the GPU.
inline float r2_cut_f32(const float *xi, const float *xj) {
    const float dx = xi[0] - xj[0], dy = xi[1] - xj[1], dz = xi[2] - xj[2];
    return std::fmaf(dz, dz, std::fmaf(dy, dy, dx * dx));
}

// Per-atom sorted partner sets derived from the topology.
struct Lowered {
    std::vector<std::vector<int>> skip_nonbonded;  // exclusions U exception partners
    std::vector<std::vector<int>> excl_12_13;      // exclusions minus exception partners
};

Lowered lower(const PrismMdTopologyV1 *t) {
    const int n = t->n_atoms;
    Lowered L;
    L.skip_nonbonded.resize(n);
    L.excl_12_13.resize(n);
    std::vector<std::vector<int>> ex(n);
    for (int e = 0; e < t->n_exceptions; ++e) {
        const int i = t->exception_atoms[2 * e], j = t->exception_atoms[2 * e + 1];
        ex[i].push_back(j);
        ex[j].push_back(i);
    }
    for (auto &v : ex) std::sort(v.begin(), v.end());
    // fused_engine.rs:11718-11731: exclusion CSR without 1-4 (exception) partners
    for (int i = 0; i < n; ++i) {
        for (int k = t->exclusion_offsets[i]; k < t->exclusion_offsets[i + 1]; ++k) {
            const int j = t->exclusion_atoms[k];
            L.skip_nonbonded[i].push_back(j);
            if (!std::binary_search(ex[i].begin(), ex[i].end(), j)) L.excl_12_13[i].push_back(j);
        }
        for (int j : ex[i]) L.skip_nonbonded[i].push_back(j);
        auto &s = L.skip_nonbonded[i];
        std::sort(s.begin(), s.end());
        s.erase(std::unique(s.begin(), s.end()), s.end());
        std::sort(L.excl_12_13[i].begin(), L.excl_12_13[i].end());
    }
    return L;
}

inline bool contains(const std::vector<int> &v, int x) { return std::binary_search(v.begin(), v.end(), x); }

// Cell hash in double with a given cell edge (>= interaction range plus margin,
// so every pair within range lies in adjacent cells).
struct Cells {
    double lo[3], edge;
    int dim[3];
    std::vector<int> head, next;
    Cells(const double *x, int n, double edge_) : edge(edge_) {
        double hi[3];
        for (int c = 0; c < 3; ++c) { lo[c] = 1e300; hi[c] = -1e300; }
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) { lo[c] = std::min(lo[c], x[3 * i + c]); hi[c] = std::max(hi[c], x[3 * i + c]); }
        for (int c = 0; c < 3; ++c) dim[c] = n ? std::max(1, (int)std::floor((hi[c] - lo[c]) / edge) + 1) : 1;
        head.assign((size_t)dim[0] * dim[1] * dim[2], -1);
        next.assign(n, -1);
        for (int i = n - 1; i >= 0; --i) {
            const size_t c = index(cell(x + 3 * i, 0), cell(x + 3 * i, 1), cell(x + 3 * i, 2));
            next[i] = head[c];
            head[c] = i;
        }
    }
    int cell(const double *p, int c) const { return std::min(dim[c] - 1, std::max(0, (int)std::floor((p[c] - lo[c]) / edge))); }
    size_t index(int a, int b, int c) const { return ((size_t)c * dim[1] + b) * dim[0] + a; }
    template <class F> void neighbours(const double *p, F &&f) const {
        const int a = cell(p, 0), b = cell(p, 1), c = cell(p, 2);
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int x = a + dx, y = b + dy, z = c + dz;
                    if (x < 0 || y < 0 || z < 0 || x >= dim[0] || y >= dim[1] || z >= dim[2]) continue;
                    for (int j = head[index(x, y, z)]; j >= 0; j = next[j]) f(j);
                }
    }
};

// ---- bonded terms -----------------------------------------------------------

// compute_bond_force, nhs_amber_fused.cu:354-386: U = k (r - r0)^2,
// rij = xj - xi, F/r = -2 k (r - r0)/r, fi -= f, fj += f.
double bond_term(const double *xi, const double *xj, double r0, double k, double *fi, double *fj) {
    const Vec rij = vsub(xj, xi);
    const double r = std::sqrt(vdot(rij, rij));
    if (r < 1e-8) return 0.0;
    const double dr = r - r0;
    const double fm = -2.0 * k * dr / r;
    vadd(fi, rij, -fm);
    vadd(fj, rij, fm);
    return k * dr * dr;
}

// compute_angle_force, nhs_amber_fused.cu:388-448: U = k (theta - theta0)^2 with
// the d cos(theta) gradients and the 1e-8 floor on sin(theta).
double angle_term(const double *xi, const double *xj, const double *xk, double theta0, double k, double *fi,
                  double *fj, double *fk) {
    const Vec rij = vsub(xi, xj), rkj = vsub(xk, xj);
    const double li = std::sqrt(vdot(rij, rij)), lk = std::sqrt(vdot(rkj, rkj));
    if (li < 1e-8 || lk < 1e-8) return 0.0;
    double c = vdot(rij, rkj) / (li * lk);
    c = std::max(-1.0, std::min(1.0, c));
    const double th = std::acos(c), dth = th - theta0;
    double s = std::sqrt(1.0 - c * c);
    if (s < 1e-8) s = 1e-8;
    const double fm = -2.0 * k * dth / s;
    const Vec di = {(rkj.x / lk - c * rij.x / li) / li, (rkj.y / lk - c * rij.y / li) / li, (rkj.z / lk - c * rij.z / li) / li};
    const Vec dk = {(rij.x / li - c * rkj.x / lk) / lk, (rij.y / li - c * rkj.y / lk) / lk, (rij.z / li - c * rkj.z / lk) / lk};
    vadd(fi, di, -fm);
    vadd(fk, dk, -fm);
    vadd(fj, Vec{di.x + dk.x, di.y + dk.y, di.z + dk.z}, fm);
    return k * dth * dth;
}

// compute_dihedral_force, nhs_amber_fused.cu:524-625: U = k (1 + cos(n phi - g))
// with OpenMM's atom ordering v0 = ri - rj, v1 = rk - rj, v2 = rk - rl and
// the complete chain rule.  phi is taken with atan2 (equal to the
// reference's acos/asin branches in exact arithmetic) and signed by v0 . cp1.
double dihedral_term(const double *xi, const double *xj, const double *xk, const double *xl, int n, double phase,
                     double k, double *fi, double *fj, double *fk, double *fl) {
    const Vec v0 = vsub(xi, xj), v1 = vsub(xk, xj), v2 = vsub(xk, xl);
    const Vec cp0 = vcross(v0, v1), cp1 = vcross(v1, v2);
    const double n1 = vdot(cp0, cp0), n2 = vdot(cp1, cp1), nb = vdot(v1, v1);
    if (n1 < 1e-16 || n2 < 1e-16 || nb < 1e-16) return 0.0;
    const Vec cc = vcross(cp0, cp1);
    double phi = std::atan2(std::sqrt(vdot(cc, cc)), vdot(cp0, cp1));
    if (vdot(v0, cp1) < 0.0) phi = -phi;
    const double dEdphi = -k * n * std::sin(n * phi - phase);
    const double lb = std::sqrt(nb);
    const double ff0 = (-dEdphi * lb) / n1;
    const double ff1 = vdot(v0, v1) / nb;
    const double ff2 = vdot(v2, v1) / nb;
    const double ff3 = (dEdphi * lb) / n2;
    const Vec f1 = {ff0 * cp0.x, ff0 * cp0.y, ff0 * cp0.z};
    const Vec f4 = {ff3 * cp1.x, ff3 * cp1.y, ff3 * cp1.z};
    const Vec s = {ff1 * f1.x - ff2 * f4.x, ff1 * f1.y - ff2 * f4.y, ff1 * f1.z - ff2 * f4.z};
    vadd(fi, f1, 1.0);
    vadd(fj, Vec{s.x - f1.x, s.y - f1.y, s.z - f1.z}, 1.0);
    vadd(fk, Vec{-s.x - f4.x, -s.y - f4.y, -s.z - f4.z}, 1.0);
    vadd(fl, f4, 1.0);
    return k * (1.0 + std::cos(n * phi - phase));
}

// Lennard-Jones + Coulomb pair, compute_nonbonded_force nhs_amber_fused.cu:643-717:
// returns F/r for d = xi - xj (force on i = (F/r) d) and the energy
// V_LJ + V_C - shift.  shift is ordinary_cutoff_potential_offset_v1
// (nhs_amber_fused.cu:631-640) for ordinary pairs and 0 for exceptions.
inline double pair_term(double r2, double sigma, double eps, double qq, double rc2, bool shifted, double *energy) {
    const double inv_r2 = 1.0 / r2, inv_r = std::sqrt(inv_r2);
    const double s2 = sigma * sigma * inv_r2, s6 = s2 * s2 * s2, s12 = s6 * s6;
    const double f = (24.0 * eps * (2.0 * s12 - s6) + kCoulomb * qq * inv_r) * inv_r2;
    double v = 4.0 * eps * (s12 - s6) + kCoulomb * qq * inv_r;
    if (shifted) {
        const double ic2 = 1.0 / rc2, ic = std::sqrt(ic2);
        const double c2 = sigma * sigma * ic2, c6 = c2 * c2 * c2;
        v -= 4.0 * eps * (c6 * c6 - c6) + kCoulomb * qq * ic;
    }
    *energy = v;
    return f;
}

}  // namespace

extern "C" int oracle_forces(const PrismMdTopologyV1 *t, const double *x, unsigned mask, int decide_f32, double *F,
                             double *E) {
    if (!t || !x || !F) return PRISM_MD_E_INVALID;
    const int n = t->n_atoms;
    std::fill(F, F + 3 * (size_t)n, 0.0);
    double e[PRISM_MD_N_TERMS] = {0, 0, 0, 0, 0};
    if (mask & PRISM_MD_TERM_BOND)
        for (int b = 0; b < t->n_bonds; ++b) {
            const int i = t->bond_atoms[2 * b], j = t->bond_atoms[2 * b + 1];
            e[0] += bond_term(x + 3 * i, x + 3 * j, t->bond_r0[b], t->bond_k[b], F + 3 * i, F + 3 * j);
        }
    if (mask & PRISM_MD_TERM_ANGLE)
        for (int a = 0; a < t->n_angles; ++a) {
            const int i = t->angle_atoms[3 * a], j = t->angle_atoms[3 * a + 1], k = t->angle_atoms[3 * a + 2];
            e[1] += angle_term(x + 3 * i, x + 3 * j, x + 3 * k, t->angle_theta0[a], t->angle_k[a], F + 3 * i,
                               F + 3 * j, F + 3 * k);
        }
    if (mask & PRISM_MD_TERM_DIHEDRAL)
        for (int d = 0; d < t->n_dihedrals; ++d) {
            const int *q = t->dihedral_atoms + 4 * d;
            e[2] += dihedral_term(x + 3 * q[0], x + 3 * q[1], x + 3 * q[2], x + 3 * q[3], t->dihedral_periodicity[d],
                                  t->dihedral_phase[d], t->dihedral_k[d], F + 3 * q[0], F + 3 * q[1], F + 3 * q[2],
                                  F + 3 * q[3]);
        }
    if (mask & PRISM_MD_TERM_EXCEPTION)
        for (int p = 0; p < t->n_exceptions; ++p) {
            const int i = t->exception_atoms[2 * p], j = t->exception_atoms[2 * p + 1];
            const Vec d = vsub(x + 3 * i, x + 3 * j);
            const double r2 = vdot(d, d);
            if (r2 < 1e-8) return PRISM_MD_E_NONFINITE;  // nhs_amber_fused.cu:3338-3343
            double v;
            const double f = pair_term(r2, t->exception_sigma[p], t->exception_epsilon[p],
                                       t->exception_charge_product[p], 0.0, false, &v);
            vadd(F + 3 * i, d, f);
            vadd(F + 3 * j, d, -f);
            e[4] += v;
        }
    if (mask & PRISM_MD_TERM_NONBONDED) {
        const Lowered L = lower(t);
        std::vector<float> xf(3 * (size_t)n);
        for (size_t k = 0; k < xf.size(); ++k) xf[k] = (float)x[k];
        const Cells cells(x, n, 10.5);
        std::vector<double> Fnb(3 * (size_t)n, 0.0), Enb(n, 0.0);
        // Per-atom gather over the full neighbourhood (each pair seen from both
        // ends, half the energy each), as the reference's atom-owned loop.
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < n; ++i) {
            double fi[3] = {0, 0, 0}, ei = 0.0;
            cells.neighbours(x + 3 * i, [&](int j) {
                if (j == i || contains(L.skip_nonbonded[i], j)) return;
                const Vec d = vsub(x + 3 * i, x + 3 * j);
                const double r2 = vdot(d, d);
                bool inside;
                if (decide_f32) {
                    const float q2 = r2_cut_f32(&xf[3 * i], &xf[3 * j]);
                    inside = !(q2 >= 100.0f || q2 < 0.01f);  // nhs_amber_fused.cu:3159
                } else {
                    inside = r2 < 100.0 && r2 >= 0.01;
                }
                if (!inside) return;
                const double sig = 0.5 * ((double)t->sigma[i] + (double)t->sigma[j]);
                const double eps = std::sqrt((double)t->epsilon[i] * (double)t->epsilon[j]);
                const double qq = (double)t->charges[i] * (double)t->charges[j];
                double v;
                const double f = pair_term(r2, sig, eps, qq, 100.0, true, &v);
                fi[0] += f * d.x;
                fi[1] += f * d.y;
                fi[2] += f * d.z;
                ei += 0.5 * v;
            });
            Fnb[3 * i] = fi[0];
            Fnb[3 * i + 1] = fi[1];
            Fnb[3 * i + 2] = fi[2];
            Enb[i] = ei;
        }
        for (size_t k = 0; k < Fnb.size(); ++k) F[k] += Fnb[k];
        for (int i = 0; i < n; ++i) e[3] += Enb[i];
    }
    if (E)
        for (int k = 0; k < PRISM_MD_N_TERMS; ++k) E[k] = e[k];
    return PRISM_MD_OK;
}

extern "C" int64_t oracle_pairs(const PrismMdTopologyV1 *t, const float *x, const float *x_grid, int mode,
                                int32_t *out_i, int32_t *out_j, int64_t cap) {
    if (!t || !x) return PRISM_MD_E_INVALID;
    const int n = t->n_atoms;
    const Lowered L = lower(t);
    std::vector<double> xd(3 * (size_t)n);
    for (size_t k = 0; k < xd.size(); ++k) xd[k] = x[k];
    // reference grid (fused_engine.rs:12285-12330), evaluated in float
    int ref_n[3] = {1, 1, 1};
    float ref_o[3] = {0, 0, 0};
    std::vector<int> ref_cell;
    if (mode == PRISM_MD_PAIRS_REFERENCE) {
        if (!x_grid) return PRISM_MD_E_INVALID;
        float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) {
                lo[c] = std::min(lo[c], x_grid[3 * i + c]);
                hi[c] = std::max(hi[c], x_grid[3 * i + c]);
            }
        for (int c = 0; c < 3; ++c) {
            ref_o[c] = lo[c] - 10.0f;
            const float extent = (hi[c] - lo[c]) + 2.0f * 10.0f;
            const long v = std::max(1L, (long)std::ceil(extent / 10.0f));
            if (v > 32) return PRISM_MD_E_CAPACITY;
            ref_n[c] = (int)v;
        }
        // build_cell_list, nhs_amber_fused.cu:11690-11728
        ref_cell.resize(3 * (size_t)n);
        std::vector<int> occupancy((size_t)ref_n[0] * ref_n[1] * ref_n[2], 0);
        for (int i = 0; i < n; ++i) {
            int id[3];
            for (int c = 0; c < 3; ++c) {
                const float g = (x[3 * i + c] - ref_o[c]) * 0.1f;
                id[c] = std::max(0, std::min((int)g, ref_n[c] - 1));
                ref_cell[3 * i + c] = id[c];
            }
            if (++occupancy[((size_t)id[2] * ref_n[1] + id[1]) * ref_n[0] + id[0]] > 512) return PRISM_MD_E_CAPACITY;
        }
    }
    const Cells cells(xd.data(), n, 12.5);
    std::vector<std::vector<int>> partners(n);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        auto &row = partners[i];
        cells.neighbours(xd.data() + 3 * i, [&](int j) {
            if (j == i) return;
            if (!(r2_f32(x + 3 * i, x + 3 * j) <= 144.0f)) return;  // keep = !(r2 > 144)
            if (contains(L.excl_12_13[i], j)) return;
            if (mode == PRISM_MD_PAIRS_REFERENCE) {
                for (int c = 0; c < 3; ++c)
                    if (std::abs(ref_cell[3 * i + c] - ref_cell[3 * j + c]) > 1) return;  // 27-cell stencil
            }
            row.push_back(j);
        });
        std::sort(row.begin(), row.end());
    }
    int64_t count = 0;
    for (int i = 0; i < n; ++i) {
        if (mode == PRISM_MD_PAIRS_REFERENCE && partners[i].size() >= 1024) return PRISM_MD_E_CAPACITY;
        for (int j : partners[i])
            if (j > i) {
                if (out_i && out_j && count < cap) {
                    out_i[count] = i;
                    out_j[count] = j;
                }
                ++count;
            }
    }
    return count;
}

extern "C" void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0, hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Langevin stream: key (seed low word, stream), counter (atom, step low word,
// step bits 32..47 << 16 | purpose << 8, 0) -- the injective packing of
// scripts/physics_refs/rng_streams.py:75-79 with replica epoch 0.  Box-Muller:
// u1 = (r0 + 1) 2^-32 in (0, 1], u2 = r1 2^-32, n0 = sqrt(-2 ln u1) cos(2 pi u2),
// n1 = ... sin(...), n2 from (r2, r3).
extern "C" void oracle_langevin_normals(uint64_t seed, uint32_t stream, uint32_t atom, uint64_t step,
                                        uint32_t purpose, double out[3]) {
    const uint32_t key[2] = {(uint32_t)seed, stream};
    const uint32_t ctr[4] = {atom, (uint32_t)step, (uint32_t)(((step >> 32) & 0xFFFFu) << 16) | ((purpose & 0xFFu) << 8),
                             0u};
    uint32_t r[4];
    oracle_philox4x32_10(ctr, key, r);
    const double s = 2.3283064365386963e-10;  // 2^-32
    const double a = std::sqrt(-2.0 * std::log(((double)r[0] + 1.0) * s)), ta = 2.0 * kPi * (double)r[1] * s;
    const double b = std::sqrt(-2.0 * std::log(((double)r[2] + 1.0) * s)), tb = 2.0 * kPi * (double)r[3] * s;
    out[0] = a * std::cos(ta);
    out[1] = a * std::sin(ta);
    out[2] = b * std::cos(tb);
}

// protocol_director.cu:58-123 (prism_protocol_baseline_v1): five-phase
// temperature table and the equilibration friction ramp; cryo scaling off.
extern "C" void oracle_schedule(const PrismMdParamsV1 *p, uint32_t step, double *T, double *gamma, int *hold) {
    const int s = (int)step;
    double temp;
    int h;
    if (s < p->cold_hold_end) { temp = p->temperature_K; h = 1; }
    else if (s < p->ramp_end) {
        const double progress = (double)(s - p->cold_hold_end) / (double)(p->ramp_end - p->cold_hold_end);
        temp = p->temperature_K + progress * (p->end_temperature_K - p->temperature_K);
        h = 0;
    } else if (s < p->warm_hold_end) { temp = p->end_temperature_K; h = 1; }
    else if (s < p->ramp_down_end) {
        const double progress = (double)(s - p->warm_hold_end) / (double)(p->ramp_down_end - p->warm_hold_end);
        temp = p->end_temperature_K - progress * (p->end_temperature_K - p->temperature_K);
        h = 0;
    } else { temp = p->warm_tail ? p->end_temperature_K : p->temperature_K; h = 1; }
    double g = p->gamma_ps;
    if (s < p->equilibration_steps) {
        const double decay = std::exp(-3.0 * (double)s / (double)p->equilibration_steps);
        g = p->equilibration_gamma_ps * decay + p->gamma_ps * (1.0 - decay);
    }
    if (T) *T = temp;
    if (gamma) *gamma = g;
    if (hold) *hold = h;
}

// shake_cluster_positions, nhs_amber_fused.cu:1889-2173 (non-compensated
// branch): Gauss-Seidel over the cluster's X-H bonds, correction along the
// step-start bond r0, velocity impulse correction/dt, relative tolerance on
// |r2 - d2|/d2, direction check |r.r0|/sqrt(r2 r0^2) > 1e-6.
extern "C" int oracle_shake_cluster(const PrismMdTopologyV1 *t, int c, double *x, const double *xs, double *v,
                                    double dt, double tol, int max_iter) {
    const int a = t->cluster_central[c];
    const double imc = t->cluster_inv_mass_central[c], imh = t->cluster_inv_mass_h[c];
    for (int it = 0; it < max_iter; ++it) {
        double max_err = 0.0;
        for (int h = 0; h < t->cluster_n_hydrogens[c]; ++h) {
            const int b = t->cluster_hydrogens[3 * c + h];
            if (b < 0) continue;
            const Vec r = vsub(x + 3 * b, x + 3 * a), r0 = vsub(xs + 3 * b, xs + 3 * a);
            const double d = t->cluster_bond_lengths[3 * c + h];
            const double r2 = vdot(r, r), d2 = d * d, diff = r2 - d2;
            if (!std::isfinite(r2) || !(d2 > 0.0)) return -1;
            const double rel = std::fabs(diff) / d2;
            max_err = std::max(max_err, rel);
            if (rel <= tol) continue;
            const double rr0 = vdot(r, r0), scale = std::sqrt(r2 * vdot(r0, r0));
            if (!std::isfinite(rr0) || !(scale > 0.0) || std::fabs(rr0) / scale <= 1e-6 || !(dt > 0.0)) return -1;
            const double lambda = diff / (2.0 * (imc + imh) * rr0);
            vadd(x + 3 * a, r0, imc * lambda);
            vadd(x + 3 * b, r0, -imh * lambda);
            vadd(v + 3 * a, r0, imc * lambda / dt);
            vadd(v + 3 * b, r0, -imh * lambda / dt);
        }
        if (max_err <= tol) return it + 1;
    }
    return -1;
}

// rattle_cluster_velocities, nhs_amber_fused.cu:2175-2360 (non-compensated
// branch): residual r.(v_h - v_c), relative error |res|/(|r| max(|v_rel|, 1)),
// lambda = res/((1/m_c + 1/m_h) r2).
extern "C" int oracle_rattle_cluster(const PrismMdTopologyV1 *t, int c, const double *x, double *v, double tol,
                                     int max_iter) {
    const int a = t->cluster_central[c];
    const double imc = t->cluster_inv_mass_central[c], imh = t->cluster_inv_mass_h[c];
    for (int it = 0; it < max_iter; ++it) {
        double max_err = 0.0;
        for (int h = 0; h < t->cluster_n_hydrogens[c]; ++h) {
            const int b = t->cluster_hydrogens[3 * c + h];
            if (b < 0) continue;
            const Vec r = vsub(x + 3 * b, x + 3 * a);
            const double r2 = vdot(r, r);
            if (!std::isfinite(r2) || r2 < 1e-12) return -1;
            const Vec vr = vsub(v + 3 * b, v + 3 * a);
            const double res = vdot(r, vr);
            const double rel = std::fabs(res) / (std::sqrt(r2) * std::max(std::sqrt(vdot(vr, vr)), 1.0));
            if (!std::isfinite(rel)) return -1;
            max_err = std::max(max_err, rel);
            if (rel <= tol) continue;
            const double lambda = res / ((imc + imh) * r2);
            vadd(v + 3 * a, r, imc * lambda);
            vadd(v + 3 * b, r, -imh * lambda);
        }
        if (max_err <= tol) return it + 1;
    }
    return -1;
}

extern "C" int oracle_energies(const PrismMdTopologyV1 *t, const double *x, const double *v, double *pe, double *ke) {
    std::vector<double> F(3 * (size_t)t->n_atoms);
    double E[PRISM_MD_N_TERMS];
    const int rc = oracle_forces(t, x, PRISM_MD_TERM_ALL, 0, F.data(), E);
    if (rc) return rc;
    if (pe) *pe = E[0] + E[1] + E[2] + E[3] + E[4];
    if (ke) {
        double k = 0.0;
        for (int i = 0; i < t->n_atoms; ++i)
            k += 0.5 * t->masses[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
        *ke = k / kKcalToAmu;
    }
    return PRISM_MD_OK;
}

namespace {

inline double ke_atom(double m, const double *v) { return 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / kKcalToAmu; }

// langevin_thermostat, nhs_amber_fused.cu:1713-1744, called with dt/2
// (nhs_amber_fused.cu:3656-3666): v <- c1 v + c2 xi per axis.
void ou_half(int n, const float *mass, double T, double gamma, double dt, double *v, const double *xi, double *work) {
    const double c1 = std::exp(-gamma * 0.5 * dt);
    for (int i = 0; i < n; ++i) {
        const double m = mass[i];
        const double c2 = std::sqrt((1.0 - c1 * c1) * kKcalToAmu * kBoltzmann * T / m);
        const double before = ke_atom(m, v + 3 * i);
        for (int c = 0; c < 3; ++c) v[3 * i + c] = c1 * v[3 * i + c] + c2 * xi[3 * i + c];
        if (work) *work += ke_atom(m, v + 3 * i) - before;
    }
}

// Half kick, nhs_amber_fused.cu:3667-3703 with
// apply_force_half_kick_with_intervention_work (:1750-1788) and
// prism_asc_used_force (prism_asc_decision.cuh:100-104): the force is clamped
// to |F| <= 1000 when gamma != 0, exact when gamma == 0.
void half_kick(int n, const float *mass, double dt, double gamma, const double *F, double *v, double *work) {
    for (int i = 0; i < n; ++i) {
        const double m = mass[i], kick = 0.5 * dt * kKcalToAmu / m;
        double q = 1.0;
        if (gamma != 0.0) {
            const double mag = std::sqrt(F[3 * i] * F[3 * i] + F[3 * i + 1] * F[3 * i + 1] + F[3 * i + 2] * F[3 * i + 2]);
            if (mag > kMaxForce) q = kMaxForce / mag;
        }
        double raw[3], used[3];
        for (int c = 0; c < 3; ++c) {
            raw[c] = v[3 * i + c] + kick * F[3 * i + c];
            used[c] = v[3 * i + c] + kick * F[3 * i + c] * q;
        }
        if (work && q != 1.0) *work += ke_atom(m, used) - ke_atom(m, raw);
        for (int c = 0; c < 3; ++c) v[3 * i + c] = used[c];
    }
}

// RATTLE over every cluster.  With work != NULL the projection's
// kinetic-energy change is booked with the cluster's masses 1/inv_mass
// (rattle_cluster_velocities, nhs_amber_fused.cu:2245-2270): the reference
// books the RATTLE after OU_N and the final one into its work ledger and
// routes the pre-drift and pre-OU_NP1 ones to separate constraint planes
// (fused_engine.rs:19273-19277, :17735-17740).
int rattle_all(const PrismMdTopologyV1 *t, const double *x, double *v, const PrismMdParamsV1 *p, double *work) {
    for (int c = 0; c < t->n_clusters; ++c) {
        const int a = t->cluster_central[c];
        const double mc = 1.0 / (double)t->cluster_inv_mass_central[c], mh = 1.0 / (double)t->cluster_inv_mass_h[c];
        auto ke = [&]() {
            double k = 0.5 * mc * vdot({v[3 * a], v[3 * a + 1], v[3 * a + 2]}, {v[3 * a], v[3 * a + 1], v[3 * a + 2]});
            for (int h = 0; h < t->cluster_n_hydrogens[c]; ++h) {
                const int b = t->cluster_hydrogens[3 * c + h];
                if (b >= 0) k += 0.5 * mh * (v[3 * b] * v[3 * b] + v[3 * b + 1] * v[3 * b + 1] + v[3 * b + 2] * v[3 * b + 2]);
            }
            return k / kKcalToAmu;
        };
        const double before = work ? ke() : 0.0;
        if (oracle_rattle_cluster(t, c, x, v, p->rattle_tol, p->constraint_max_iter) < 0) return PRISM_MD_E_CONSTRAINT;
        if (work) *work += ke() - before;
    }
    return PRISM_MD_OK;
}

}  // namespace

// One constrained OVRVO step per iteration in the reference's phase order
// (fused_engine.rs:6223-6235): FORCE_N, OU_HALF_N, RATTLE_AFTER_OU_N, KICK_N,
// RATTLE_PRE_DRIFT, DRIFT_SAVE_XN, SHAKE_POSITION, FORCE_NP1, KICK_NP1,
// RATTLE_BEFORE_OU_NP1, OU_HALF_NP1, then the post-phase velocity safety clamp
// (nhs_amber_fused.cu:11148-11186) and final RATTLE (:11128-11146)
// (fused_engine.rs:17688-17760).  COM velocity removal precedes step s when
// s > 0 and s % com_remove_every == 0 (fused_engine.rs:19023; housekeeping.cu:128-199).
// The director advances its counter before evaluating, so step s runs with
// the schedule of step s + 1 (protocol_director.cu:308-312).
extern "C" int oracle_run(const PrismMdTopologyV1 *t, const PrismMdParamsV1 *p, double *x, double *v, int64_t step0,
                          int64_t n_steps, const float *noise, double *work) {
    if (!t || !p || !x || !v) return PRISM_MD_E_INVALID;
    const int n = t->n_atoms;
    const double dt = p->dt_ps;
    std::vector<double> F(3 * (size_t)n), xs(3 * (size_t)n), xi(3 * (size_t)n);
    // cutoff membership decided in FP32 on the float-rounded coordinates, as
    // the reference decides it (nhs_amber_fused.cu:3156-3159)
    auto force = [&]() { return oracle_forces(t, x, PRISM_MD_TERM_ALL, 1, F.data(), nullptr); };
    auto draw = [&](int64_t s_local, int64_t s_global, int half) {
        for (int i = 0; i < n; ++i) {
            if (noise) {
                for (int c = 0; c < 3; ++c)
                    xi[3 * i + c] = noise[(((size_t)s_local * 2 + half) * n + i) * 3 + c];
            } else {
                oracle_langevin_normals(p->seed, p->stream_id, (uint32_t)i, (uint64_t)s_global, 1u + (uint32_t)half,
                                        &xi[3 * i]);
            }
        }
    };
    for (int64_t k = 0; k < n_steps; ++k) {
        const int64_t s = step0 + k;
        if (p->com_remove_every > 0 && s > 0 && s % p->com_remove_every == 0) {
            double pm[3] = {0, 0, 0}, mt = 0.0;
            for (int i = 0; i < n; ++i) {
                for (int c = 0; c < 3; ++c) pm[c] += t->masses[i] * v[3 * i + c];
                mt += t->masses[i];
            }
            if (mt > 0.0)
                for (int i = 0; i < n; ++i)
                    for (int c = 0; c < 3; ++c) v[3 * i + c] -= pm[c] / mt;
        }
        double T, gamma;
        oracle_schedule(p, (uint32_t)(s + 1), &T, &gamma, nullptr);
        int rc = force();
        if (rc) return rc;
        if (gamma != 0.0) {
            draw(k, s, 0);
            ou_half(n, t->masses, T, gamma, dt, v, xi.data(), work);
        }
        if ((rc = rattle_all(t, x, v, p, work))) return rc;  // RATTLE_AFTER_OU_N: booked
        half_kick(n, t->masses, dt, gamma, F.data(), v, work);
        if ((rc = rattle_all(t, x, v, p, nullptr))) return rc;  // RATTLE_PRE_DRIFT
        for (size_t q = 0; q < xs.size(); ++q) {
            xs[q] = x[q];
            x[q] += dt * v[q];
        }
        for (int c = 0; c < t->n_clusters; ++c)
            if (oracle_shake_cluster(t, c, x, xs.data(), v, dt, p->shake_tol, p->constraint_max_iter) < 0)
                return PRISM_MD_E_CONSTRAINT;
        if ((rc = force())) return rc;
        half_kick(n, t->masses, dt, gamma, F.data(), v, work);
        if ((rc = rattle_all(t, x, v, p, nullptr))) return rc;  // RATTLE_BEFORE_OU_NP1
        if (gamma != 0.0) {
            draw(k, s, 1);
            ou_half(n, t->masses, T, gamma, dt, v, xi.data(), work);
        }
        for (int i = 0; i < n; ++i) {
            const double m = t->masses[i];
            const double vmax = std::max(20.0 * std::sqrt(kKcalToAmu * kBoltzmann * T / std::max(m, 1.0)), 2.0);
            const double sp = std::sqrt(v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
            if (sp > vmax) {
                const double before = ke_atom(m, v + 3 * i);
                for (int c = 0; c < 3; ++c) v[3 * i + c] *= vmax / sp;
                if (work) *work += ke_atom(m, v + 3 * i) - before;
            }
        }
        if ((rc = rattle_all(t, x, v, p, work))) return rc;  // final RATTLE: booked
        for (size_t q = 0; q < xs.size(); ++q)
            if (!std::isfinite(x[q]) || !std::isfinite(v[q])) return PRISM_MD_E_NONFINITE;
    }
    return PRISM_MD_OK;
}
