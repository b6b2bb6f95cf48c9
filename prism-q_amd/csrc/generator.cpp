// Seeded synthetic AMBER-typed chain generator (host C++, no device code).
//
// Builds the inputs of the classical step in the prism_md.h layout.  The
// heavy-atom backbone is grown as a self-avoiding walk: each new atom sits at
// 1.53 A from its predecessor, at the configured bond angle, and its torsion
// is the best of n_candidates trial azimuths (largest distance to every atom
// already placed, capped at 4 A), restricted to the confining sphere.
// Hydrogens are placed tetrahedrally once both backbone neighbours of their
// heavy atom exist.  Masses follow hydrogen mass repartitioning exactly as
// apply_hmr does (crates/prism-nhs/src/input.rs:2482-2565): m_H' = f m_H and
// the heavy atom gives up n_H (m_H' - m_H), unless it would drop below 1 amu.
#include "prism_gen.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <unordered_map>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kBoltzmann = 0.001987204;   // kcal/mol/K
constexpr double kKcalToAmuA2Ps2 = 418.4;

struct Rng {  // xoshiro256** seeded through splitmix64
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (auto &w : s) {
            z += 0x9E3779B97F4A7C15ull;
            uint64_t x = z;
            x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
            x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
            w = x ^ (x >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    int below(int n) { return (int)(uniform() * n); }
    double normal() {  // Box-Muller, one value per call
        double u1 = uniform();
        while (u1 <= 0.0) u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    }
};

struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 unit(V3 a) { return (1.0 / norm(a)) * a; }

// Orthonormal pair perpendicular to unit vector u.
void perp_basis(V3 u, V3 &n1, V3 &n2) {
    V3 t = std::fabs(u.x) < 0.9 ? V3{1, 0, 0} : V3{0, 1, 0};
    n1 = unit(cross(u, t));
    n2 = cross(u, n1);
}

// Spatial hash of placed atoms for the self-avoidance score.
struct Grid {
    double cell, origin;
    int dim;
    std::vector<std::vector<int>> cells;
    Grid(double radius, double cell_) : cell(cell_) {
        origin = -radius - 8.0;
        dim = (int)std::ceil((2.0 * radius + 16.0) / cell) + 1;
        cells.resize((size_t)dim * dim * dim);
    }
    int coord(double v) const { return std::max(0, std::min(dim - 1, (int)std::floor((v - origin) / cell))); }
    size_t index(int ix, int iy, int iz) const { return ((size_t)iz * dim + iy) * dim + ix; }
    void insert(int atom, V3 p) { cells[index(coord(p.x), coord(p.y), coord(p.z))].push_back(atom); }
    void remove(int atom, V3 p) {
        auto &c = cells[index(coord(p.x), coord(p.y), coord(p.z))];
        for (size_t i = 0; i < c.size(); ++i)
            if (c[i] == atom) {
                c[i] = c.back();
                c.pop_back();
                return;
            }
    }
    // Minimum distance from p to placed atoms not in `skip`, capped at `cap`.
    double min_distance(V3 p, const std::vector<V3> &pos, const int *skip, int n_skip, double cap) const {
        double best2 = cap * cap;
        const int cx = coord(p.x), cy = coord(p.y), cz = coord(p.z);
        const int reach = (int)std::ceil(cap / cell);
        for (int dz = -reach; dz <= reach; ++dz)
            for (int dy = -reach; dy <= reach; ++dy)
                for (int dx = -reach; dx <= reach; ++dx) {
                    const int x = cx + dx, y = cy + dy, z = cz + dz;
                    if (x < 0 || y < 0 || z < 0 || x >= dim || y >= dim || z >= dim) continue;
                    for (int a : cells[index(x, y, z)]) {
                        bool skipped = false;
                        for (int s = 0; s < n_skip; ++s) skipped |= (skip[s] == a);
                        if (skipped) continue;
                        const V3 d = pos[a] - p;
                        best2 = std::min(best2, dot(d, d));
                    }
                }
        return std::sqrt(best2);
    }
};

}  // namespace

struct PrismGenHandle {
    PrismGenParamsV1 params{};
    PrismGenStatsV1 stats{};
    PrismMdTopologyV1 topo{};
    std::vector<float> positions, masses, charges, sigma, epsilon;
    std::vector<int32_t> bond_atoms;
    std::vector<float> bond_r0, bond_k;
    std::vector<int32_t> angle_atoms;
    std::vector<float> angle_theta0, angle_k;
    std::vector<int32_t> dihedral_atoms, dihedral_periodicity;
    std::vector<float> dihedral_phase, dihedral_k;
    std::vector<uint8_t> dihedral_improper;
    std::vector<int32_t> exclusion_offsets, exclusion_atoms;
    std::vector<int32_t> exception_atoms;
    std::vector<float> exception_qq, exception_sigma, exception_epsilon;
    std::vector<int32_t> cluster_central, cluster_hydrogens, cluster_n_hydrogens;
    std::vector<float> cluster_bond_lengths, cluster_inv_mass_central, cluster_inv_mass_h;
};

extern "C" int prism_gen_default_params(PrismGenParamsV1 *p) {
    if (!p) return PRISM_MD_E_INVALID;
    std::memset(p, 0, sizeof(*p));
    p->n_atoms = 2000;
    p->seed = 0x50524953ull;
    p->density = 0.1;
    p->heavy_bond_A = 1.53;
    p->h_bond_A = 1.09;
    p->heavy_angle_deg = 111.0;
    p->hmr_factor = 3.0;
    p->charge_max = 0.8;
    p->sigma_min = 1.0;
    p->sigma_max = 3.5;
    p->epsilon_min = 0.01;
    p->epsilon_max = 0.2;
    p->bond_k_min = 300.0;
    p->bond_k_max = 500.0;
    p->angle_k_min = 40.0;
    p->angle_k_max = 80.0;
    p->torsion_vn_max = 2.0;
    p->scale14_lj = 0.5;
    p->scale14_qq = 1.0 / 1.2;
    p->n_candidates = 24;
    return PRISM_MD_OK;
}

// FIRE relaxation (Bitzek et al., PRL 97, 170201 (2006)) of the grown chain:
// soft harmonic repulsion between atoms more than two bonds apart, stiff
// springs on bonds and 1-3 distances, and a harmonic wall at the sphere
// radius.  Afterwards every hydrogen is put back at exactly h_bond from its
// heavy atom so the X-H constraints hold at step 0.
static void relax(std::vector<V3> &pos, const std::vector<std::vector<int>> &adj, const std::vector<int> &heavy_index,
                  const std::vector<int> &n_h, const std::vector<int> &first_h, const std::vector<float> &sigma,
                  double radius, double b, double bh, double theta, double tet) {
    const int N = (int)pos.size();
    std::vector<char> is_h(N, 0);
    for (size_t k = 0; k < heavy_index.size(); ++k)
        for (int h = 0; h < n_h[k]; ++h) is_h[first_h[k] + h] = 1;
    // springs: (i, j, target, k)
    struct Spring { int i, j; double d, k; };
    std::vector<Spring> springs;
    for (int i = 0; i < N; ++i)
        for (int j : adj[i])
            if (i < j) springs.push_back({i, j, (is_h[i] || is_h[j]) ? bh : b, 200.0});
    for (int j = 0; j < N; ++j)
        for (size_t x = 0; x < adj[j].size(); ++x)
            for (size_t y = x + 1; y < adj[j].size(); ++y) {
                const int i = adj[j][x], k = adj[j][y];
                const double a = (is_h[i] ? bh : b), c = (is_h[k] ? bh : b);
                const double th = (is_h[i] || is_h[k]) ? tet : theta;
                springs.push_back({i, k, std::sqrt(a * a + c * c - 2 * a * c * std::cos(th)), 60.0});
            }
    // topological distance <= 2 partners are exempt from the repulsion
    std::vector<std::vector<int>> near(N);
    for (int i = 0; i < N; ++i) {
        for (int j : adj[i]) {
            near[i].push_back(j);
            for (int k : adj[j])
                if (k != i) near[i].push_back(k);
        }
        std::sort(near[i].begin(), near[i].end());
    }
    // contact distance: element-based floor, raised to 0.8 sigma_ij so the
    // Lennard-Jones force of the relaxed geometry stays bounded
    auto contact = [&](int i, int j) {
        const double base = (is_h[i] && is_h[j]) ? 2.1 : ((is_h[i] || is_h[j]) ? 2.4 : 3.0);
        return std::max(base, 0.4 * ((double)sigma[i] + (double)sigma[j]));
    };
    const double rep_k = 20.0, cell = 3.0;  // cell >= largest contact (2.8 or 3.0)
    const double origin = -radius - 10.0;
    const int dim = (int)std::ceil((2.0 * radius + 20.0) / cell) + 1;
    std::vector<int> head, next(N);
    std::vector<V3> vel(N, V3{0, 0, 0}), frc(N), rep(N);
    auto cidx = [&](double v) { return std::max(0, std::min(dim - 1, (int)std::floor((v - origin) / cell))); };
    auto forces = [&]() {
        head.assign((size_t)dim * dim * dim, -1);
        for (int i = 0; i < N; ++i) {
            const size_t c = ((size_t)cidx(pos[i].z) * dim + cidx(pos[i].y)) * dim + cidx(pos[i].x);
            next[i] = head[c];
            head[c] = i;
        }
        for (int i = 0; i < N; ++i) frc[i] = V3{0, 0, 0};
        for (const Spring &sp : springs) {
            const V3 d = pos[sp.j] - pos[sp.i];
            const double r = std::max(norm(d), 1e-6);
            const V3 f = (2.0 * sp.k * (r - sp.d) / r) * d;
            frc[sp.i] = frc[sp.i] + f;
            frc[sp.j] = frc[sp.j] - f;
        }
#pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i) {
            V3 fi{0, 0, 0};
            const int cx = cidx(pos[i].x), cy = cidx(pos[i].y), cz = cidx(pos[i].z);
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int x = cx + dx, y = cy + dy, z = cz + dz;
                        if (x < 0 || y < 0 || z < 0 || x >= dim || y >= dim || z >= dim) continue;
                        for (int j = head[((size_t)z * dim + y) * dim + x]; j >= 0; j = next[j]) {
                            if (j == i) continue;
                            const V3 d = pos[j] - pos[i];
                            const double r2 = dot(d, d), c0 = contact(i, j);
                            if (r2 >= c0 * c0) continue;
                            if (std::binary_search(near[i].begin(), near[i].end(), j)) continue;
                            const double r = std::max(std::sqrt(r2), 1e-6);
                            fi = fi + (2.0 * rep_k * (r - c0) / r) * d;
                        }
                    }
            const double r = norm(pos[i]);
            if (r > radius) fi = fi - (2.0 * 10.0 * (r - radius) / r) * pos[i];
            rep[i] = fi;
        }
        for (int i = 0; i < N; ++i) frc[i] = frc[i] + rep[i];
    };
    // FIRE with unit masses
    double dt = 0.02, alpha = 0.1;
    const double dt_max = 0.1;
    int positive = 0;
    for (int it = 0; it < 400; ++it) {
        forces();
        double P = 0.0, vn = 0.0, fn = 0.0;
        for (int i = 0; i < N; ++i) {
            P += dot(frc[i], vel[i]);
            vn += dot(vel[i], vel[i]);
            fn += dot(frc[i], frc[i]);
        }
        vn = std::sqrt(vn);
        fn = std::sqrt(fn);
        if (P > 0.0) {
            for (int i = 0; i < N; ++i) vel[i] = (1.0 - alpha) * vel[i] + (fn > 0 ? alpha * vn / fn : 0.0) * frc[i];
            if (++positive > 5) {
                dt = std::min(dt * 1.1, dt_max);
                alpha *= 0.99;
            }
        } else {
            positive = 0;
            dt *= 0.5;
            alpha = 0.1;
            for (int i = 0; i < N; ++i) vel[i] = V3{0, 0, 0};
        }
        for (int i = 0; i < N; ++i) {
            vel[i] = vel[i] + dt * frc[i];
            V3 step = dt * vel[i];
            const double sn = norm(step);
            if (sn > 0.2) step = (0.2 / sn) * step;
            pos[i] = pos[i] + step;
        }
    }
    for (size_t k = 0; k < heavy_index.size(); ++k) {
        const V3 pc = pos[heavy_index[k]];
        for (int h = 0; h < n_h[k]; ++h) {
            const int a = first_h[k] + h;
            pos[a] = pc + bh * unit(pos[a] - pc);
        }
    }
}

static int generate(PrismGenHandle &g) {
    const PrismGenParamsV1 &P = g.params;
    const int N = P.n_atoms;
    Rng rng(P.seed);

    // ---- hydrogen counts: every heavy atom carries 1 or 2 hydrogens ----------
    std::vector<int> n_h;
    for (int remaining = N; remaining > 0;) {
        int h;
        if (remaining == 2) h = 1;
        else if (remaining == 3) h = 2;
        else if (remaining == 4) h = 1;  // avoid leaving a single atom
        else h = 1 + rng.below(2);
        n_h.push_back(h);
        remaining -= 1 + h;
    }
    const int H = (int)n_h.size();
    const double radius = std::cbrt(3.0 * N / (4.0 * kPi * P.density));

    // atom numbering: heavy atom k, then its hydrogens, chain order
    std::vector<int> heavy_index(H), first_h(H);
    {
        int a = 0;
        for (int k = 0; k < H; ++k) {
            heavy_index[k] = a;
            first_h[k] = a + 1;
            a += 1 + n_h[k];
        }
    }
    std::vector<V3> pos(N, V3{0, 0, 0});
    std::vector<char> placed(N, 0);
    Grid grid(radius, 3.0);
    auto place = [&](int atom, V3 p) {
        pos[atom] = p;
        placed[atom] = 1;
        grid.insert(atom, p);
    };

    const double b = P.heavy_bond_A, bh = P.h_bond_A;
    const double theta = P.heavy_angle_deg * kPi / 180.0;
    const int K = std::max(4, P.n_candidates);
    const double cap = 3.0;

    const double tet = 109.4712206 * kPi / 180.0;

    // Hydrogen positions of heavy atom k given its backbone neighbours (one or
    // two), and their clearance: the smallest distance to any placed atom that
    // is not bonded to or across the centre.  Returns the clearance.
    auto hydrogen_set = [&](int k, V3 pc, const V3 *nb, const int *nb_atoms, int n_nb, V3 *out) -> double {
        const int c = heavy_index[k];
        int skip[3] = {c, nb_atoms[0], n_nb > 1 ? nb_atoms[1] : c};
        double best = -1.0;
        if (n_nb == 2) {
            const V3 u1 = unit(nb[0] - pc), u2 = unit(nb[1] - pc);
            const V3 w = u1 + u2;
            V3 bis, nrm = cross(u1, u2);
            if (norm(nrm) < 1e-9 || norm(w) < 1e-9) {
                perp_basis(u1, nrm, bis);
            } else {
                nrm = unit(nrm);
                bis = unit(-1.0 * w);
            }
            const double half = 0.5 * tet;
            const V3 h2[2] = {pc + bh * (std::cos(half) * bis + std::sin(half) * nrm),
                              pc + bh * (std::cos(half) * bis - std::sin(half) * nrm)};
            const double d0 = grid.min_distance(h2[0], pos, skip, 3, cap);
            const double d1 = grid.min_distance(h2[1], pos, skip, 3, cap);
            if (n_h[k] == 2) {
                out[0] = h2[0];
                out[1] = h2[1];
                best = std::min(d0, d1);
            } else {
                out[0] = d0 >= d1 ? h2[0] : h2[1];
                best = std::max(d0, d1);
            }
        } else {
            // one backbone neighbour: hydrogens on the tetrahedral cone around
            // that bond, best of K azimuths
            const V3 u = unit(nb[0] - pc);
            V3 n1, n2;
            perp_basis(u, n1, n2);
            const double off = rng.uniform(0.0, 2.0 * kPi);
            for (int t = 0; t < K; ++t) {
                const double phi = off + 2.0 * kPi * t / K;
                V3 set[3];
                double score = cap;
                for (int h = 0; h < n_h[k]; ++h) {
                    const double ph = phi + h * 2.0 * kPi / 3.0;
                    set[h] = pc + bh * (std::cos(tet) * u + std::sin(tet) * (std::cos(ph) * n1 + std::sin(ph) * n2));
                    score = std::min(score, grid.min_distance(set[h], pos, skip, 2, cap));
                }
                if (score > best) {
                    best = score;
                    for (int h = 0; h < n_h[k]; ++h) out[h] = set[h];
                }
            }
        }
        return best;
    };
    auto place_hydrogens = [&](int k, const V3 *hpos) {
        for (int h = 0; h < n_h[k]; ++h) place(first_h[k] + h, hpos[h]);
    };
    auto unplace = [&](int atom) {
        if (!placed[atom]) return;
        grid.remove(atom, pos[atom]);
        placed[atom] = 0;
    };

    // ---- backbone growth --------------------------------------------------------
    // Step k places heavy atom k and the hydrogens of heavy atom k-1 (whose two
    // backbone neighbours are then known).  A step whose best candidate leaves
    // less than `clearance` to the placed atoms is undone together with a few
    // preceding steps and regrown, within a fixed budget.
    const double clearance = 2.0;
    long budget = 0;  // backtracking off: the relaxation below resolves contacts
    if (H >= 1) place(heavy_index[0], V3{0, 0, 0});
    if (H >= 2) {
        V3 d{rng.normal(), rng.normal(), rng.normal()};
        place(heavy_index[1], b * unit(d));
        V3 hp[3];
        const V3 nb0[1] = {pos[heavy_index[1]]};
        const int nba[1] = {heavy_index[1]};
        hydrogen_set(0, pos[heavy_index[0]], nb0, nba, 1, hp);
        place_hydrogens(0, hp);
    }
    for (int k = 2; k < H;) {
        const V3 pa = pos[heavy_index[k - 2]], pb = pos[heavy_index[k - 1]];
        const V3 u = unit(pb - pa);
        V3 n1, n2;
        perp_basis(u, n1, n2);
        const double off = rng.uniform(0.0, 2.0 * kPi);
        int skip[2] = {heavy_index[k - 1], heavy_index[k - 2]};
        double best_score = -1.0, best_r = 1e300;
        V3 best_p{0, 0, 0}, best_h[3], best_out{0, 0, 0};
        bool have_inside = false;
        for (int t = 0; t < K; ++t) {
            const double phi = off + 2.0 * kPi * t / K;
            const V3 dir = std::cos(kPi - theta) * u + std::sin(kPi - theta) * (std::cos(phi) * n1 + std::sin(phi) * n2);
            const V3 q = pb + b * dir;
            const double rq = norm(q);
            if (rq <= radius) {
                double s = grid.min_distance(q, pos, skip, 2, cap);
                if (have_inside && s <= best_score) continue;
                V3 hp[3];
                const V3 nb[2] = {pa, q};
                const int nba[2] = {heavy_index[k - 2], heavy_index[k]};
                s = std::min(s, hydrogen_set(k - 1, pb, nb, nba, 2, hp));
                if (!have_inside || s > best_score) {
                    best_score = s;
                    best_p = q;
                    for (int h = 0; h < 3; ++h) best_h[h] = hp[h];
                    have_inside = true;
                }
            } else if (!have_inside && rq < best_r) {
                best_r = rq;
                best_out = q;
            }
        }
        if ((!have_inside || best_score < clearance) && budget > 0 && k > 3) {
            // undo 1..4 steps and regrow
            const int m = std::min(k - 3, 1 + rng.below(4));
            for (int s = 0; s < m; ++s) {
                --k;
                unplace(heavy_index[k]);
                for (int h = 0; h < n_h[k - 1]; ++h) unplace(first_h[k - 1] + h);
            }
            budget -= m;
            continue;
        }
        if (!have_inside) {
            best_p = best_out;
            const V3 nb[2] = {pa, best_p};
            const int nba[2] = {heavy_index[k - 2], heavy_index[k]};
            hydrogen_set(k - 1, pb, nb, nba, 2, best_h);
        }
        place(heavy_index[k], best_p);
        place_hydrogens(k - 1, best_h);
        ++k;
    }
    if (H >= 2) {
        V3 hp[3];
        const V3 nb[1] = {pos[heavy_index[H - 2]]};
        const int nba[1] = {heavy_index[H - 2]};
        hydrogen_set(H - 1, pos[heavy_index[H - 1]], nb, nba, 1, hp);
        place_hydrogens(H - 1, hp);
    } else if (H == 1) {
        V3 hp[3];
        const V3 nb[1] = {V3{1, 0, 0}};
        const int nba[1] = {heavy_index[0]};
        hydrogen_set(0, pos[heavy_index[0]], nb, nba, 1, hp);
        place_hydrogens(0, hp);
    }

    // ---- bond graph -------------------------------------------------------------
    std::vector<std::vector<int>> adj(N);
    auto add_bond = [&](int i, int j, double r0) {
        adj[i].push_back(j);
        adj[j].push_back(i);
        g.bond_atoms.push_back(i);
        g.bond_atoms.push_back(j);
        g.bond_r0.push_back((float)r0);
        g.bond_k.push_back((float)rng.uniform(P.bond_k_min, P.bond_k_max));
    };
    for (int k = 0; k < H; ++k) {
        if (k > 0) add_bond(heavy_index[k - 1], heavy_index[k], b);
        for (int h = 0; h < n_h[k]; ++h) add_bond(heavy_index[k], first_h[k] + h, bh);
    }
    for (auto &a : adj) std::sort(a.begin(), a.end());

    // ---- per-atom parameters -------------------------------------------------
    g.masses.assign(N, 0.0f);
    g.charges.assign(N, 0.0f);
    g.sigma.assign(N, 0.0f);
    g.epsilon.assign(N, 0.0f);
    std::vector<double> mass(N), q(N);
    for (int k = 0; k < H; ++k) {
        const double r = rng.uniform();
        mass[heavy_index[k]] = r < 0.6 ? 12.011 : (r < 0.8 ? 14.007 : 15.999);
        for (int h = 0; h < n_h[k]; ++h) mass[first_h[k] + h] = 1.008;
    }
    // The sampled values are exactly the stated uniform draws -- N values of
    // sigma ~ U(sigma_min, sigma_max), q ~ U(-charge_max, charge_max),
    // epsilon ~ U(epsilon_min, epsilon_max) -- but they are handed out by
    // rank, as a force field would: hydrogens receive the smaller sigma and
    // epsilon draws and the more positive charges, heavy atoms the larger
    // sigma and epsilon and the more negative charges, and within each class
    // |q| grows with sigma.  Drawn per atom independently, oppositely charged
    // small hydrogens pair up against the r^-12 wall (a sigma = 1 A atom next
    // to an opposite 0.8 e charge balances at ~0.7 A; 2.2 A hydrogens with
    // +-0.4 e meet at 1.8 A), and those 3 amu contacts vibrate with periods of
    // a few femtoseconds, which a 4 fs integrator heats until SHAKE fails.
    std::vector<int> order(N);
    for (int a = 0; a < N; ++a) order[a] = a;
    std::vector<double> key(N);
    for (int a = 0; a < N; ++a) key[a] = (mass[a] < 2.0 ? 0.0 : 1.0) + rng.uniform();  // H first, random within
    std::sort(order.begin(), order.end(), [&](int x, int y) { return key[x] < key[y]; });
    int n_hyd = 0;
    for (int a = 0; a < N; ++a) n_hyd += mass[a] < 2.0;
    std::vector<double> us(N), ue(N), uq(N);
    for (int a = 0; a < N; ++a) {
        us[a] = rng.uniform();
        ue[a] = rng.uniform();
        uq[a] = P.charge_max * (2.0 * rng.uniform() - 1.0);
    }
    std::sort(us.begin(), us.end());
    std::sort(ue.begin(), ue.end());
    std::sort(uq.begin(), uq.end(), [](double x, double y) { return x > y; });  // most positive first
    // hydrogens: the n_hyd most positive charges, ascending with sigma rank;
    // heavy atoms: the rest, descending (more negative) with sigma rank
    std::sort(uq.begin(), uq.begin() + n_hyd);
    double qsum = 0.0;
    for (int r = 0; r < N; ++r) {
        const int a = order[r];
        q[a] = uq[r];
        qsum += q[a];
        g.sigma[a] = (float)(P.sigma_min + (P.sigma_max - P.sigma_min) * us[r]);
        g.epsilon[a] = (float)(P.epsilon_min + (P.epsilon_max - P.epsilon_min) * ue[r]);
    }
    for (int a = 0; a < N; ++a) g.charges[a] = (float)(q[a] - qsum / N);

    relax(pos, adj, heavy_index, n_h, first_h, g.sigma, radius, b, bh, theta, tet);

    // angles i-j-k for every pair of bonds sharing j; theta0 = built angle
    for (int j = 0; j < N; ++j)
        for (size_t x = 0; x < adj[j].size(); ++x)
            for (size_t y = x + 1; y < adj[j].size(); ++y) {
                const int i = adj[j][x], k = adj[j][y];
                const V3 a = unit(pos[i] - pos[j]), c = unit(pos[k] - pos[j]);
                const double th = std::acos(std::max(-1.0, std::min(1.0, dot(a, c))));
                g.angle_atoms.insert(g.angle_atoms.end(), {i, j, k});
                g.angle_theta0.push_back((float)th);
                g.angle_k.push_back((float)rng.uniform(P.angle_k_min, P.angle_k_max));
            }

    // proper torsions i-j-k-l around every bond j-k
    auto add_torsion = [&](int i, int j, int k, int l, bool improper) {
        g.dihedral_atoms.insert(g.dihedral_atoms.end(), {i, j, k, l});
        if (improper) {
            g.dihedral_periodicity.push_back(2);
            g.dihedral_phase.push_back((float)kPi);
        } else {
            g.dihedral_periodicity.push_back(1 + rng.below(4));
            g.dihedral_phase.push_back(rng.below(2) ? (float)kPi : 0.0f);
        }
        g.dihedral_k.push_back((float)(0.5 * rng.uniform(0.0, P.torsion_vn_max)));
        g.dihedral_improper.push_back(improper ? 1 : 0);
    };
    for (size_t t = 0; t < g.bond_r0.size(); ++t) {
        const int j = g.bond_atoms[2 * t], k = g.bond_atoms[2 * t + 1];
        for (int i : adj[j]) {
            if (i == k) continue;
            for (int l : adj[k]) {
                if (l == j || l == i) continue;
                add_torsion(i, j, k, l, false);
            }
        }
    }
    // impropers on three-coordinate centres (AMBER order: centre third)
    for (int c = 0; c < N; ++c)
        if (adj[c].size() == 3) add_torsion(adj[c][0], adj[c][1], c, adj[c][2], true);

    // hydrogen clusters + HMR (apply_hmr semantics)
    for (int k = 0; k < H; ++k) {
        const int c = heavy_index[k];
        const double mh = 1.008, mh_new = mh * P.hmr_factor;
        double mc_new = mass[c] - n_h[k] * (mh_new - mh);
        double mh_used = mh_new;
        if (mc_new < 1.0) {
            mc_new = mass[c];
            mh_used = mh;
        }
        mass[c] = mc_new;
        for (int h = 0; h < n_h[k]; ++h) mass[first_h[k] + h] = mh_used;
        g.cluster_central.push_back(c);
        for (int h = 0; h < 3; ++h) {
            g.cluster_hydrogens.push_back(h < n_h[k] ? first_h[k] + h : -1);
            g.cluster_bond_lengths.push_back(h < n_h[k] ? (float)bh : 0.0f);
        }
        g.cluster_n_hydrogens.push_back(n_h[k]);
        g.cluster_inv_mass_central.push_back((float)(1.0 / mc_new));
        g.cluster_inv_mass_h.push_back((float)(1.0 / mh_used));
    }
    for (int a = 0; a < N; ++a) g.masses[a] = (float)mass[a];

    // ---- exclusions (1-2, 1-3, 1-4) and exact 1-4 exceptions ----------------
    g.exclusion_offsets.assign(N + 1, 0);
    std::vector<int> dist(N, -1), touched;
    for (int a = 0; a < N; ++a) {
        touched.clear();
        dist[a] = 0;
        touched.push_back(a);
        std::vector<int> frontier{a};
        std::vector<int> partners;
        for (int depth = 1; depth <= 3; ++depth) {
            std::vector<int> next;
            for (int u : frontier)
                for (int v : adj[u])
                    if (dist[v] < 0) {
                        dist[v] = depth;
                        touched.push_back(v);
                        next.push_back(v);
                        partners.push_back(v);
                    }
            frontier.swap(next);
        }
        std::sort(partners.begin(), partners.end());
        for (int p : partners) {
            g.exclusion_atoms.push_back(p);
            if (dist[p] == 3 && a < p) {
                g.exception_atoms.push_back(a);
                g.exception_atoms.push_back(p);
                g.exception_qq.push_back((float)(P.scale14_qq * (double)g.charges[a] * (double)g.charges[p]));
                g.exception_sigma.push_back((float)(0.5 * ((double)g.sigma[a] + (double)g.sigma[p])));
                g.exception_epsilon.push_back(
                    (float)(P.scale14_lj * std::sqrt((double)g.epsilon[a] * (double)g.epsilon[p])));
            }
        }
        g.exclusion_offsets[a + 1] = (int32_t)g.exclusion_atoms.size();
        for (int t : touched) dist[t] = -1;
    }

    // ---- positions and statistics --------------------------------------------
    g.positions.resize(3 * (size_t)N);
    for (int a = 0; a < N; ++a) {
        g.positions[3 * a] = (float)pos[a].x;
        g.positions[3 * a + 1] = (float)pos[a].y;
        g.positions[3 * a + 2] = (float)pos[a].z;
    }
    {
        Grid fresh(radius + 20.0, 2.0);
        for (int a = 0; a < N; ++a) fresh.insert(a, pos[a]);
        double dmin = cap, dsum = 0.0, tq = 0.0, tm = 0.0;
        for (int a = 0; a < N; ++a) {
            int skip[64];
            int n_skip = 0;
            skip[n_skip++] = a;
            for (int e = g.exclusion_offsets[a]; e < g.exclusion_offsets[a + 1] && n_skip < 64; ++e)
                skip[n_skip++] = g.exclusion_atoms[e];
            const double d = fresh.min_distance(pos[a], pos, skip, n_skip, cap);
            dmin = std::min(dmin, d);
            dsum += d;
            tq += g.charges[a];
            tm += g.masses[a];
        }
        g.stats.n_heavy = H;
        g.stats.n_hydrogen = N - H;
        g.stats.sphere_radius_A = radius;
        g.stats.min_nonbonded_distance_A = dmin;
        g.stats.mean_nearest_nonbonded_A = N ? dsum / N : 0.0;
        g.stats.total_charge = tq;
        g.stats.total_mass = tm;
    }

    PrismMdTopologyV1 &t = g.topo;
    t.n_atoms = N;
    t.positions = g.positions.data();
    t.masses = g.masses.data();
    t.charges = g.charges.data();
    t.sigma = g.sigma.data();
    t.epsilon = g.epsilon.data();
    t.n_bonds = (int32_t)g.bond_r0.size();
    t.bond_atoms = g.bond_atoms.data();
    t.bond_r0 = g.bond_r0.data();
    t.bond_k = g.bond_k.data();
    t.n_angles = (int32_t)g.angle_theta0.size();
    t.angle_atoms = g.angle_atoms.data();
    t.angle_theta0 = g.angle_theta0.data();
    t.angle_k = g.angle_k.data();
    t.n_dihedrals = (int32_t)g.dihedral_k.size();
    t.dihedral_atoms = g.dihedral_atoms.data();
    t.dihedral_periodicity = g.dihedral_periodicity.data();
    t.dihedral_phase = g.dihedral_phase.data();
    t.dihedral_k = g.dihedral_k.data();
    t.dihedral_improper = g.dihedral_improper.data();
    t.exclusion_offsets = g.exclusion_offsets.data();
    t.exclusion_atoms = g.exclusion_atoms.data();
    t.n_exceptions = (int32_t)g.exception_qq.size();
    t.exception_atoms = g.exception_atoms.data();
    t.exception_charge_product = g.exception_qq.data();
    t.exception_sigma = g.exception_sigma.data();
    t.exception_epsilon = g.exception_epsilon.data();
    t.n_clusters = (int32_t)g.cluster_central.size();
    t.cluster_central = g.cluster_central.data();
    t.cluster_hydrogens = g.cluster_hydrogens.data();
    t.cluster_bond_lengths = g.cluster_bond_lengths.data();
    t.cluster_n_hydrogens = g.cluster_n_hydrogens.data();
    t.cluster_inv_mass_central = g.cluster_inv_mass_central.data();
    t.cluster_inv_mass_h = g.cluster_inv_mass_h.data();
    return PRISM_MD_OK;
}

extern "C" int prism_gen_create(const PrismGenParamsV1 *params, PrismGenHandle **out) {
    if (!params || !out) return PRISM_MD_E_INVALID;
    if (params->n_atoms < 3 || params->density <= 0.0 || params->heavy_bond_A <= 0.0 || params->h_bond_A <= 0.0)
        return PRISM_MD_E_INVALID;
    PrismGenHandle *g = new (std::nothrow) PrismGenHandle();
    if (!g) return PRISM_MD_E_NOMEM;
    g->params = *params;
    const int rc = generate(*g);
    if (rc != PRISM_MD_OK) {
        delete g;
        return rc;
    }
    *out = g;
    return PRISM_MD_OK;
}

extern "C" const PrismMdTopologyV1 *prism_gen_topology(const PrismGenHandle *g) { return g ? &g->topo : nullptr; }

extern "C" int prism_gen_stats(const PrismGenHandle *g, PrismGenStatsV1 *stats) {
    if (!g || !stats) return PRISM_MD_E_INVALID;
    *stats = g->stats;
    return PRISM_MD_OK;
}

extern "C" int prism_gen_destroy(PrismGenHandle *g) {
    delete g;
    return PRISM_MD_OK;
}

extern "C" int prism_gen_maxwell_boltzmann(const float *masses, int32_t n, double T, uint64_t seed, float *v) {
    if (!masses || !v || n < 0 || !(T >= 0.0)) return PRISM_MD_E_INVALID;
    Rng rng(seed ^ 0x4D415857454C4C42ull);
    std::vector<double> vd(3 * (size_t)n, 0.0);
    double p[3] = {0, 0, 0}, mt = 0.0;
    for (int i = 0; i < n; ++i) {
        const double m = masses[i];
        if (!(m > 0.0)) continue;
        const double s = std::sqrt(kKcalToAmuA2Ps2 * kBoltzmann * T / m);
        for (int c = 0; c < 3; ++c) {
            vd[3 * i + c] = s * rng.normal();
            p[c] += m * vd[3 * i + c];
        }
        mt += m;
    }
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) v[3 * i + c] = (float)(vd[3 * i + c] - (mt > 0 ? p[c] / mt : 0.0));
    return PRISM_MD_OK;
}
