// engine.hip -- host engine and C ABI of the MI355X classical MD step.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <numeric>
#include <utility>

namespace prism {

// Staging of prism_md_build_pairs_static: the library's own device memory for
// callers that cannot allocate.  Layout: the pair counter (u64) at byte 0, the
// first atoms of STATIC_PAIRS pairs from byte 256, the second atoms after them.
// Each GPU the library runs on holds one copy (32 MiB + 256 B) until the
// process ends.  In host code `scratch` names a host-side stand-in, not a
// device address: its per-device address comes from hipGetSymbolAddress
// (static_pair_staging) and only that address is handed to kernels and copies.
constexpr size_t STATIC_PAIRS = (size_t)1 << 22;  // 4,194,304 pairs
constexpr size_t STATIC_BYTES = 256 + 2 * STATIC_PAIRS * sizeof(int);
static __device__ __align__(256) unsigned char scratch[STATIC_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_addr[MAX_DEVICES] = {};

// The current device's address of `scratch`, looked up once per device.
hipError_t static_pair_staging(void **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_addr[dev] = p;
    }
    *out = g_scratch_addr[dev];
    return hipSuccess;
}

// Clears the staging's counter: a kernel, not a memset node, so the call has
// no memset in it.
__global__ void k_clear_count(unsigned long long *count) { *count = 0ull; }

int pad64(int n) { return (n + BLOCK - 1) / BLOCK * BLOCK; }

// Host Maxwell-Boltzmann draw from the Langevin Philox stream (purpose 3),
// COM velocity removed (role of generate_maxwell_boltzmann_velocities,
// fused_engine.rs:16066-16116; the reference's StdRng stream is not reproduced).
void host_philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = (uint32_t)p1;
        c[2] = n2;
        c[3] = (uint32_t)p0;
    }
}

void maxwell_boltzmann(const PrismMdTopologyV1 *t, const PrismMdParamsV1 &p, std::vector<float> &v) {
    const int n = t->n_atoms;
    v.assign(3 * (size_t)n, 0.0f);
    std::vector<double> vd(3 * (size_t)n);
    double mom[3] = {0, 0, 0}, mt = 0.0;
    for (int i = 0; i < n; ++i) {
        uint32_t c[4] = {(uint32_t)i, 0u, 3u << 8, 0u};
        host_philox(c, (uint32_t)p.seed, p.stream_id);
        const double s = 2.3283064365386963e-10;
        const double a = std::sqrt(-2.0 * std::log(((double)c[0] + 1.0) * s)), b = std::sqrt(-2.0 * std::log(((double)c[2] + 1.0) * s));
        const double z[3] = {a * std::cos(6.283185307179586 * c[1] * s), a * std::sin(6.283185307179586 * c[1] * s),
                             b * std::cos(6.283185307179586 * c[3] * s)};
        const double m = t->masses[i];
        const double sig = std::sqrt(KCAL_TO_AMU_A2_PS2 * KB * p.temperature_K / m);
        for (int k = 0; k < 3; ++k) {
            vd[3 * i + k] = sig * z[k];
            mom[k] += m * vd[3 * i + k];
        }
        mt += m;
    }
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) v[3 * i + k] = (float)(vd[3 * i + k] - mom[k] / mt);
}

}  // namespace

template <class T> T *Engine::dalloc(size_t n) {
    void *p = nullptr;
    if (n == 0) n = 1;
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) {
        alloc_failed_ = true;
        return nullptr;
    }
    allocs_.push_back(p);
    bytes_ += n * sizeof(T);
    hipMemsetAsync(p, 0, n * sizeof(T), stream_);
    return (T *)p;
}

template <class T> int Engine::upload(T *d, const std::vector<T> &h) {
    if (h.empty()) return PRISM_MD_OK;
    return hipcheck(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, stream_), "upload");
}

int Engine::fail(int rc, const std::string &msg) {
    err_ = msg;
    return rc;
}

int Engine::hipcheck(hipError_t e, const char *what) {
    if (e == hipSuccess) return PRISM_MD_OK;
    return fail(PRISM_MD_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

int Engine::d2h(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return PRISM_MD_OK;
    if (int rc = hipcheck(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream_), "device-to-host copy")) return rc;
    return hipcheck(hipStreamSynchronize(stream_), "device-to-host sync");
}

Engine::~Engine() {
    if (stream_) hipStreamSynchronize(stream_);
    for (hipEvent_t e : ev_pool_) hipEventDestroy(e);
    for (auto &g : graphs_) hipGraphExecDestroy(g.second);
    for (void *p : allocs_) hipFree(p);
    if (own_stream_ && stream_) hipStreamDestroy(stream_);
}

static int validate(const PrismMdTopologyV1 *t, const PrismMdParamsV1 &p, std::string &why) {
    const int n = t->n_atoms;
    auto bad = [&](const char *m) { why = m; return PRISM_MD_E_INVALID; };
    if (n <= 0 || !t->positions || !t->masses || !t->charges || !t->sigma || !t->epsilon) return bad("atom arrays");
    if (!t->exclusion_offsets || t->exclusion_offsets[0] != 0) return bad("exclusion offsets");
    for (int i = 0; i < n; ++i) {
        if (!(t->masses[i] > 0.0f) || !std::isfinite(t->positions[3 * i]) || !std::isfinite(t->positions[3 * i + 1]) ||
            !std::isfinite(t->positions[3 * i + 2]))
            return bad("mass or position");
        if (t->exclusion_offsets[i + 1] < t->exclusion_offsets[i]) return bad("exclusion offsets not monotone");
    }
    const int ne = t->exclusion_offsets[n];
    if (ne > 0 && !t->exclusion_atoms) return bad("exclusion atoms");
    for (int e = 0; e < ne; ++e)
        if (t->exclusion_atoms[e] < 0 || t->exclusion_atoms[e] >= n) return bad("exclusion index");
    auto idx_ok = [&](const int32_t *a, int count, int arity) {
        for (int k = 0; k < count * arity; ++k)
            if (a[k] < 0 || a[k] >= n) return false;
        return true;
    };
    if (t->n_bonds > 0 && (!t->bond_atoms || !t->bond_r0 || !t->bond_k || !idx_ok(t->bond_atoms, t->n_bonds, 2))) return bad("bonds");
    if (t->n_angles > 0 && (!t->angle_atoms || !t->angle_theta0 || !t->angle_k || !idx_ok(t->angle_atoms, t->n_angles, 3)))
        return bad("angles");
    if (t->n_dihedrals > 0 && (!t->dihedral_atoms || !t->dihedral_periodicity || !t->dihedral_phase || !t->dihedral_k ||
                               !idx_ok(t->dihedral_atoms, t->n_dihedrals, 4)))
        return bad("dihedrals");
    if (t->n_exceptions > 0 && (!t->exception_atoms || !t->exception_charge_product || !t->exception_sigma ||
                                !t->exception_epsilon || !idx_ok(t->exception_atoms, t->n_exceptions, 2)))
        return bad("exceptions");
    std::vector<char> owned(n, 0);
    for (int c = 0; c < t->n_clusters; ++c) {
        const int nh = t->cluster_n_hydrogens[c], a = t->cluster_central[c];
        if (nh < 1 || nh > 3 || a < 0 || a >= n || owned[a]) return bad("cluster central");
        owned[a] = 1;
        if (!(t->cluster_inv_mass_central[c] > 0.0f) || !(t->cluster_inv_mass_h[c] > 0.0f)) return bad("cluster masses");
        for (int h = 0; h < nh; ++h) {
            const int b = t->cluster_hydrogens[3 * c + h];
            if (b < 0 || b >= n || owned[b] || !(t->cluster_bond_lengths[3 * c + h] > 0.0f)) return bad("cluster hydrogen");
            owned[b] = 1;
        }
    }
    if (!(p.dt_ps > 0.0) || !(p.cutoff_A > 0.0) || !(p.skin_A >= 0.0) || p.rebuild_every < 1 || !(p.temperature_K >= 0.0) ||
        !(p.gamma_ps >= 0.0) || p.constraint_max_iter < 1 || !(p.shake_tol > 0.0) || !(p.rattle_tol > 0.0) ||
        !(p.sort_cell_A == 0.0f || (p.sort_cell_A >= 0.25f && p.sort_cell_A <= 64.0f)) ||
        !(p.chunk_tiles == 0 || (p.chunk_tiles >= CHUNK_MIN && p.chunk_tiles <= CHUNK_MAX)))
        return bad("parameters");
    return PRISM_MD_OK;
}

// ---- arena layout ----------------------------------------------------------
// Every system lives in a slot: fixed index ranges of every arena array sized
// by the slot's capacities.  A slot holds at least one padding atom after its
// last atom, so the CSR end offset of an atom never reaches the next slot.
// The host keeps a mirror of the arena arrays so a slot can be rewritten and
// uploaded alone (pool admission) without touching the other slots.

static PrismMdSlotCapsV1 exact_caps(const PrismMdTopologyV1 *t) {
    PrismMdSlotCapsV1 c{};
    c.atoms = t->n_atoms + 1;
    c.bonds = t->n_bonds;
    c.angles = t->n_angles;
    c.dihedrals = t->n_dihedrals;
    c.exceptions = t->n_exceptions;
    c.exclusions = t->exclusion_offsets[t->n_atoms];
    return c;
}

// Constraint groups of a topology: its clusters, then one per atom outside
// every cluster (the grouping lower() builds).
static int group_count(const PrismMdTopologyV1 *t) {
    std::vector<char> in_cluster(t->n_atoms, 0);
    for (int c = 0; c < t->n_clusters; ++c) {
        in_cluster[t->cluster_central[c]] = 1;
        for (int k = 0; k < t->cluster_n_hydrogens[c]; ++k) in_cluster[t->cluster_hydrogens[3 * c + k]] = 1;
    }
    int n = t->n_clusters;
    for (int i = 0; i < t->n_atoms; ++i) n += in_cluster[i] ? 0 : 1;
    return n;
}

static std::string fits(const PrismMdSlotCapsV1 &c, const PrismMdTopologyV1 *t) {
    if (t->n_atoms >= c.atoms) return "atoms";
    if (t->n_bonds > c.bonds) return "bonds";
    if (t->n_angles > c.angles) return "angles";
    if (t->n_dihedrals > c.dihedrals) return "dihedrals";
    if (t->n_exceptions > c.exceptions) return "exceptions";
    if (t->exclusion_offsets[t->n_atoms] > c.exclusions) return "exclusions";
    return "";
}

int Engine::layout(const std::vector<PrismMdSlotCapsV1> &caps, void *stream, const std::vector<int> &group_waves) {
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0) return fail(PRISM_MD_E_NODEVICE, "no HIP device");
    if (stream) {
        stream_ = (hipStream_t)stream;
    } else {
        if (int rc = hipcheck(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "stream")) return rc;
        own_stream_ = true;
    }
    // the static pair staging of this device, looked up once per device
    if (int rc = hipcheck(static_pair_staging(&static_staging_), "static pair staging address")) return rc;
    {
        int dev = 0, khz = 0;
        if (int rc = hipcheck(hipGetDevice(&dev), "device")) return rc;
        if (int rc = hipcheck(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev), "wall clock rate"))
            return rc;
        ms_per_tick_ = khz > 0 ? 1.0 / (double)khz : 0.0;
    }
    int A = 0, W = 0;
    size_t NB = 0, NA = 0, ND = 0, NE = 0, SK = 0, EX = 0, CS = 0;
    slots_.clear();
    for (const auto &c : caps) {
        if (c.atoms < 2 || c.bonds < 0 || c.angles < 0 || c.dihedrals < 0 || c.exceptions < 0 || c.exclusions < 0)
            return fail(PRISM_MD_E_INVALID, "slot capacities");
        Slot S{};
        S.atom_base = A;
        S.atom_cap = pad64(c.atoms);
        S.wave_base = W;
        // constraint-group waves: exact when the system is known at create
        // (a created engine), else one group per atom at worst (a pool slot)
        S.wave_cap = slots_.size() < group_waves.size() ? std::max(1, group_waves[slots_.size()]) : S.atom_cap / 64;
        S.bond_base = (int)NB; S.bond_cap = c.bonds;
        S.angle_base = (int)NA; S.angle_cap = c.angles;
        S.dih_base = (int)ND; S.dih_cap = c.dihedrals;
        S.exc_base = (int)NE; S.exc_cap = c.exceptions;
        S.skip_base = (int)SK; S.skip_cap = c.exclusions + 2 * c.exceptions;
        S.ex_base = (int)EX; S.ex_cap = c.exclusions;
        S.contrib_base = (int)CS;
        S.contrib_cap = 2 * c.bonds + 3 * c.angles + 4 * c.dihedrals + 2 * c.exceptions;
        S.caps = c;
        A += S.atom_cap;
        W += S.wave_cap;
        NB += S.bond_cap; NA += S.angle_cap; ND += S.dih_cap; NE += S.exc_cap;
        SK += S.skip_cap; EX += S.ex_cap; CS += S.contrib_cap;
        slots_.push_back(S);
    }
    const int B = A / BLOCK;
    const int n_sys = (int)caps.size();
    // host mirror, all padding
    HostArena &h = h_;
    h.x.assign(A, make_float4(0, 0, 0, 0));
    h.v.assign(A, make_float4(0, 0, 0, 0));
    h.mass.assign(A, 1.0f);
    h.q.assign(A, 0.0f);
    h.sig.assign(A, 0.0f);
    h.eps.assign(A, 0.0f);
    h.atom_sys.assign(A, -1);
    h.block_sys.assign(B, -1);
    h.wave_sys.assign(W, -1);
    h.skip_off.assign(A + 1, 0);
    h.ex_off.assign(A + 1, 0);
    h.inc_off.assign(A + 1, 0);
    for (const auto &S : slots_)
        for (int a = S.atom_base; a <= S.atom_base + S.atom_cap && a <= A; ++a) {
            h.skip_off[a] = S.skip_base;
            h.ex_off[a] = S.ex_base;
            h.inc_off[a] = S.contrib_base;
        }
    h.skip_idx.assign(std::max<size_t>(SK, 1), 0);
    h.ex_idx.assign(std::max<size_t>(EX, 1), 0);
    h.groups.assign((size_t)W * 64, Group{{-1, -1, -1, -1}});
    h.gp.assign((size_t)W * 64, GroupParam{});
    h.bonds.assign(NB, Bond{-1, -1, 0.0f, 0.0f});
    h.angles.assign(NA, Angle{-1, -1, -1, 0.0f, 0.0f, {0, 0, 0}});
    h.dih.assign(ND, Dihedral{-1, -1, -1, -1, 0, 1.0f, 0.0f, 0.0f});
    h.exc.assign(NE, Exception14{-1, -1, 0.0f, 0.0f, 0.0f, {0, 0, 0}});
    h.dst_b.assign(2 * NB, 0);
    h.dst_a.assign(3 * NA, 0);
    h.dst_d.assign(4 * ND, 0);
    h.dst_e.assign(2 * NE, 0);
    h.inc_idx.assign(std::max<size_t>(CS, 1), 0);
    sys_.assign(n_sys, HostSys{});
    desc_.assign(n_sys, SysDesc{});
    lat_.assign(n_sys, LatencyLog{});
    for (int k = 0; k < n_sys; ++k) {
        desc_[k].atom_begin = slots_[k].atom_base;
        desc_[k].block_begin = slots_[k].atom_base / BLOCK;
        desc_[k].wave_begin = slots_[k].wave_base;
        desc_[k].bond_begin = slots_[k].bond_base;
        desc_[k].angle_begin = slots_[k].angle_base;
        desc_[k].dih_begin = slots_[k].dih_base;
        desc_[k].exc_begin = slots_[k].exc_base;
        desc_[k].resort_every = 1;
        desc_[k].inv_total_mass = 1.0;
    }
    const size_t n_terms = NB + NA + ND + NE;
    // ---- device allocation
    P_.n_sys = n_sys;
    P_.n_atoms_arena = A;
    P_.n_blocks_arena = B;
    P_.n_waves_arena = W;
    P_.x = dalloc<float4>(A);
    P_.v = dalloc<float4>(A);
    P_.xs = dalloc<float4>(A);
    P_.x_build = dalloc<float4>(A);
    P_.f = dalloc<long long>((size_t)A * 4);
    d_mass_ = dalloc<float>(A); d_q_ = dalloc<float>(A); d_sig_ = dalloc<float>(A); d_eps_ = dalloc<float>(A);
    d_atom_sys_ = dalloc<int>(A);
    P_.perm = dalloc<int>(A);
    P_.slot_of = dalloc<int>(A);
    P_.xq = dalloc<float4>(A);
    P_.lj = dalloc<float2>(A);
    P_.bb_c = dalloc<float4>(B);
    P_.bb_h = dalloc<float4>(B);
    P_.qbox = dalloc<float4>((size_t)B * QBOX_PER_BLOCK * 2);
    d_block_sys_ = dalloc<int>(B);
    P_.row_n = dalloc<int>(B);
    P_.row_j = dalloc<int>((size_t)B * MAX_ROW);
    P_.row_m = dalloc<int>((size_t)B * MAX_ROW);
    P_.masks = dalloc<unsigned long long>((size_t)B * MAX_MASKS * 64);
    P_.pblk = dalloc<int>((size_t)B * MAX_MASKS);
    P_.n_pblk = dalloc<int>(B);
    P_.unit_count = dalloc<int>(B);
    P_.units = dalloc<int>((size_t)B * UNITS_PER_BLOCK);
    P_.unit_off = dalloc<int>(B);
    P_.rows_total = dalloc<int>(B);  // at most one row-building workgroup per block
    P_.unit_f = dalloc<float4>((size_t)B * UNITS_PER_BLOCK * SUB);
    P_.contrib = dalloc<float>(std::max<size_t>(CS, 1) * 3);
    d_inc_off_ = dalloc<int>(A + 1);
    d_dst_b_ = dalloc<int>(2 * NB); d_dst_a_ = dalloc<int>(3 * NA);
    d_dst_d_ = dalloc<int>(4 * ND); d_dst_e_ = dalloc<int>(2 * NE);
    d_inc_idx_ = dalloc<int>(std::max<size_t>(CS, 1));
    term_live_.assign((NB + NA + ND + NE) / TERM_CHUNK + 1, 0);
    d_term_live_ = dalloc<unsigned char>(term_live_.size());
    P_.term_e = dalloc<double>(n_terms);
    P_.atom_e = dalloc<double>(A);
    d_skip_off_ = dalloc<int>(A + 1); d_skip_idx_ = dalloc<int>(h.skip_idx.size());
    d_ex_off_ = dalloc<int>(A + 1); d_ex_idx_ = dalloc<int>(h.ex_idx.size());
    d_groups_ = dalloc<Group>((size_t)W * 64);
    d_gp_ = dalloc<GroupParam>((size_t)W * 64);
    d_wave_sys_ = dalloc<int>(W);
    d_bonds_ = dalloc<Bond>(NB);
    d_angles_ = dalloc<Angle>(NA);
    d_dih_ = dalloc<Dihedral>(ND);
    d_exc_ = dalloc<Exception14>(NE);
    d_desc_ = dalloc<SysDesc>(n_sys);
    P_.sst = dalloc<SysState>(n_sys);
    P_.keys_a = dalloc<unsigned int>(A);
    P_.keys_b = dalloc<unsigned int>(A);
    P_.vals_a = dalloc<int>(A);
    P_.vals_b = dalloc<int>(A);
    P_.com_partial = dalloc<double>((size_t)W * 12);
    P_.work_g = dalloc<double>((size_t)W * 64);
    P_.sweeps = dalloc<int>((size_t)W * 2);
    P_.counters = dalloc<EngineCounters>(1);
    P_.nb_claim = dalloc<unsigned>((size_t)NB_CLAIM_COUNTERS * NB_CLAIM_STRIDE);  // zeroed
    P_.step_ts = dalloc<unsigned long long>(TS_RING);
    P_.hs_group = dalloc<unsigned>((size_t)W / 64 + 2);  // >= second-half workgroups / 128, zeroed
    d_count_ = dalloc<unsigned long long>(1);
    d_red_ = dalloc<double>((size_t)n_sys * RED_STRIDE);
    d_noise_ = dalloc<float>((size_t)2 * A * 3);
    d_vmin_ = dalloc<float4>(A);
    d_fire_ = dalloc<double>((size_t)(A / 256 + 1) * 4);
    if (alloc_failed_) return fail(PRISM_MD_E_NOMEM, "device allocation failed");
    P_.mass = d_mass_; P_.charge = d_q_; P_.sigma = d_sig_; P_.eps = d_eps_; P_.atom_sys = d_atom_sys_;
    P_.block_sys = d_block_sys_;
    P_.skip_off = d_skip_off_; P_.skip_idx = d_skip_idx_; P_.ex_off = d_ex_off_; P_.ex_idx = d_ex_idx_;
    P_.groups = d_groups_; P_.gparams = d_gp_; P_.wave_sys = d_wave_sys_;
    P_.bonds = d_bonds_; P_.angles = d_angles_; P_.dih = d_dih_; P_.exc = d_exc_;
    P_.inc_off = d_inc_off_;
    P_.dst_bond = d_dst_b_; P_.dst_angle = d_dst_a_; P_.dst_dih = d_dst_d_; P_.dst_exc = d_dst_e_;
    P_.inc_idx = d_inc_idx_;
    P_.term_live = d_term_live_;
    // the contribution layout: term-major from 100,000 arena atoms (measured:
    // +3 % at 250k atoms, +7 % for the 1.33 M-atom ensemble, -1 % at 25k)
    P_.term_major = A >= 100000 ? 1 : 0;
    P_.n_bonds = (int)NB; P_.n_angles = (int)NA; P_.n_dih = (int)ND; P_.n_exc = (int)NE;
    P_.sys = d_desc_;
    seam_ = seam_fits(P_);
    // a single system's step builds its rows inside the nonbonded launch (no
    // rows launch per step); a batch keeps the rows kernel, whose re-sort
    // steps (many systems at once) cost far less there
    P_.rows_in_nonbonded = n_sys == 1 ? 1 : 0;
    const std::vector<int> minus1(A, -1);
    int rc = PRISM_MD_OK;
    if ((rc = upload(P_.x, h.x)) || (rc = upload(P_.v, h.v)) || (rc = upload(d_mass_, h.mass)) ||
        (rc = upload(d_q_, h.q)) || (rc = upload(d_sig_, h.sig)) || (rc = upload(d_eps_, h.eps)) ||
        (rc = upload(d_atom_sys_, h.atom_sys)) || (rc = upload(d_block_sys_, h.block_sys)) ||
        (rc = upload(d_skip_off_, h.skip_off)) || (rc = upload(d_skip_idx_, h.skip_idx)) ||
        (rc = upload(d_ex_off_, h.ex_off)) || (rc = upload(d_ex_idx_, h.ex_idx)) || (rc = upload(d_groups_, h.groups)) ||
        (rc = upload(d_gp_, h.gp)) || (rc = upload(d_wave_sys_, h.wave_sys)) || (rc = upload(d_bonds_, h.bonds)) ||
        (rc = upload(d_angles_, h.angles)) || (rc = upload(d_dih_, h.dih)) || (rc = upload(d_exc_, h.exc)) ||
        (rc = upload(d_desc_, desc_)) || (rc = upload(d_inc_off_, h.inc_off)) || (rc = upload(d_dst_b_, h.dst_b)) ||
        (rc = upload(d_dst_a_, h.dst_a)) || (rc = upload(d_dst_d_, h.dst_d)) || (rc = upload(d_dst_e_, h.dst_e)) ||
        (rc = upload(d_inc_idx_, h.inc_idx)) ||
        (rc = upload(P_.perm, minus1)))
        return rc;
    std::vector<SysState> st(n_sys);
    for (auto &S : st) {
        std::memset(&S, 0, sizeof(S));
        S.rebuild_step = -1;
        S.fault_atom = -1;
    }
    return upload(P_.sst, st);
}

// Lower one topology into slot s of the host mirror (padding after it) and
// fill the slot's descriptor.  Nothing is uploaded here.
// FNV-1a 64 over byte ranges: identity of a system for checkpoint restore.
struct Fnv {
    uint64_t h = 1469598103934665603ull;
    void add(const void *p, size_t n) {
        const unsigned char *b = static_cast<const unsigned char *>(p);
        for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    }
    template <class T> void arr(const T *a, size_t n) {
        add(&n, sizeof n);
        if (a && n) add(a, n * sizeof(T));
    }
};

// Everything of a topology except its positions.
static uint64_t topology_digest(const PrismMdTopologyV1 *t) {
    Fnv f;
    const size_t n = (size_t)t->n_atoms;
    f.arr(t->masses, n); f.arr(t->charges, n); f.arr(t->sigma, n); f.arr(t->epsilon, n);
    f.arr(t->bond_atoms, 2 * (size_t)t->n_bonds); f.arr(t->bond_r0, t->n_bonds); f.arr(t->bond_k, t->n_bonds);
    f.arr(t->angle_atoms, 3 * (size_t)t->n_angles); f.arr(t->angle_theta0, t->n_angles); f.arr(t->angle_k, t->n_angles);
    f.arr(t->dihedral_atoms, 4 * (size_t)t->n_dihedrals); f.arr(t->dihedral_periodicity, t->n_dihedrals);
    f.arr(t->dihedral_phase, t->n_dihedrals); f.arr(t->dihedral_k, t->n_dihedrals);
    f.arr(t->exclusion_offsets, n + 1); f.arr(t->exclusion_atoms, (size_t)t->exclusion_offsets[n]);
    f.arr(t->exception_atoms, 2 * (size_t)t->n_exceptions); f.arr(t->exception_charge_product, t->n_exceptions);
    f.arr(t->exception_sigma, t->n_exceptions); f.arr(t->exception_epsilon, t->n_exceptions);
    f.arr(t->cluster_central, t->n_clusters); f.arr(t->cluster_hydrogens, 3 * (size_t)t->n_clusters);
    f.arr(t->cluster_bond_lengths, 3 * (size_t)t->n_clusters); f.arr(t->cluster_n_hydrogens, t->n_clusters);
    f.arr(t->cluster_inv_mass_central, t->n_clusters); f.arr(t->cluster_inv_mass_h, t->n_clusters);
    return f.h;
}

// The run parameters of a descriptor, without its arena placement.
static uint64_t params_digest(const SysDesc &D) {
    SysDesc d = D;
    d.atom_begin = d.block_begin = d.wave_begin = 0;
    d.active = 0;
    d.noise_explicit = 0;
    d.bond_begin = d.angle_begin = d.dih_begin = d.exc_begin = 0;
    Fnv f;
    f.add(&d, sizeof d);
    return f.h;
}

int Engine::lower(int s, const PrismMdTopologyV1 *t, const PrismMdParamsV1 &p) {
    const Slot &L = slots_[s];
    const std::string over = fits(L.caps, t);
    if (!over.empty()) return fail(PRISM_MD_E_CAPACITY, "system does not fit its slot: " + over);
    HostArena &h = h_;
    const int n = t->n_atoms;
    const int o = L.atom_base;
    HostSys H{};
    H.atom_begin = o;
    H.n_atoms = n;
    H.n_slots = L.atom_cap;
    H.block_begin = o / BLOCK;
    H.n_blocks = (n + BLOCK - 1) / BLOCK;
    std::vector<float> v0;
    maxwell_boltzmann(t, p, v0);
    double mt = 0.0;
    for (int i = 0; i < L.atom_cap; ++i) {
        const int a = o + i;
        if (i < n) {
            h.x[a] = make_float4(t->positions[3 * i], t->positions[3 * i + 1], t->positions[3 * i + 2], 0.0f);
            h.v[a] = make_float4(v0[3 * i], v0[3 * i + 1], v0[3 * i + 2], 0.0f);
            h.mass[a] = t->masses[i];
            h.q[a] = t->charges[i];
            h.sig[a] = t->sigma[i];
            h.eps[a] = t->epsilon[i];
            h.atom_sys[a] = s;
            mt += t->masses[i];
        } else {
            h.x[a] = h.v[a] = make_float4(0, 0, 0, 0);
            h.mass[a] = 1.0f;
            h.q[a] = h.sig[a] = h.eps[a] = 0.0f;
            h.atom_sys[a] = -1;
        }
    }
    for (int b = 0; b < L.atom_cap / BLOCK; ++b) h.block_sys[H.block_begin + b] = b < H.n_blocks ? s : -1;
    // skip lists (exclusions U exception partners) and pair-list exclusions
    // (exclusions minus exception partners, fused_engine.rs:11718-11731)
    std::vector<std::vector<int>> expart(n);
    for (int e = 0; e < t->n_exceptions; ++e) {
        const int i = t->exception_atoms[2 * e], j = t->exception_atoms[2 * e + 1];
        expart[i].push_back(j);
        expart[j].push_back(i);
    }
    int sk_n = 0, ex_n = 0;
    for (int i = 0; i < L.atom_cap; ++i) {
        h.skip_off[o + i] = L.skip_base + sk_n;
        h.ex_off[o + i] = L.ex_base + ex_n;
        if (i >= n) continue;
        std::vector<int> sk, ex;
        auto &xp = expart[i];
        std::sort(xp.begin(), xp.end());
        for (int e = t->exclusion_offsets[i]; e < t->exclusion_offsets[i + 1]; ++e) {
            const int j = t->exclusion_atoms[e];
            sk.push_back(j);
            if (!std::binary_search(xp.begin(), xp.end(), j)) ex.push_back(j);
        }
        for (int j : xp) sk.push_back(j);
        std::sort(sk.begin(), sk.end());
        sk.erase(std::unique(sk.begin(), sk.end()), sk.end());
        std::sort(ex.begin(), ex.end());
        ex.erase(std::unique(ex.begin(), ex.end()), ex.end());
        if (sk_n + (int)sk.size() > L.skip_cap || ex_n + (int)ex.size() > L.ex_cap)
            return fail(PRISM_MD_E_CAPACITY, "system does not fit its slot: exclusion lists");
        for (int j : sk) h.skip_idx[L.skip_base + sk_n++] = o + j;
        for (int j : ex) h.ex_idx[L.ex_base + ex_n++] = o + j;
    }
    // constraint groups: clusters, then lone atoms, then padding (topology
    // order keeps a wave's atoms close in memory; ordering the clusters by
    // hydrogen count to give waves one solver shape was measured slower)
    if (group_count(t) > L.wave_cap * 64) return fail(PRISM_MD_E_CAPACITY, "system does not fit its slot: constraint groups");
    std::vector<char> in_cluster(n, 0);
    int ng = 0;
    auto put_group = [&](const Group &G, const GroupParam &Q) {
        h.groups[(size_t)L.wave_base * 64 + ng] = G;
        h.gp[(size_t)L.wave_base * 64 + ng] = Q;
        ++ng;
    };
    for (int c = 0; c < t->n_clusters; ++c) {
        Group G{{-1, -1, -1, -1}};
        GroupParam Q{};
        const int nh = t->cluster_n_hydrogens[c];
        G.atom[0] = o + t->cluster_central[c];
        in_cluster[t->cluster_central[c]] = 1;
        for (int k = 0; k < nh; ++k) {
            G.atom[1 + k] = o + t->cluster_hydrogens[3 * c + k];
            in_cluster[t->cluster_hydrogens[3 * c + k]] = 1;
            Q.d[k] = t->cluster_bond_lengths[3 * c + k];
        }
        Q.n_h = nh;
        Q.inv_mass_c = t->cluster_inv_mass_central[c];
        Q.inv_mass_h = t->cluster_inv_mass_h[c];
        put_group(G, Q);
    }
    for (int i = 0; i < n; ++i)
        if (!in_cluster[i]) {
            GroupParam Q{};
            Q.inv_mass_c = 1.0f / t->masses[i];
            put_group(Group{{o + i, -1, -1, -1}}, Q);
        }
    H.wave_begin = L.wave_base;
    H.n_waves = (ng + 63) / 64;
    while (ng < L.wave_cap * 64) put_group(Group{{-1, -1, -1, -1}}, GroupParam{});
    for (int w = 0; w < L.wave_cap; ++w) h.wave_sys[L.wave_base + w] = w < H.n_waves ? s : -1;
    // terms (padding records have i = -1)
    for (int b = 0; b < L.bond_cap; ++b)
        h.bonds[L.bond_base + b] = b < t->n_bonds
            ? Bond{o + t->bond_atoms[2 * b], o + t->bond_atoms[2 * b + 1], t->bond_r0[b], t->bond_k[b]}
            : Bond{-1, -1, 0.0f, 0.0f};
    for (int a = 0; a < L.angle_cap; ++a)
        h.angles[L.angle_base + a] = a < t->n_angles
            ? Angle{o + t->angle_atoms[3 * a], o + t->angle_atoms[3 * a + 1], o + t->angle_atoms[3 * a + 2],
                    t->angle_theta0[a], t->angle_k[a], {0, 0, 0}}
            : Angle{-1, -1, -1, 0.0f, 0.0f, {0, 0, 0}};
    for (int d = 0; d < L.dih_cap; ++d)
        h.dih[L.dih_base + d] = d < t->n_dihedrals
            ? Dihedral{o + t->dihedral_atoms[4 * d], o + t->dihedral_atoms[4 * d + 1], o + t->dihedral_atoms[4 * d + 2],
                       o + t->dihedral_atoms[4 * d + 3], t->dihedral_periodicity[d],
                       (float)std::cos((double)t->dihedral_phase[d]), (float)std::sin((double)t->dihedral_phase[d]),
                       t->dihedral_k[d]}
            : Dihedral{-1, -1, -1, -1, 0, 1.0f, 0.0f, 0.0f};
    for (int e = 0; e < L.exc_cap; ++e)
        h.exc[L.exc_base + e] = e < t->n_exceptions
            ? Exception14{o + t->exception_atoms[2 * e], o + t->exception_atoms[2 * e + 1],
                          t->exception_charge_product[e], t->exception_sigma[e], t->exception_epsilon[e], {0, 0, 0}}
            : Exception14{-1, -1, 0.0f, 0.0f, 0.0f, {0, 0, 0}};
    // bonded contribution slots, atom-major: the slots of one atom are
    // contiguous and ordered by term kind, term index and position, so the
    // gather adds every atom's bonded forces in one fixed order
    std::vector<int> cnt(L.atom_cap + 1, 0);
    auto count_atoms = [&](const int *atoms, int n_terms, int arity) {
        for (int k = 0; k < n_terms; ++k)
            for (int q = 0; q < arity; ++q) ++cnt[atoms[k * arity + q] + 1];
    };
    count_atoms(t->bond_atoms, t->n_bonds, 2);
    count_atoms(t->angle_atoms, t->n_angles, 3);
    count_atoms(t->dihedral_atoms, t->n_dihedrals, 4);
    count_atoms(t->exception_atoms, t->n_exceptions, 2);
    for (int i = 0; i < L.atom_cap; ++i) cnt[i + 1] += cnt[i];
    for (int i = 0; i < L.atom_cap; ++i) h.inc_off[o + i] = L.contrib_base + cnt[i];
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    // both layouts (the arena uses one): dst_* for atom-major slots, inc_idx
    // lists each atom's term-major slots in the same order
    const long long kind0[4] = {0, 2ll * (long long)h.bonds.size(),
                                2ll * (long long)h.bonds.size() + 3ll * (long long)h.angles.size(),
                                2ll * (long long)h.bonds.size() + 3ll * (long long)h.angles.size() + 4ll * (long long)h.dih.size()};
    auto assign = [&](const int *atoms, int n_terms, int arity, std::vector<int> &dst, int kind, int base) {
        for (int k = 0; k < n_terms; ++k)
            for (int q = 0; q < arity; ++q) {
                const int slot = L.contrib_base + fill[atoms[k * arity + q]]++;
                dst[(size_t)(base + k) * arity + q] = slot;
                h.inc_idx[(size_t)slot] = (int)(kind0[kind] + (long long)(base + k) * arity + q);
            }
    };
    assign(t->bond_atoms, t->n_bonds, 2, h.dst_b, 0, L.bond_base);
    assign(t->angle_atoms, t->n_angles, 3, h.dst_a, 1, L.angle_base);
    assign(t->dihedral_atoms, t->n_dihedrals, 4, h.dst_d, 2, L.dih_base);
    assign(t->exception_atoms, t->n_exceptions, 2, h.dst_e, 3, L.exc_base);
    // reference cell grid from the creation-time coordinates, float
    // arithmetic of fused_engine.rs:12285-12330
    H.x0.assign(t->positions, t->positions + 3 * (size_t)n);
    float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            lo[c] = std::min(lo[c], t->positions[3 * i + c]);
            hi[c] = std::max(hi[c], t->positions[3 * i + c]);
        }
    int dims[3];
    H.grid_ok = 1;
    for (int c = 0; c < 3; ++c) {
        const float extent = (hi[c] - lo[c]) + 2.0f * 10.0f;
        const long v = std::max(1L, (long)std::ceil(extent / 10.0f));
        if (v > 32) H.grid_ok = 0;
        dims[c] = (int)v;
    }
    H.grid = RefGrid{lo[0] - 10.0f, lo[1] - 10.0f, lo[2] - 10.0f, dims[0], dims[1], dims[2]};
    H.dof = std::max(1.0, 3.0 * n - (double)std::accumulate(t->cluster_n_hydrogens, t->cluster_n_hydrogens + t->n_clusters, 0) - 3.0);
    SysDesc D{};
    D.atom_begin = H.atom_begin;
    D.n_atoms = n;
    D.block_begin = H.block_begin;
    D.n_blocks = H.n_blocks;
    D.wave_begin = H.wave_begin;
    D.n_waves = H.n_waves;
    D.active = 1;
    D.rebuild_every = p.rebuild_every;
    D.com_every = p.com_remove_every;
    D.constraint_max_iter = p.constraint_max_iter;
    D.sort_bits = 10;
    D.resort_every = std::max(1, p.resort_every);
    D.sort_cell = p.sort_cell_A > 0.0f ? p.sort_cell_A : 2.0f;
    // tiles per nonbonded unit: short units balance the persistent grid when
    // units are few, long ones save per-unit start-up when they are many
    // (measured best with units claimed at run time: 8 at 25k atoms, 24 at 250k)
    D.chunk = p.chunk_tiles > 0 ? p.chunk_tiles : (t->n_atoms < 100000 ? 8 : 24);
    D.dt = (float)p.dt_ps;
    D.cutoff = (float)p.cutoff_A;
    D.cutoff2 = (float)(p.cutoff_A * p.cutoff_A);
    D.list_cutoff = (float)(p.cutoff_A + p.skin_A);
    D.half_skin2 = (float)(0.25 * p.skin_A * p.skin_A);
    D.inv_rc2 = 1.0f / D.cutoff2;
    D.inv_rc = 1.0f / std::sqrt(D.cutoff2);
    D.reach2 = (D.cutoff + 0.001f) * (D.cutoff + 0.001f);
    D.shake_tol = p.shake_tol;
    D.rattle_tol = p.rattle_tol;
    D.inv_total_mass = 1.0 / mt;
    D.t_start = (float)p.temperature_K;
    D.t_end = (float)p.end_temperature_K;
    D.gamma_base = (float)p.gamma_ps;
    D.gamma_eq = (float)p.equilibration_gamma_ps;
    D.cold_hold_end = p.cold_hold_end;
    D.ramp_end = p.ramp_end;
    D.warm_hold_end = p.warm_hold_end;
    D.ramp_down_end = p.ramp_down_end;
    D.warm_tail = p.warm_tail;
    D.equilibration_steps = p.equilibration_steps;
    D.seed_lo = (uint32_t)p.seed;
    D.stream = p.stream_id;
    D.bond_begin = L.bond_base;
    D.n_bonds = t->n_bonds;
    D.angle_begin = L.angle_base;
    D.n_angles = t->n_angles;
    D.dih_begin = L.dih_base;
    D.n_dih = t->n_dihedrals;
    D.exc_begin = L.exc_base;
    D.n_exc = t->n_exceptions;
    desc_[s] = D;
    lat_[s] = LatencyLog{};
    H.topo_digest = topology_digest(t);
    H.params_digest = params_digest(D);
    sys_[s] = std::move(H);
    return PRISM_MD_OK;
}

template <class T> int Engine::upload_range(T *d, const std::vector<T> &h, size_t begin, size_t count) {
    if (count == 0) return PRISM_MD_OK;
    return hipcheck(hipMemcpyAsync(d + begin, h.data() + begin, count * sizeof(T), hipMemcpyHostToDevice, stream_),
                    "slot upload");
}

// Upload slot s of the host mirror and its descriptor; reset its state.
int Engine::upload_slot(int s) {
    const Slot &L = slots_[s];
    const HostArena &h = h_;
    const size_t a0 = L.atom_base, na = L.atom_cap, b0 = a0 / BLOCK, nb = na / BLOCK;
    const size_t g0 = (size_t)L.wave_base * 64, ng = (size_t)L.wave_cap * 64;
    int rc = PRISM_MD_OK;
    if ((rc = upload_range(P_.x, h.x, a0, na)) || (rc = upload_range(P_.v, h.v, a0, na)) ||
        (rc = upload_range(d_mass_, h.mass, a0, na)) || (rc = upload_range(d_q_, h.q, a0, na)) ||
        (rc = upload_range(d_sig_, h.sig, a0, na)) || (rc = upload_range(d_eps_, h.eps, a0, na)) ||
        (rc = upload_range(d_atom_sys_, h.atom_sys, a0, na)) || (rc = upload_range(d_block_sys_, h.block_sys, b0, nb)) ||
        (rc = upload_range(d_skip_off_, h.skip_off, a0, na)) || (rc = upload_range(d_ex_off_, h.ex_off, a0, na)) ||
        (rc = upload_range(d_inc_off_, h.inc_off, a0, na)) ||
        (rc = upload_range(d_skip_idx_, h.skip_idx, L.skip_base, L.skip_cap)) ||
        (rc = upload_range(d_ex_idx_, h.ex_idx, L.ex_base, L.ex_cap)) || (rc = upload_range(d_groups_, h.groups, g0, ng)) ||
        (rc = upload_range(d_gp_, h.gp, g0, ng)) || (rc = upload_range(d_wave_sys_, h.wave_sys, L.wave_base, L.wave_cap)) ||
        (rc = upload_range(d_bonds_, h.bonds, L.bond_base, L.bond_cap)) ||
        (rc = upload_range(d_angles_, h.angles, L.angle_base, L.angle_cap)) ||
        (rc = upload_range(d_dih_, h.dih, L.dih_base, L.dih_cap)) ||
        (rc = upload_range(d_exc_, h.exc, L.exc_base, L.exc_cap)) ||
        (rc = upload_range(d_dst_b_, h.dst_b, 2 * (size_t)L.bond_base, 2 * (size_t)L.bond_cap)) ||
        (rc = upload_range(d_dst_a_, h.dst_a, 3 * (size_t)L.angle_base, 3 * (size_t)L.angle_cap)) ||
        (rc = upload_range(d_dst_d_, h.dst_d, 4 * (size_t)L.dih_base, 4 * (size_t)L.dih_cap)) ||
        (rc = upload_range(d_dst_e_, h.dst_e, 2 * (size_t)L.exc_base, 2 * (size_t)L.exc_cap)) ||
        (rc = upload_range(d_inc_idx_, h.inc_idx, L.contrib_base, L.contrib_cap)) ||
        (rc = upload_range(d_desc_, desc_, (size_t)s, 1)))
        return rc;
    // no stale slot permutation: every slot of the range maps to no atom until re-sorted
    const std::vector<int> minus1(na, -1);
    if ((rc = hipcheck(hipMemcpyAsync(P_.perm + a0, minus1.data(), na * sizeof(int), hipMemcpyHostToDevice, stream_),
                       "slot perm")))
        return rc;
    SysState S;
    std::memset(&S, 0, sizeof(S));
    S.rebuild_step = -1;
    S.fault_atom = -1;
    if ((rc = hipcheck(hipMemcpyAsync(P_.sst + s, &S, sizeof(S), hipMemcpyHostToDevice, stream_), "slot state")))
        return rc;
    const std::vector<int> zeros((size_t)L.wave_cap * 2, 0);
    if ((rc = hipcheck(hipMemcpyAsync(P_.sweeps + (size_t)L.wave_base * 2, zeros.data(), zeros.size() * sizeof(int),
                                      hipMemcpyHostToDevice, stream_),
                       "slot sweeps")))
        return rc;
    // the work ledger starts at zero with the system
    const std::vector<double> zw(ng, 0.0);
    if ((rc = hipcheck(hipMemcpyAsync(P_.work_g + g0, zw.data(), ng * sizeof(double), hipMemcpyHostToDevice, stream_),
                       "slot work")))
        return rc;
    // the host staging buffers above must outlive the copies
    return hipcheck(hipStreamSynchronize(stream_), "slot upload");
}

int Engine::init(const std::vector<SystemSpec> &systems, void *stream) {
    std::vector<PrismMdSlotCapsV1> caps;
    std::vector<int> waves;
    for (const auto &sp : systems) {
        std::string why;
        if (validate(sp.topo, sp.params, why) != PRISM_MD_OK) return fail(PRISM_MD_E_INVALID, "invalid topology/params: " + why);
        caps.push_back(exact_caps(sp.topo));
        waves.push_back((group_count(sp.topo) + 63) / 64);
    }
    if (int rc = layout(caps, stream, waves)) return rc;
    for (int s = 0; s < (int)systems.size(); ++s) {
        if (int rc = lower(s, systems[s].topo, systems[s].params)) return rc;
        if (int rc = upload_slot(s)) return rc;
    }
    if (int rc = upload_term_live()) return rc;
    launch_kinetic(P_, d_red_, -2, stream_);  // initial COM velocity (zero by construction)
    if (int rc = refresh_all()) return rc;
    return check_status();
}

int Engine::init_pool(const std::vector<PrismMdSlotCapsV1> &caps, void *stream) {
    if (caps.empty()) return fail(PRISM_MD_E_INVALID, "no slots");
    if (int rc = layout(caps, stream)) return rc;
    P_.terms_sparse = 1;
    if (int rc = upload_term_live()) return rc;  // no system yet: every chunk off
    return check_status();
}

// Admit a system into a free slot: lower, upload the slot, rebuild only it.
int Engine::admit(int s, const PrismMdTopologyV1 *t, const PrismMdParamsV1 &p) {
    if (s < 0 || s >= (int)slots_.size()) return fail(PRISM_MD_E_INVALID, "slot index");
    if (desc_[s].active) return fail(PRISM_MD_E_INVALID, "slot is occupied");
    std::string why;
    if (validate(t, p, why) != PRISM_MD_OK) return fail(PRISM_MD_E_INVALID, "invalid topology/params: " + why);
    if (int rc = lower(s, t, p)) {
        desc_[s].active = 0;
        return rc;
    }
    if (int rc = upload_slot(s)) return rc;
    if (int rc = upload_term_live()) return rc;
    launch_kinetic(P_, d_red_, s, stream_);
    if (int rc = refresh(s)) return rc;
    return check_status();
}

int Engine::retire(int s) {
    if (s < 0 || s >= (int)slots_.size()) return fail(PRISM_MD_E_INVALID, "slot index");
    return set_active(s, false);
}

// Rebuild the lists at the current positions and recompute the carried force.
int Engine::refresh_all() { return refresh(-1); }

// Rebuild the lists of system `only` (-1: every system) at the current
// positions with a re-sort, re-claim the unit list and recompute the carried
// force.  Systems not rebuilt keep their rows, so their forces -- and their
// trajectories -- are bitwise unchanged.
int Engine::refresh(int only) {
    static const int ones[2] = {1, 1};  // rebuild_flag, resort_flag
    static_assert(offsetof(SysState, resort_flag) == offsetof(SysState, rebuild_flag) + sizeof(int), "flag layout");
    for (int s = 0; s < (int)sys_.size(); ++s) {
        if (only >= 0 && s != only) continue;
        if (int rc = hipcheck(hipMemcpyAsync((char *)(P_.sst + s) + offsetof(SysState, rebuild_flag), ones, sizeof(ones),
                                             hipMemcpyHostToDevice, stream_),
                              "flag"))
            return rc;
    }
    if (int rc = hipcheck(hipMemcpyAsync((char *)P_.counters + offsetof(EngineCounters, any_rebuild), ones, sizeof(int),
                                         hipMemcpyHostToDevice, stream_),
                          "flag"))
        return rc;
    launch_sort(P_, false, stream_);
    launch_rows(P_, stream_);
    launch_consume(P_, -1, stream_);
    full_force();
    return hipcheck(hipGetLastError(), "refresh");
}

// The step's force evaluation: nonbonded (sampled by the kernel timer),
// bonded, then the fixed-order gather.  (Running the bonded terms on a forked
// second stream beside the nonbonded kernel was measured slower: the fork and
// join cost more than the overlap gains.)
void Engine::launch_forces(bool step, bool sample) {
    sample = sample && 2 * ev_used_ + 1 < ev_pool_.size();
    // a sampled launch carries its own start / stop events
    hipEvent_t start = sample ? ev_pool_[2 * ev_used_] : nullptr, stop = sample ? ev_pool_[2 * ev_used_ + 1] : nullptr;
    if (sample) ++ev_used_;
    if (step) launch_nonbonded_direct(P_, stream_, start, stop);
    else launch_nonbonded(P_, false, stream_, start, stop);
    if (!step) {
        launch_bonded(P_, PRISM_MD_TERM_ALL, false, stream_);
        launch_gather(P_, PRISM_MD_TERM_ALL, false, stream_);
    }
}

void Engine::full_force() {
    launch_forces(false, false);
    ++force_evals_;
}

// Four kernels (five for a batch, whose rows kernel runs every step): the
// bonded terms ride in the sort kernel's launch (spare workgroups after the
// sort and refresh ones), a single system's nonbonded kernel builds the rows
// of a rebuild step ahead of its units, and the nonbonded kernel adds every
// contribution straight into the accumulators the first half zeroed, so no
// bonded or gather kernel runs.
void Engine::launch_step(const float *noise_n, const float *noise_np1, bool sample) {
    launch_first_half(P_, noise_n, stream_);
    launch_sort(P_, true, stream_);
    if (!P_.rows_in_nonbonded) launch_rows(P_, stream_);
    launch_forces(true, sample);
    launch_second_half(P_, noise_np1, stream_);
    ++force_evals_;
}

// The rest of a step whose first half has run: the sort kernel, the rows
// (a batch's rows kernel, or inside a single system's nonbonded launch), the
// forces, then either the seam, which also runs the next step's first half,
// or the second half alone.
void Engine::launch_tail(bool seam, bool sample) {
    launch_sort(P_, true, stream_);
    if (!P_.rows_in_nonbonded) launch_rows(P_, stream_);
    launch_forces(true, sample);
    if (seam) launch_seam(P_, stream_);
    else launch_second_half(P_, nullptr, stream_);
    ++force_evals_;
}

// The graph and eager unit of a run: a whole step, or with the seam a step
// tail that both starts and ends inside a seam.
void Engine::launch_unit() {
    if (seam_) launch_tail(true);
    else launch_step(nullptr, nullptr);
}

// n_steps steps without explicit noise.  With the seam (seam_fits: a
// four-lane arena whose integrator workgroups are all resident at once) each
// step of the call but the last ends in a seam launch, so the call opens with
// one first half and closes with one second half; otherwise every step is the
// five launches of launch_step.  Graphs of graph_steps units replay the bulk.
// With nonbonded sampling on (`sample`, kernel_timing), every
// timing_every-th step runs eagerly with its nonbonded launch bracketed by
// events, and the units between two sampled steps replay as one graph.
int Engine::run_steps(uint64_t n_steps, int graph_steps, bool sample) {
    if (n_steps == 0) return PRISM_MD_OK;
    const uint64_t period = sample && timing_every_ > 0 ? (uint64_t)timing_every_ : 0;
    if (!seam_) {
        if (!period) return replay(n_steps, graph_steps);
        for (uint64_t done = 0; done < n_steps;) {
            launch_step(nullptr, nullptr, true);  // sampled
            const uint64_t rest = std::min<uint64_t>(period - 1, n_steps - done - 1);
            if (rest)
                if (int rc = replay(rest, (int)(period - 1))) return rc;
            done += 1 + rest;
        }
        return PRISM_MD_OK;
    }
    launch_first_half(P_, nullptr, stream_);
    for (uint64_t done = 0; done < n_steps;) {  // done: step tails launched
        const bool sampled = period && done % period == 0;
        if (sampled || done + 1 == n_steps) {
            launch_tail(done + 1 < n_steps, sampled);
            ++done;
            continue;
        }
        // the unsampled tails before the next sampled one or the last
        uint64_t upto = n_steps - 1;
        if (period) upto = std::min(upto, (done / period + 1) * period);
        if (int rc = replay(upto - done, period ? (int)(period - 1) : graph_steps)) return rc;
        done = upto;
    }
    return PRISM_MD_OK;
}

// Turn the device's step-end stamps since the last harvest into per-step
// latencies of every system that advanced.  Systems only stop (limit, fault)
// or start (admission, activation) between host calls, so a system that
// advanced d steps in a window of global steps [lo, hi) ran its first d.  A
// step's latency is the time from the previous step's end to its own: it
// includes whatever the device or the host did between the two (admissions,
// retirements, copies).  A system's first step after admission has no
// previous step of its own and is not counted.  Only the TS_RING most recent
// stamps are kept, so a single call of more steps loses its older ones.
int Engine::harvest_step_ends(const std::vector<SysState> &st) {
    unsigned long long hi = 0;
    if (int rc = d2h(&hi, &P_.counters->step_ends, sizeof(hi))) return rc;
    const unsigned long long lo = ts_harvested_;
    if (hi <= lo) return PRISM_MD_OK;
    // stamps [first, hi) are still in the ring; one before lo gives lo's latency
    const unsigned long long first = std::max<unsigned long long>(lo == 0 ? 0 : lo - 1, hi > TS_RING ? hi - TS_RING : 0);
    std::vector<unsigned long long> ts(hi - first);
    for (unsigned long long k = first; k < hi;) {  // at most two contiguous pieces of the ring
        const unsigned long long idx = k % TS_RING, len = std::min<unsigned long long>(hi - k, TS_RING - idx);
        if (int rc = d2h(ts.data() + (k - first), P_.step_ts + idx, len * sizeof(unsigned long long))) return rc;
        k += len;
    }
    for (size_t s = 0; s < sys_.size(); ++s) {
        LatencyLog &L = lat_[s];
        const unsigned long long now = st[s].step;
        if (!desc_[s].active || now <= L.steps_seen) {
            L.steps_seen = now;
            continue;
        }
        const unsigned long long d = now - L.steps_seen;
        for (unsigned long long k = lo; k < lo + d && k < hi; ++k) {
            const bool first_of_system = L.steps_seen == 0 && k == lo;
            if (first_of_system || k == 0 || k - 1 < first) continue;
            L.ms.push_back((float)((double)(ts[k - first] - ts[k - 1 - first]) * ms_per_tick_));
        }
        L.steps_seen = now;
    }
    ts_harvested_ = hi;
    return PRISM_MD_OK;
}

int Engine::latency(int s, double *p50, double *p99, uint64_t *n, float *samples, size_t cap) {
    if (s < 0 || s >= (int)sys_.size()) return fail(PRISM_MD_E_INVALID, "system index");
    std::vector<SysState> st;  // harvest what the device has done; a latched fault is no error here
    if (int rc = sync_states(st)) return rc;
    const std::vector<float> &v = lat_[s].ms;
    if (n) *n = v.size();
    if (samples) std::copy(v.begin(), v.begin() + std::min(cap, v.size()), samples);
    auto pct = [&](double q) {
        if (v.empty()) return 0.0;
        std::vector<float> w(v);
        const size_t k = std::min(w.size() - 1, (size_t)std::floor(q * (double)(w.size() - 1) + 0.5));
        std::nth_element(w.begin(), w.begin() + (ptrdiff_t)k, w.end());
        return (double)w[k];
    };
    if (p50) *p50 = pct(0.50);
    if (p99) *p99 = pct(0.99);
    return PRISM_MD_OK;
}

int Engine::set_step_limit(int s, uint64_t end_step) {
    if (s < 0 || s >= (int)sys_.size()) return fail(PRISM_MD_E_INVALID, "system index");
    desc_[s].end_step = end_step;
    return hipcheck(hipMemcpyAsync((char *)(d_desc_ + s) + offsetof(SysDesc, end_step), &desc_[s].end_step,
                                   sizeof(desc_[s].end_step), hipMemcpyHostToDevice, stream_),
                    "step limit");
}

// Wait for the stream, read every system's state and harvest the step ends.
int Engine::sync_states(std::vector<SysState> &st) {
    if (int rc = hipcheck(hipStreamSynchronize(stream_), "synchronize")) return rc;
    st.resize(sys_.size());
    if (int rc = d2h(st.data(), P_.sst, st.size() * sizeof(SysState))) return rc;
    return harvest_step_ends(st);
}

int Engine::check_status() {
    std::vector<SysState> st;
    if (int rc = sync_states(st)) return rc;
    for (size_t s = 0; s < st.size(); ++s) {
        const int b = st[s].status;
        // a retired or deactivated system keeps its latched fault for
        // get_stats but no longer fails calls on the systems still running
        if (!b || !desc_[s].active) continue;
        static const char *kinds[] = {"?", "RATTLE after OU_N", "RATTLE before drift", "SHAKE", "RATTLE before OU_NP1",
                                      "final RATTLE", "non-finite state", "coordinate beyond 1000 A"};
        const int fk = st[s].fault_kind >= 0 && st[s].fault_kind <= 7 ? st[s].fault_kind : 0;
        const std::string where = " (system " + std::to_string(s) + ", " + kinds[fk] + ", atom " +
                                  std::to_string(st[s].fault_atom - sys_[s].atom_begin) + ", step " +
                                  std::to_string(st[s].fault_step) + ")";
        // the first fault decides the class (later steps may turn a failed
        // constraint into non-finite values)
        if ((b & ST_CONSTRAINT) && fk >= FAULT_RATTLE_1 && fk <= FAULT_RATTLE_4)
            return fail(PRISM_MD_E_CONSTRAINT, "constraint iteration did not converge" + where);
        if (b & ST_NONFINITE) return fail(PRISM_MD_E_NONFINITE, "non-finite state" + where);
        if (b & ST_DIVERGED) return fail(PRISM_MD_E_DIVERGED, "system diverged: a coordinate beyond 1000 A" + where);
        if (b & ST_CONSTRAINT) return fail(PRISM_MD_E_CONSTRAINT, "constraint iteration did not converge" + where);
        if (b & ST_EXCEPTION_COINCIDENT) return fail(PRISM_MD_E_NONFINITE, "coincident exception atoms" + where);
        if (b & ST_ROW_OVERFLOW) return fail(PRISM_MD_E_CAPACITY, "tile row overflow" + where);
        if (b & ST_MASK_OVERFLOW) return fail(PRISM_MD_E_CAPACITY, "exclusion-mask overflow" + where);
        if (b & ST_BARRIER) return fail(PRISM_MD_E_HIP, "grid barrier timed out (step seam or nonbonded rows)" + where);
        return fail(PRISM_MD_E_INVALID, "device status " + std::to_string(b) + where);
    }
    return PRISM_MD_OK;
}

int Engine::set_state(int s, const float *x, const float *v, int flags) {
    if (s < 0 || s >= (int)sys_.size()) return fail(PRISM_MD_E_INVALID, "system index");
    const HostSys &H = sys_[s];
    const int n = H.n_atoms;
    std::vector<float> hx, hv;
    auto fetch = [&](const float *src, std::vector<float> &dst) -> int {
        dst.resize(3 * (size_t)n);
        if (flags == PRISM_MD_DEVICE_PTR)
            return hipcheck(hipMemcpy(dst.data(), src, dst.size() * sizeof(float), hipMemcpyDeviceToHost), "state fetch");
        std::memcpy(dst.data(), src, dst.size() * sizeof(float));
        return PRISM_MD_OK;
    };
    std::vector<float4> pk(n);
    if (x) {
        if (int rc = fetch(x, hx)) return rc;
        for (int i = 0; i < n; ++i) {
            if (!std::isfinite(hx[3 * i]) || !std::isfinite(hx[3 * i + 1]) || !std::isfinite(hx[3 * i + 2]))
                return fail(PRISM_MD_E_NONFINITE, "non-finite position");
            pk[i] = make_float4(hx[3 * i], hx[3 * i + 1], hx[3 * i + 2], 0.0f);
        }
        if (int rc = hipcheck(hipMemcpyAsync(P_.x + H.atom_begin, pk.data(), n * sizeof(float4), hipMemcpyHostToDevice, stream_),
                              "set x"))
            return rc;
    }
    if (v) {
        if (int rc = fetch(v, hv)) return rc;
        for (int i = 0; i < n; ++i) pk[i] = make_float4(hv[3 * i], hv[3 * i + 1], hv[3 * i + 2], 0.0f);
        if (int rc = hipcheck(hipMemcpyAsync(P_.v + H.atom_begin, pk.data(), n * sizeof(float4), hipMemcpyHostToDevice, stream_),
                              "set v"))
            return rc;
        launch_kinetic(P_, d_red_, s, stream_);
    }
    if (x)
        if (int rc = refresh(s)) return rc;
    return check_status();
}

int Engine::get_state(int s, float *x, float *v, int flags) {
    if (s < 0 || s >= (int)sys_.size()) return fail(PRISM_MD_E_INVALID, "system index");
    const HostSys &H = sys_[s];
    const int n = H.n_atoms;
    if (int rc = hipcheck(hipStreamSynchronize(stream_), "sync")) return rc;
    std::vector<float4> pk(n);
    std::vector<float> flat(3 * (size_t)n);
    for (int which = 0; which < 2; ++which) {
        float *dst = which ? v : x;
        if (!dst) continue;
        if (int rc = d2h(pk.data(), (which ? P_.v : P_.x) + H.atom_begin, n * sizeof(float4))) return rc;
        for (int i = 0; i < n; ++i) {
            flat[3 * i] = pk[i].x;
            flat[3 * i + 1] = pk[i].y;
            flat[3 * i + 2] = pk[i].z;
        }
        if (flags == PRISM_MD_DEVICE_PTR) {
            if (int rc = hipcheck(hipMemcpyAsync(dst, flat.data(), flat.size() * sizeof(float), hipMemcpyHostToDevice, stream_),
                                  "put state"))
                return rc;
            if (int rc = hipcheck(hipStreamSynchronize(stream_), "put state")) return rc;
        } else {
            std::memcpy(dst, flat.data(), flat.size() * sizeof(float));
        }
    }
    return PRISM_MD_OK;
}

int Engine::compute_forces(unsigned mask, int s, float *F_out, int64_t *F_fixed, double *E_terms) {
    if (s < 0 || s >= (int)sys_.size()) return fail(PRISM_MD_E_INVALID, "system index");
    mask &= PRISM_MD_TERM_ALL;
    const bool energy = E_terms != nullptr;
    if (mask & PRISM_MD_TERM_NONBONDED) launch_nonbonded(P_, energy, stream_);
    launch_bonded(P_, mask, energy, stream_);
    launch_gather(P_, mask, energy, stream_);
    if (energy) launch_energy_reduce(P_, stream_);
    if (int rc = hipcheck(hipGetLastError(), "force launch")) return rc;
    if (int rc = check_status()) return rc;
    const HostSys &H = sys_[s];
    const int n = H.n_atoms;
    std::vector<long long> acc((size_t)n * 4);
    if (int rc = d2h(acc.data(), P_.f + (size_t)H.atom_begin * 4, acc.size() * sizeof(long long))) return rc;
    const double inv = 1.0 / (double)(1ll << FORCE_FIXED_BITS);
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            if (F_out) F_out[3 * i + c] = (float)((double)acc[(size_t)i * 4 + c] * inv);
            if (F_fixed) F_fixed[3 * i + c] = acc[(size_t)i * 4 + c];
        }
    if (E_terms) {
        SysState S;
        if (int rc = d2h(&S, P_.sst + s, sizeof(S))) return rc;
        for (int k = 0; k < PRISM_MD_N_TERMS; ++k) E_terms[k] = S.energy[k];
    }
    if (mask != PRISM_MD_TERM_ALL) full_force();  // restore the carried F(x)
    return check_status();
}

// Steps without explicit noise replay captured graphs (run_steps).  With
// nonbonded sampling on (kernel_timing), each sampling period is one eager
// step whose nonbonded launch is bracketed by events, then a graph of the
// period's other steps.  Explicit noise is staged per step, so those steps
// run eagerly, each as its own five launches.
int Engine::step(uint64_t n_steps, const float *noise) {
    if (noise && sys_.size() != 1) return fail(PRISM_MD_E_UNSUPPORTED, "explicit noise needs a single system");
    if (!noise) {
        if (int rc = run_steps(n_steps, SINGLE_GRAPH_STEPS, true)) return rc;
        host_steps_ += n_steps;
        if (int rc = hipcheck(hipGetLastError(), "step launch")) return rc;
        const int rc = check_status();
        harvest_timing();
        return rc;
    }
    const int n = sys_[0].n_atoms;
    const size_t per_half = 3 * (size_t)n;
    for (uint64_t k = 0; k < n_steps; ++k) {
        if (int rc = hipcheck(hipMemcpyAsync(d_noise_, noise + (size_t)k * 2 * per_half, 2 * per_half * sizeof(float),
                                             hipMemcpyHostToDevice, stream_),
                              "noise upload"))
            return rc;
        // with sampling on, every timing_every-th step of the call is sampled
        launch_step(d_noise_, d_noise_ + per_half, timing_every_ > 0 && k % (uint64_t)timing_every_ == 0);
        // the staging buffer is reused next step
        if (int rc = hipcheck(hipStreamSynchronize(stream_), "noise sync")) return rc;
    }
    host_steps_ += n_steps;
    if (int rc = hipcheck(hipGetLastError(), "step launch")) return rc;
    const int rc = check_status();
    harvest_timing();
    return rc;
}

// Durations of the sampled nonbonded launches of the last step() call (the
// stream is synchronised by then).
void Engine::harvest_timing() {
    for (size_t k = 0; k < ev_used_; ++k) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, ev_pool_[2 * k], ev_pool_[2 * k + 1]) == hipSuccess) {
            timed_ms_ += ms;
            ++timed_n_;
        }
    }
    ev_used_ = 0;
}

int Engine::kernel_timing(int every, double *mean_ms, uint64_t *samples) {
    if (every < 0) return fail(PRISM_MD_E_INVALID, "timing interval");
    harvest_timing();
    if (mean_ms) *mean_ms = timed_n_ ? timed_ms_ / (double)timed_n_ : 0.0;
    if (samples) *samples = timed_n_;
    if (every != timing_every_ || every == 0) {
        timed_ms_ = 0.0;
        timed_n_ = 0;
    }
    if (every > 0 && ev_pool_.empty()) {
        ev_pool_.resize(2048);
        for (auto &e : ev_pool_)
            if (int rc = hipcheck(hipEventCreate(&e), "timing event")) return rc;
    }
    timing_every_ = every;
    return PRISM_MD_OK;
}

// n_steps units (launch_unit) as replays of a captured graph of graph_steps
// units (captured once per length and kept), the remainder eagerly.  No event
// records are captured.
int Engine::replay(uint64_t n_steps, int graph_steps) {
    if (graph_steps < 1) return fail(PRISM_MD_E_INVALID, "graph_steps");
    hipGraphExec_t exec = nullptr;
    for (auto &g : graphs_)
        if (g.first == graph_steps) exec = g.second;
    if (!exec && n_steps >= (uint64_t)graph_steps) {
        if (graphs_.size() >= MAX_GRAPHS) {  // keep the cache small: drop the oldest
            hipGraphExecDestroy(graphs_.front().second);
            graphs_.erase(graphs_.begin());
        }
        hipGraph_t g;
        const int every = timing_every_;
        const uint64_t evals = force_evals_;  // capturing evaluates nothing
        timing_every_ = 0;  // no event records inside the captured graph
        if (int rc = hipcheck(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "capture")) {
            timing_every_ = every;
            return rc;
        }
        for (int k = 0; k < graph_steps; ++k) launch_unit();
        const hipError_t ce = hipStreamEndCapture(stream_, &g);
        timing_every_ = every;
        force_evals_ = evals;
        if (int rc = hipcheck(ce, "end capture")) return rc;
        const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (int rc = hipcheck(e, "instantiate")) return rc;
        graphs_.emplace_back(graph_steps, exec);
    }
    uint64_t done = 0;
    if (exec)
        for (; done + graph_steps <= n_steps; done += graph_steps)
            if (int rc = hipcheck(hipGraphLaunch(exec, stream_), "graph launch")) return rc;
    force_evals_ += done;
    for (; done < n_steps; ++done) launch_unit();  // counts its own evaluation
    return PRISM_MD_OK;
}

int Engine::step_graph(uint64_t n_steps, int graph_steps) {
    if (graph_steps < 1) return fail(PRISM_MD_E_INVALID, "graph_steps");
    if (int rc = run_steps(n_steps, graph_steps, false)) return rc;
    host_steps_ += n_steps;
    if (int rc = hipcheck(hipGetLastError(), "graph step")) return rc;
    return check_status();
}

int Engine::energies(int s, double out[3]) {
    double E[PRISM_MD_N_TERMS];
    if (int rc = compute_forces(PRISM_MD_TERM_ALL, s, nullptr, nullptr, E)) return rc;
    double red[RED_STRIDE];
    if (int rc = reduce_system(s, red)) return rc;
    out[0] = E[0] + E[1] + E[2] + E[3] + E[4];
    out[1] = red[0];
    out[2] = red[4];
    return PRISM_MD_OK;
}

// Kinetic energy, COM velocity and cumulative work of system s (FP64, fixed
// order), without touching the carried COM velocity.
int Engine::reduce_system(int s, double red[RED_STRIDE]) {
    if (s < 0 || s >= (int)sys_.size()) return fail(PRISM_MD_E_INVALID, "system index");
    launch_kinetic(P_, d_red_, -1, stream_);
    if (int rc = hipcheck(hipMemcpyAsync(red, d_red_ + (size_t)RED_STRIDE * s, RED_STRIDE * sizeof(double),
                                         hipMemcpyDeviceToHost, stream_),
                          "reduce copy"))
        return rc;
    return hipcheck(hipStreamSynchronize(stream_), "sync");
}

int Engine::temperature(int s, double *T) {
    double red[RED_STRIDE];
    if (int rc = reduce_system(s, red)) return rc;
    *T = 2.0 * red[0] / (sys_[s].dof * KB);
    return PRISM_MD_OK;
}

// The pair list of system s in the given mode, emitted into device staging
// (i_d, j_d, counter count_d, at most cap pairs) and kept sorted on the host.
int Engine::emit_pairs(int s, int mode, int *i_d, int *j_d, unsigned long long *count_d, size_t cap) {
    const HostSys &H = sys_[s];
    hipLaunchKernelGGL(k_clear_count, dim3(1), dim3(1), 0, stream_, count_d);
    launch_emit_pairs(P_, mode, H.grid, i_d, j_d, count_d, (unsigned long long)cap, stream_);
    unsigned long long count = 0;
    if (int rc = d2h(&count, count_d, sizeof(count))) return rc;
    if (count > cap) return fail(PRISM_MD_E_CAPACITY, "pair list of " + std::to_string(count) +
                                                          " pairs exceeds the staging of " + std::to_string(cap));
    std::vector<int> hi(count), hj(count);
    if (int rc = d2h(hi.data(), i_d, count * sizeof(int))) return rc;
    if (int rc = d2h(hj.data(), j_d, count * sizeof(int))) return rc;
    // the reference audits per-atom list saturation (fused_engine.rs:16722-16735)
    std::vector<std::pair<int, int>> pr(count);
    std::vector<int> degree(H.n_atoms, 0);
    for (size_t k = 0; k < count; ++k) {
        pr[k] = {hi[k] - H.atom_begin, hj[k] - H.atom_begin};
        ++degree[pr[k].first];
        ++degree[pr[k].second];
    }
    if (mode == PRISM_MD_PAIRS_REFERENCE)
        for (int d : degree)
            if (d >= 1024) return fail(PRISM_MD_E_CAPACITY, "reference neighbour row saturated (>= 1024)");
    std::sort(pr.begin(), pr.end());
    pairs_i_.resize(count);
    pairs_j_.resize(count);
    for (size_t k = 0; k < count; ++k) {
        pairs_i_[k] = pr[k].first;
        pairs_j_[k] = pr[k].second;
    }
    return PRISM_MD_OK;
}

// Rebuild the lists at the current positions (parity entry points).
int Engine::pairs_prologue(int s, int mode) {
    if (s < 0 || s >= (int)sys_.size()) return fail(PRISM_MD_E_INVALID, "system index");
    if (mode != PRISM_MD_PAIRS_REFERENCE && mode != PRISM_MD_PAIRS_BRUTE) return fail(PRISM_MD_E_INVALID, "pair mode");
    if (mode == PRISM_MD_PAIRS_REFERENCE && !sys_[s].grid_ok)
        return fail(PRISM_MD_E_CAPACITY, "reference cell grid needs more than 32 cells per axis");
    pairs_i_.clear();
    pairs_j_.clear();
    if (int rc = refresh_all()) return rc;
    return check_status();
}

int Engine::build_pairs(int s, int mode) {
    if (int rc = pairs_prologue(s, mode)) return rc;
    // count first, then stage exactly that many pairs in memory of this call
    hipLaunchKernelGGL(k_clear_count, dim3(1), dim3(1), 0, stream_, d_count_);
    launch_emit_pairs(P_, mode, sys_[s].grid, nullptr, nullptr, d_count_, 0ull, stream_);
    unsigned long long count = 0;
    if (int rc = d2h(&count, d_count_, sizeof(count))) return rc;
    int *di = nullptr, *dj = nullptr;
    if (int rc = hipcheck(hipMalloc(&di, std::max<size_t>(1, count) * sizeof(int)), "pair alloc")) return rc;
    if (int rc = hipcheck(hipMalloc(&dj, std::max<size_t>(1, count) * sizeof(int)), "pair alloc")) {
        hipFree(di);
        return rc;
    }
    const int rc = emit_pairs(s, mode, di, dj, d_count_, count);
    hipFree(di);
    hipFree(dj);
    return rc;
}

// The same list staged in the library's static device array: no allocation.
int Engine::build_pairs_static(int s, int mode) {
    if (!static_staging_) return fail(PRISM_MD_E_HIP, "static pair staging not set up");
    if (int rc = pairs_prologue(s, mode)) return rc;
    char *base = static_cast<char *>(static_staging_);
    int *si = reinterpret_cast<int *>(base + 256);
    int *sj = si + STATIC_PAIRS;
    return emit_pairs(s, mode, si, sj, reinterpret_cast<unsigned long long *>(base), STATIC_PAIRS);
}

int Engine::get_pairs(int32_t *i_out, int32_t *j_out, size_t cap, size_t *n_out) {
    if (!n_out) return fail(PRISM_MD_E_INVALID, "n_out");
    *n_out = pairs_i_.size();
    if (cap < pairs_i_.size()) return PRISM_MD_E_CAPACITY;
    if (!pairs_i_.empty() && (!i_out || !j_out)) return fail(PRISM_MD_E_INVALID, "output arrays");
    std::copy(pairs_i_.begin(), pairs_i_.end(), i_out);
    std::copy(pairs_j_.begin(), pairs_j_.end(), j_out);
    return PRISM_MD_OK;
}

int Engine::count_pairs(int s, uint64_t *n) {
    (void)s;
    unsigned long long c = 0;
    hipMemsetAsync(d_count_, 0, sizeof(c), stream_);
    launch_count_pairs(P_, d_count_, stream_);
    if (int rc = d2h(&c, d_count_, sizeof(c))) return rc;
    *n = c;
    return PRISM_MD_OK;
}

int Engine::time_nonbonded(int iters, double *mean_ms) {
    hipEvent_t a, b;
    if (int rc = hipcheck(hipEventCreate(&a), "event")) return rc;
    if (int rc = hipcheck(hipEventCreate(&b), "event")) return rc;
    launch_nonbonded(P_, false, stream_);  // warm
    hipEventRecord(a, stream_);
    for (int k = 0; k < iters; ++k) launch_nonbonded(P_, false, stream_, nullptr, nullptr, false);
    hipEventRecord(b, stream_);
    hipEventSynchronize(b);
    float ms = 0.0f;
    hipEventElapsedTime(&ms, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    *mean_ms = (double)ms / iters;
    full_force();  // restore the carried F(x)
    return check_status();
}

// FIRE minimisation of the potential energy over all systems (unit FIRE time,
// dt 0.002 -> at most 0.02, steps limited to 0.1 A), then every X-H bond set
// to its constraint length.  Velocities of the MD state are untouched.
int Engine::minimize(int max_iter, double f_tol, double *e_final, double *f_max) {
    if (max_iter < 0 || !(f_tol > 0.0)) return fail(PRISM_MD_E_INVALID, "minimize arguments");
    hipMemsetAsync(d_vmin_, 0, sizeof(float4) * P_.n_atoms_arena, stream_);
    const int nb = fire_partial_blocks(P_);
    std::vector<double> part((size_t)nb * 4);
    double dt = 0.002, alpha = 0.1, fmax = 0.0;
    int n_pos = 0;
    for (int it = 0; it < max_iter; ++it) {
        full_force();
        launch_fire_partials(P_, d_vmin_, d_fire_, stream_);
        if (int rc = d2h(part.data(), d_fire_, part.size() * sizeof(double))) return rc;
        double pw = 0.0, vv = 0.0, ff = 0.0, fm2 = 0.0;
        for (int b = 0; b < nb; ++b) {
            pw += part[4 * b];
            vv += part[4 * b + 1];
            ff += part[4 * b + 2];
            fm2 = std::max(fm2, part[4 * b + 3]);
        }
        fmax = std::sqrt(fm2);
        if (!std::isfinite(ff)) return fail(PRISM_MD_E_NONFINITE, "non-finite force during minimisation");
        if (fmax < f_tol) break;
        float cv = 0.0f, cf = 0.0f;
        if (pw > 0.0) {
            cv = (float)(1.0 - alpha);
            cf = (float)(alpha * std::sqrt(vv) / std::sqrt(ff));
            if (++n_pos > 5) {
                dt = std::min(dt * 1.1, 0.02);
                alpha *= 0.99;
            }
        } else {
            n_pos = 0;
            dt *= 0.5;
            alpha = 0.1;
        }
        launch_fire_update(P_, d_vmin_, cv, cf, (float)dt, 0.1f, stream_);
        launch_sort(P_, false, stream_);
        launch_rows(P_, stream_);
        launch_consume(P_, -1, stream_);
    }
    launch_snap_hydrogens(P_, stream_);
    if (int rc = refresh_all()) return rc;
    if (int rc = check_status()) return rc;
    if (e_final) {
        double e[3];
        if (int rc = energies(0, e)) return rc;
        *e_final = e[0];
    }
    if (f_max) *f_max = fmax;
    return PRISM_MD_OK;
}

uint64_t Engine::steps_done(int s) {
    SysState S{};
    d2h(&S, P_.sst + s, sizeof(S));
    return S.step;
}

int Engine::stats(int s, PrismMdStatsV1 *out) {
    if (s < 0 || s >= (int)sys_.size() || !out) return fail(PRISM_MD_E_INVALID, "stats");
    if (int rc = hipcheck(hipStreamSynchronize(stream_), "sync")) return rc;
    SysState S;
    if (int rc = d2h(&S, P_.sst + s, sizeof(S))) return rc;
    const HostSys &H = sys_[s];
    std::vector<int> rn(H.n_blocks);
    if (int rc = d2h(rn.data(), P_.row_n + H.block_begin, rn.size() * sizeof(int))) return rc;
    std::memset(out, 0, sizeof(*out));
    out->steps_done = S.step;
    out->pair_rebuilds = (uint64_t)S.n_rebuilds;
    out->force_evaluations = force_evals_;
    out->n_tiles = std::accumulate(rn.begin(), rn.end(), 0);
    out->n_blocks = H.n_blocks;
    out->status = S.status;
    std::vector<int> sw((size_t)H.n_waves * 2);
    if (!sw.empty())
        if (int rc = d2h(sw.data(), P_.sweeps + (size_t)H.wave_begin * 2, sw.size() * sizeof(int))) return rc;
    for (size_t k = 0; k < sw.size(); k += 2) {
        out->max_shake_sweeps = std::max(out->max_shake_sweeps, sw[k]);
        out->max_rattle_sweeps = std::max(out->max_rattle_sweeps, sw[k + 1]);
    }
    return PRISM_MD_OK;
}

// Live flags of the bonded term chunks: a chunk is computed when one of its
// terms belongs to an active system (the term ranges in each descriptor).
int Engine::upload_term_live() {
    std::fill(term_live_.begin(), term_live_.end(), 0);
    const long long sec[4] = {0, P_.n_bonds, (long long)P_.n_bonds + P_.n_angles,
                              (long long)P_.n_bonds + P_.n_angles + P_.n_dih};
    for (const SysDesc &D : desc_) {
        if (!D.active) continue;
        const int begin[4] = {D.bond_begin, D.angle_begin, D.dih_begin, D.exc_begin};
        const int count[4] = {D.n_bonds, D.n_angles, D.n_dih, D.n_exc};
        for (int k = 0; k < 4; ++k)
            if (count[k] > 0)
                for (long long c = (sec[k] + begin[k]) / TERM_CHUNK; c <= (sec[k] + begin[k] + count[k] - 1) / TERM_CHUNK; ++c)
                    term_live_[(size_t)c] = 1;
    }
    return hipcheck(hipMemcpyAsync(d_term_live_, term_live_.data(), term_live_.size(), hipMemcpyHostToDevice, stream_),
                    "term flags");
}

int Engine::set_active(int s, bool active) {
    if (s < 0 || s >= (int)sys_.size()) return fail(PRISM_MD_E_INVALID, "system index");
    if (sys_[s].n_atoms == 0) return fail(PRISM_MD_E_INVALID, "slot holds no system");
    if ((desc_[s].active != 0) == active) return PRISM_MD_OK;
    desc_[s].active = active ? 1 : 0;
    if (int rc = hipcheck(hipMemcpyAsync(d_desc_ + s, &desc_[s], sizeof(SysDesc), hipMemcpyHostToDevice, stream_),
                          "activate"))
        return rc;
    if (int rc = upload_term_live()) return rc;
    EngineCounters c{};  // source of an asynchronous copy: lives until the synchronisation below
    if (active) {
        // re-sort and rebuild this system, re-claim the unit list, and give it
        // its carried force before its first step
        if (int rc = refresh(s)) return rc;
    } else {
        // drop its units: every active block re-claims, no row changes
        c.any_dirty = 1;
        // the flags only: the step-end count behind them keeps running
        if (int rc = hipcheck(hipMemcpyAsync(P_.counters, &c, offsetof(EngineCounters, seam_gen), hipMemcpyHostToDevice, stream_),
                              "counters"))
            return rc;
        launch_rows(P_, stream_);
        launch_consume(P_, -1, stream_);
    }
    return check_status();
}

// ---- checkpoint -------------------------------------------------------------
// Analogue of capture_process_checkpoint_fields / restore_process_checkpoint_fields
// (fused_engine.rs:24719, :25076): the engine's mutable per-system state.
// Noise is counter-based (Philox keyed by seed and stream, counter = atom and
// step), so the step counter is the whole RNG state.
namespace {
constexpr char CKPT_MAGIC[8] = {'P', 'R', 'I', 'S', 'M', 'C', 'K', '1'};
constexpr uint32_t CKPT_VERSION = 2;  // 2: the cumulative work ledger
struct CheckpointHeader {
    char magic[8];
    uint32_t version, header_bytes;
    int32_t n_atoms, n_blocks;
    uint64_t topo_digest, params_digest, payload_bytes;
    double work;                    // cumulative non-Hamiltonian work, kcal/mol
    SysState state;
};
size_t payload_bytes(int n, int nb) {
    return 3 * (size_t)n * sizeof(float4) + (size_t)nb * BLOCK * sizeof(int) + 2 * (size_t)nb * sizeof(float4);
}
}  // namespace

size_t Engine::checkpoint_bytes(int s) const {
    if (s < 0 || s >= (int)sys_.size()) return 0;
    return sizeof(CheckpointHeader) + payload_bytes(sys_[s].n_atoms, sys_[s].n_blocks);
}

int Engine::checkpoint_save(int s, void *buf, size_t cap, size_t *written) {
    if (s < 0 || s >= (int)sys_.size() || !buf) return fail(PRISM_MD_E_INVALID, "checkpoint: system index or buffer");
    if (!desc_[s].active) return fail(PRISM_MD_E_INVALID, "checkpoint: slot holds no running system");
    const HostSys &H = sys_[s];
    const int n = H.n_atoms, nb = H.n_blocks;
    const size_t need = checkpoint_bytes(s);
    if (cap < need) return fail(PRISM_MD_E_CAPACITY, "checkpoint: buffer smaller than checkpoint_size");
    if (int rc = hipcheck(hipStreamSynchronize(stream_), "checkpoint sync")) return rc;
    CheckpointHeader hd{};
    std::memcpy(hd.magic, CKPT_MAGIC, 8);
    hd.version = CKPT_VERSION;
    hd.header_bytes = sizeof(CheckpointHeader);
    hd.n_atoms = n;
    hd.n_blocks = nb;
    hd.topo_digest = H.topo_digest;
    hd.params_digest = H.params_digest;
    hd.payload_bytes = payload_bytes(n, nb);
    if (int rc = d2h(&hd.state, P_.sst + s, sizeof(SysState))) return rc;
    {
        double red[RED_STRIDE];
        if (int rc = reduce_system(s, red)) return rc;
        hd.work = red[4];
    }
    char *out = static_cast<char *>(buf);
    std::memcpy(out, &hd, sizeof hd);
    char *p = out + sizeof hd;
    for (const float4 *src : {P_.x, P_.v, P_.x_build}) {
        if (int rc = d2h(p, src + H.atom_begin, (size_t)n * sizeof(float4))) return rc;
        p += (size_t)n * sizeof(float4);
    }
    std::vector<int> perm((size_t)nb * BLOCK);
    if (int rc = d2h(perm.data(), P_.perm + (size_t)H.block_begin * BLOCK, perm.size() * sizeof(int))) return rc;
    for (int &a : perm) a = a >= 0 ? a - H.atom_begin : -1;  // position independent
    std::memcpy(p, perm.data(), perm.size() * sizeof(int));
    p += perm.size() * sizeof(int);
    for (const float4 *src : {P_.bb_c, P_.bb_h}) {
        if (int rc = d2h(p, src + H.block_begin, (size_t)nb * sizeof(float4))) return rc;
        p += (size_t)nb * sizeof(float4);
    }
    if (written) *written = need;
    return PRISM_MD_OK;
}

int Engine::checkpoint_restore(int s, const void *buf, size_t len) {
    if (s < 0 || s >= (int)sys_.size() || !buf) return fail(PRISM_MD_E_INVALID, "restore: system index or buffer");
    if (!desc_[s].active) return fail(PRISM_MD_E_INVALID, "restore: slot holds no running system");
    const HostSys &H = sys_[s];
    const int n = H.n_atoms, nb = H.n_blocks;
    CheckpointHeader hd;
    if (len < sizeof hd) return fail(PRISM_MD_E_INVALID, "restore: buffer shorter than a checkpoint header");
    std::memcpy(&hd, buf, sizeof hd);
    if (std::memcmp(hd.magic, CKPT_MAGIC, 8) != 0 || hd.version != CKPT_VERSION || hd.header_bytes != sizeof hd)
        return fail(PRISM_MD_E_INVALID, "restore: not a prism_md checkpoint of this version");
    if (hd.n_atoms != n || hd.n_blocks != nb || hd.topo_digest != H.topo_digest)
        return fail(PRISM_MD_E_INVALID, "restore: checkpoint of a different topology");
    if (hd.params_digest != H.params_digest)
        return fail(PRISM_MD_E_INVALID, "restore: checkpoint taken with different run parameters");
    if (hd.payload_bytes != payload_bytes(n, nb) || len != sizeof hd + hd.payload_bytes)
        return fail(PRISM_MD_E_INVALID, "restore: checkpoint length does not match its header");
    const char *p = static_cast<const char *>(buf) + sizeof hd;
    std::vector<float4> x3(3 * (size_t)n);
    std::memcpy(x3.data(), p, x3.size() * sizeof(float4));
    p += x3.size() * sizeof(float4);
    std::vector<int> perm((size_t)nb * BLOCK);
    std::memcpy(perm.data(), p, perm.size() * sizeof(int));
    p += perm.size() * sizeof(int);
    std::vector<float4> boxes(2 * (size_t)nb);
    std::memcpy(boxes.data(), p, boxes.size() * sizeof(float4));
    std::vector<char> seen(n, 0);
    for (int &a : perm) {  // a permutation of the system's atoms plus padding
        if (a < -1 || a >= n || (a >= 0 && seen[a]++)) return fail(PRISM_MD_E_INVALID, "restore: corrupt slot order");
        if (a >= 0) a += H.atom_begin;
    }
    for (const float4 &q : x3)
        if (!std::isfinite(q.x) || !std::isfinite(q.y) || !std::isfinite(q.z))
            return fail(PRISM_MD_E_NONFINITE, "restore: non-finite state");
    int rc = PRISM_MD_OK;
    const size_t a0 = H.atom_begin;
    if ((rc = hipcheck(hipMemcpyAsync(P_.x + a0, x3.data(), (size_t)n * sizeof(float4), hipMemcpyHostToDevice, stream_), "restore x")) ||
        (rc = hipcheck(hipMemcpyAsync(P_.v + a0, x3.data() + n, (size_t)n * sizeof(float4), hipMemcpyHostToDevice, stream_), "restore v")) ||
        (rc = hipcheck(hipMemcpyAsync(P_.x_build + a0, x3.data() + 2 * (size_t)n, (size_t)n * sizeof(float4), hipMemcpyHostToDevice, stream_), "restore x_build")) ||
        (rc = hipcheck(hipMemcpyAsync(P_.perm + (size_t)H.block_begin * BLOCK, perm.data(), perm.size() * sizeof(int), hipMemcpyHostToDevice, stream_), "restore perm")) ||
        (rc = hipcheck(hipMemcpyAsync(P_.bb_c + H.block_begin, boxes.data(), (size_t)nb * sizeof(float4), hipMemcpyHostToDevice, stream_), "restore boxes")) ||
        (rc = hipcheck(hipMemcpyAsync(P_.bb_h + H.block_begin, boxes.data() + nb, (size_t)nb * sizeof(float4), hipMemcpyHostToDevice, stream_), "restore boxes")))
        return rc;
    // the work ledger: the saved total in the system's first group, zeros elsewhere
    std::vector<double> wg((size_t)H.n_waves * 64, 0.0);
    wg[0] = hd.work;
    if ((rc = hipcheck(hipMemcpyAsync(P_.work_g + (size_t)H.wave_begin * 64, wg.data(), wg.size() * sizeof(double),
                                      hipMemcpyHostToDevice, stream_),
                       "restore work")))
        return rc;
    launch_restore_slots(P_, s, nb * BLOCK, stream_);
    // rows and exclusion masks of the restored blocks from the restored boxes
    // (flags make this system's rows rebuild with masks); every block re-claims
    // its units
    SysState st = hd.state;
    st.rebuild_flag = 1;
    st.resort_flag = 1;
    EngineCounters c{};
    c.any_dirty = 1;
    if ((rc = hipcheck(hipMemcpyAsync(P_.sst + s, &st, sizeof st, hipMemcpyHostToDevice, stream_), "restore state")) ||
        (rc = hipcheck(hipMemcpyAsync(P_.counters, &c, offsetof(EngineCounters, seam_gen), hipMemcpyHostToDevice, stream_),
                       "restore counters")))
        return rc;
    launch_rows(P_, stream_);
    SysState after;
    if ((rc = d2h(&after, P_.sst + s, sizeof after))) return rc;
    if (after.status & (ST_ROW_OVERFLOW | ST_MASK_OVERFLOW)) return fail(PRISM_MD_E_CAPACITY, "restore: tile rows overflow");
    // the saved bookkeeping replaces the rebuild flags; the unit list claimed
    // above stays, so the dirty flag is cleared here -- left set, the next
    // step's row kernel would claim every block's units a second time
    const int clean = 0;
    if ((rc = hipcheck(hipMemcpyAsync(P_.sst + s, &hd.state, sizeof hd.state, hipMemcpyHostToDevice, stream_), "restore state")) ||
        (rc = hipcheck(hipMemcpyAsync(&P_.counters->any_dirty, &clean, sizeof clean, hipMemcpyHostToDevice, stream_), "restore counters")))
        return rc;
    full_force();  // the carried force F(x)
    if ((rc = hipcheck(hipGetLastError(), "restore"))) return rc;
    return check_status();
}

}  // namespace prism

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
struct PrismMdHandle {
    prism::Engine eng;
    std::string err;
};

// Message of the last failed prism_md_create on this thread (the handle is
// gone by then); returned by prism_md_last_error(NULL).
static thread_local std::string g_create_error = "no error";

extern "C" {

int prism_md_abi_version(void) { return PRISM_MD_ABI_VERSION; }

int prism_md_default_params(PrismMdParamsV1 *p) {
    if (!p) return PRISM_MD_E_INVALID;
    std::memset(p, 0, sizeof(*p));
    p->dt_ps = 0.004;
    p->temperature_K = 300.0;
    p->end_temperature_K = 300.0;
    p->gamma_ps = 10.0;  // gamma_base, fused_engine.rs:13314
    p->cutoff_A = 10.0;
    p->skin_A = 2.0;
    p->rebuild_every = 20;
    p->com_remove_every = 100;
    p->seed = 0x50524953ull;
    p->stream_id = 0;
    p->constraint_max_iter = 100;
    p->shake_tol = 1e-10;
    p->rattle_tol = 1e-12;
    p->equilibration_steps = 10000;  // fused_engine.rs:15370-15378
    p->equilibration_gamma_ps = 1000.0;
    p->resort_every = 8;
    p->sort_cell_A = 2.0f;
    return PRISM_MD_OK;
}

int prism_md_create(const PrismMdTopologyV1 *topology, const PrismMdParamsV1 *params, void *hip_stream, PrismMdHandle **out) {
    if (!topology || !params || !out) return PRISM_MD_E_INVALID;
    *out = nullptr;
    PrismMdHandle *h = new (std::nothrow) PrismMdHandle();
    if (!h) return PRISM_MD_E_NOMEM;
    const int rc = h->eng.init({prism::SystemSpec{topology, *params}}, hip_stream);
    if (rc != PRISM_MD_OK) {
        g_create_error = h->eng.last_error();
        delete h;
        return rc;
    }
    *out = h;
    return PRISM_MD_OK;
}

int prism_md_destroy(PrismMdHandle *h) {
    delete h;
    return PRISM_MD_OK;
}

int prism_md_set_state(PrismMdHandle *h, const float *x, const float *v, int flags) {
    return h ? h->eng.set_state(0, x, v, flags) : PRISM_MD_E_INVALID;
}
int prism_md_get_state(PrismMdHandle *h, float *x, float *v, int flags) {
    return h ? h->eng.get_state(0, x, v, flags) : PRISM_MD_E_INVALID;
}
int prism_md_build_pairs(PrismMdHandle *h, int mode) { return h ? h->eng.build_pairs(0, mode) : PRISM_MD_E_INVALID; }
int prism_md_set_step_limit(PrismMdHandle *h, uint64_t end_step) {
    if (!h) return PRISM_MD_E_INVALID;
    if (int rc = h->eng.set_step_limit(0, end_step)) return rc;
    return h->eng.check_status();
}
int prism_md_query_latency(PrismMdHandle *h, double *p50_ms, double *p99_ms, uint64_t *samples) {
    return h ? h->eng.latency(0, p50_ms, p99_ms, samples, nullptr, 0) : PRISM_MD_E_INVALID;
}
int prism_md_build_pairs_static(PrismMdHandle *h, int mode) {
    return h ? h->eng.build_pairs_static(0, mode) : PRISM_MD_E_INVALID;
}
int prism_md_get_pairs(PrismMdHandle *h, int32_t *i_out, int32_t *j_out, size_t cap, size_t *n_out) {
    return h ? h->eng.get_pairs(i_out, j_out, cap, n_out) : PRISM_MD_E_INVALID;
}
int prism_md_compute_forces(PrismMdHandle *h, unsigned term_mask, float *F_out, double *E_terms_out) {
    return h ? h->eng.compute_forces(term_mask, 0, F_out, nullptr, E_terms_out) : PRISM_MD_E_INVALID;
}
int prism_md_compute_forces_fixed(PrismMdHandle *h, unsigned term_mask, int64_t *F_fixed_out) {
    return h ? h->eng.compute_forces(term_mask, 0, nullptr, F_fixed_out, nullptr) : PRISM_MD_E_INVALID;
}
int prism_md_step(PrismMdHandle *h, uint64_t n_steps, const float *noise) {
    return h ? h->eng.step(n_steps, noise) : PRISM_MD_E_INVALID;
}
int prism_md_energies(PrismMdHandle *h, double out[3]) { return h && out ? h->eng.energies(0, out) : PRISM_MD_E_INVALID; }
int prism_md_kinetic_temperature(PrismMdHandle *h, double *T) {
    return h && T ? h->eng.temperature(0, T) : PRISM_MD_E_INVALID;
}
int prism_md_get_stats(PrismMdHandle *h, PrismMdStatsV1 *stats) { return h ? h->eng.stats(0, stats) : PRISM_MD_E_INVALID; }
const char *prism_md_last_error(PrismMdHandle *h) { return h ? h->eng.last_error() : g_create_error.c_str(); }
int prism_md_count_pairs(PrismMdHandle *h, uint64_t *n) { return h && n ? h->eng.count_pairs(0, n) : PRISM_MD_E_INVALID; }
int prism_md_minimize(PrismMdHandle *h, int max_iter, double f_tol, double *e_final, double *f_max) {
    return h ? h->eng.minimize(max_iter, f_tol, e_final, f_max) : PRISM_MD_E_INVALID;
}
int prism_md_kernel_timing(PrismMdHandle *h, int every, double *mean_ms, uint64_t *samples) {
    return h ? h->eng.kernel_timing(every, mean_ms, samples) : PRISM_MD_E_INVALID;
}

// ---- batched ensemble ---------------------------------------------------------
struct PrismMdBatch {
    prism::Engine eng;
};

int prism_md_batch_create(const PrismMdTopologyV1 *const *topologies, const PrismMdParamsV1 *params, int32_t n_systems,
                          void *hip_stream, PrismMdBatch **out) {
    if (!topologies || !params || !out || n_systems < 1) return PRISM_MD_E_INVALID;
    *out = nullptr;
    std::vector<prism::SystemSpec> specs;
    specs.reserve((size_t)n_systems);
    for (int32_t s = 0; s < n_systems; ++s) {
        if (!topologies[s]) return PRISM_MD_E_INVALID;
        specs.push_back(prism::SystemSpec{topologies[s], params[s]});
    }
    PrismMdBatch *b = new (std::nothrow) PrismMdBatch();
    if (!b) return PRISM_MD_E_NOMEM;
    const int rc = b->eng.init(specs, hip_stream);
    if (rc != PRISM_MD_OK) {
        g_create_error = b->eng.last_error();
        delete b;
        return rc;
    }
    *out = b;
    return PRISM_MD_OK;
}

int prism_md_batch_destroy(PrismMdBatch *b) {
    delete b;
    return PRISM_MD_OK;
}

int prism_md_batch_size(PrismMdBatch *b, int32_t *n_systems, int64_t *n_atoms) {
    if (!b) return PRISM_MD_E_INVALID;
    int64_t total = 0;
    for (int s = 0; s < b->eng.n_systems(); ++s)
        if (b->eng.occupied(s)) total += b->eng.n_atoms(s);
    if (n_systems) *n_systems = b->eng.n_systems();
    if (n_atoms) *n_atoms = total;
    return PRISM_MD_OK;
}

int prism_md_batch_step(PrismMdBatch *b, uint64_t n_steps, int32_t graph_steps) {
    if (!b) return PRISM_MD_E_INVALID;
    return graph_steps > 0 ? b->eng.step_graph(n_steps, graph_steps) : b->eng.step(n_steps, nullptr);
}

int prism_md_batch_minimize(PrismMdBatch *b, int max_iter, double f_tol, double *f_max) {
    return b ? b->eng.minimize(max_iter, f_tol, nullptr, f_max) : PRISM_MD_E_INVALID;
}

int prism_md_batch_set_active(PrismMdBatch *b, int32_t system, int active) {
    return b ? b->eng.set_active(system, active != 0) : PRISM_MD_E_INVALID;
}

int prism_md_batch_set_state(PrismMdBatch *b, int32_t system, const float *x, const float *v, int flags) {
    return b ? b->eng.set_state(system, x, v, flags) : PRISM_MD_E_INVALID;
}

int prism_md_batch_get_state(PrismMdBatch *b, int32_t system, float *x, float *v, int flags) {
    return b ? b->eng.get_state(system, x, v, flags) : PRISM_MD_E_INVALID;
}

int prism_md_batch_compute_forces(PrismMdBatch *b, int32_t system, unsigned term_mask, float *F_out,
                                  double *E_terms_out) {
    return b ? b->eng.compute_forces(term_mask, system, F_out, nullptr, E_terms_out) : PRISM_MD_E_INVALID;
}

int prism_md_batch_kinetic_temperature(PrismMdBatch *b, int32_t system, double *T) {
    return b && T ? b->eng.temperature(system, T) : PRISM_MD_E_INVALID;
}
int prism_md_batch_set_step_limit(PrismMdBatch *b, int32_t system, uint64_t end_step) {
    if (!b) return PRISM_MD_E_INVALID;
    if (int rc = b->eng.set_step_limit(system, end_step)) return rc;
    return b->eng.check_status();
}
int prism_md_batch_query_latency(PrismMdBatch *b, int32_t system, double *p50_ms, double *p99_ms, uint64_t *samples) {
    return b ? b->eng.latency(system, p50_ms, p99_ms, samples, nullptr, 0) : PRISM_MD_E_INVALID;
}
int prism_md_batch_latency_samples(PrismMdBatch *b, int32_t system, float *ms, size_t cap, size_t *n) {
    if (!b || !n || (cap && !ms)) return PRISM_MD_E_INVALID;
    uint64_t k = 0;
    if (int rc = b->eng.latency(system, nullptr, nullptr, &k, ms, cap)) return rc;
    *n = (size_t)k;
    return PRISM_MD_OK;
}
int prism_md_batch_energies(PrismMdBatch *b, int32_t system, double out[3]) {
    return b && out ? b->eng.energies(system, out) : PRISM_MD_E_INVALID;
}

int prism_md_batch_get_stats(PrismMdBatch *b, int32_t system, PrismMdStatsV1 *stats) {
    return b ? b->eng.stats(system, stats) : PRISM_MD_E_INVALID;
}

const char *prism_md_batch_last_error(PrismMdBatch *b) { return b ? b->eng.last_error() : g_create_error.c_str(); }

int prism_md_slot_caps_for(const PrismMdTopologyV1 *t, PrismMdSlotCapsV1 *caps) {
    if (!t || !caps || t->n_atoms < 1 || !t->exclusion_offsets) return PRISM_MD_E_INVALID;
    *caps = prism::exact_caps(t);
    return PRISM_MD_OK;
}

int prism_md_pool_create(const PrismMdSlotCapsV1 *caps, int32_t n_slots, void *hip_stream, PrismMdBatch **out) {
    if (!caps || !out || n_slots < 1) return PRISM_MD_E_INVALID;
    *out = nullptr;
    PrismMdBatch *b = new (std::nothrow) PrismMdBatch();
    if (!b) return PRISM_MD_E_NOMEM;
    const int rc = b->eng.init_pool(std::vector<PrismMdSlotCapsV1>(caps, caps + n_slots), hip_stream);
    if (rc != PRISM_MD_OK) {
        g_create_error = b->eng.last_error();
        delete b;
        return rc;
    }
    *out = b;
    return PRISM_MD_OK;
}

int prism_md_batch_admit(PrismMdBatch *b, int32_t slot, const PrismMdTopologyV1 *t, const PrismMdParamsV1 *p) {
    return b && t && p ? b->eng.admit(slot, t, *p) : PRISM_MD_E_INVALID;
}

int prism_md_batch_retire(PrismMdBatch *b, int32_t slot) { return b ? b->eng.retire(slot) : PRISM_MD_E_INVALID; }

int prism_md_batch_slot_info(PrismMdBatch *b, int32_t slot, PrismMdSlotCapsV1 *caps, int32_t *occupied) {
    if (!b || slot < 0 || slot >= b->eng.n_systems()) return PRISM_MD_E_INVALID;
    if (caps) *caps = b->eng.slot_caps(slot);
    if (occupied) *occupied = b->eng.occupied(slot) ? 1 : 0;
    return PRISM_MD_OK;
}

int prism_md_batch_device_bytes(PrismMdBatch *b, uint64_t *bytes) {
    if (!b || !bytes) return PRISM_MD_E_INVALID;
    *bytes = b->eng.device_bytes();
    return PRISM_MD_OK;
}

int prism_md_checkpoint_size(PrismMdHandle *h, size_t *bytes) {
    if (!h || !bytes) return PRISM_MD_E_INVALID;
    *bytes = h->eng.checkpoint_bytes(0);
    return PRISM_MD_OK;
}
int prism_md_checkpoint_save(PrismMdHandle *h, void *buf, size_t cap, size_t *written) {
    return h ? h->eng.checkpoint_save(0, buf, cap, written) : PRISM_MD_E_INVALID;
}
int prism_md_checkpoint_restore(PrismMdHandle *h, const void *buf, size_t len) {
    return h ? h->eng.checkpoint_restore(0, buf, len) : PRISM_MD_E_INVALID;
}
int prism_md_batch_checkpoint_size(PrismMdBatch *b, int32_t system, size_t *bytes) {
    if (!b || !bytes || system < 0 || system >= b->eng.n_systems()) return PRISM_MD_E_INVALID;
    *bytes = b->eng.checkpoint_bytes(system);
    return PRISM_MD_OK;
}
int prism_md_batch_checkpoint_save(PrismMdBatch *b, int32_t system, void *buf, size_t cap, size_t *written) {
    return b ? b->eng.checkpoint_save(system, buf, cap, written) : PRISM_MD_E_INVALID;
}
int prism_md_batch_checkpoint_restore(PrismMdBatch *b, int32_t system, const void *buf, size_t len) {
    return b ? b->eng.checkpoint_restore(system, buf, len) : PRISM_MD_E_INVALID;
}

int prism_md_time_nonbonded(PrismMdHandle *h, int iters, double *mean_ms) {
    return h && mean_ms && iters > 0 ? h->eng.time_nonbonded(iters, mean_ms) : PRISM_MD_E_INVALID;
}

}  // extern "C"
