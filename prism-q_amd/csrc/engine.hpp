// engine.hpp -- host engine of the MI355X classical MD step.
//
// Plays the role of NhsAmberFusedEngine (crates/prism-nhs/src/fused_engine.rs:6554):
// lowers topologies, owns every device buffer, sequences the phase kernels on
// one HIP stream.  One Engine packs S >= 1 independent systems into shared
// arenas (md_layout.hpp); the single-system C API is the S = 1 case.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/prism_md.h"
#include "kernels.hpp"

namespace prism {

struct SystemSpec {
    const PrismMdTopologyV1 *topo;
    PrismMdParamsV1 params;
};

class Engine {
  public:
    Engine() = default;
    ~Engine();
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    // Capacity in atoms / groups / terms is fixed at construction; systems may
    // be (de)activated later through set_active.
    int init(const std::vector<SystemSpec> &systems, void *stream);
    // Empty slots of the given capacities (serving pool); systems enter with
    // admit and leave with retire, without reallocation and without
    // invalidating a captured graph.
    int init_pool(const std::vector<PrismMdSlotCapsV1> &caps, void *stream);
    int admit(int s, const PrismMdTopologyV1 *t, const PrismMdParamsV1 &p);
    int retire(int s);
    bool occupied(int s) const { return desc_[s].active != 0; }
    const PrismMdSlotCapsV1 &slot_caps(int s) const { return slots_[s].caps; }
    size_t device_bytes() const { return bytes_; }

    int n_systems() const { return (int)sys_.size(); }
    int n_atoms(int s) const { return sys_[s].n_atoms; }

    int set_state(int s, const float *x, const float *v, int flags);
    int get_state(int s, float *x, float *v, int flags);
    int compute_forces(unsigned mask, int s, float *F_out, int64_t *F_fixed, double *E_terms);
    int step(uint64_t n_steps, const float *noise);  // noise: single-system explicit array or NULL
    int energies(int s, double out[3]);   // potential, kinetic, cumulative work
    int temperature(int s, double *T);    // kinetic temperature over the system's degrees of freedom
    int build_pairs(int s, int mode);
    // per-step latency of system s from the device's step-end stamps (ms);
    // samples (cap entries at most) may be NULL
    int latency(int s, double *p50, double *p99, uint64_t *n, float *samples, size_t cap);
    // system s stops advancing when its step count reaches end_step (0: no limit)
    int set_step_limit(int s, uint64_t end_step);
    // build_pairs staged in the library's static device array (no allocation)
    int build_pairs_static(int s, int mode);
    int get_pairs(int32_t *i_out, int32_t *j_out, size_t cap, size_t *n_out);
    int count_pairs(int s, uint64_t *n);
    int time_nonbonded(int iters, double *mean_ms);
    // sample the nonbonded launch of every `every`-th step with HIP events
    // (0 disables); returns the mean duration over the samples so far
    int kernel_timing(int every, double *mean_ms, uint64_t *samples);
    int minimize(int max_iter, double f_tol, double *e_final, double *f_max);
    int stats(int s, PrismMdStatsV1 *out);
    int set_active(int s, bool active);
    // Dynamic state of system s as a self-describing blob (positions,
    // velocities, build-time positions, sorted order, block boxes, step and
    // rebuild bookkeeping); restoring it into an engine holding the same
    // topology and parameters continues the trajectory bitwise.
    size_t checkpoint_bytes(int s) const;
    int checkpoint_save(int s, void *buf, size_t cap, size_t *written);
    int checkpoint_restore(int s, const void *buf, size_t len);
    int check_status();           // synchronise, read device status, map to rc
    const char *last_error() const { return err_.c_str(); }
    hipStream_t stream() const { return stream_; }
    uint64_t steps_done(int s);

    // graph-captured stepping (batch path); n_steps must be a multiple of
    // graph_steps once the graph exists
    int step_graph(uint64_t n_steps, int graph_steps);

  private:
    int fail(int rc, const std::string &msg);
    int hipcheck(hipError_t e, const char *what);
    int d2h(void *dst, const void *src, size_t bytes);  // stream-ordered copy + synchronise
    int reduce_system(int s, double red[RED_STRIDE]);
    int pairs_prologue(int s, int mode);
    int harvest_step_ends(const std::vector<SysState> &st);
    int sync_states(std::vector<SysState> &st);
    struct LatencyLog {
        unsigned long long steps_seen = 0;  // the system's step count at the last harvest
        std::vector<float> ms;              // per-step latencies since admission
    };
    std::vector<LatencyLog> lat_;
    unsigned long long ts_harvested_ = 0;   // step ends already turned into latencies
    double ms_per_tick_ = 0.0;              // device wall clock period
    int emit_pairs(int s, int mode, int *i_d, int *j_d, unsigned long long *count_d, size_t cap);
    int refresh_all();            // rebuild lists + forces at current positions
    int refresh(int only);        // the same for one system (-1: all)
    int layout(const std::vector<PrismMdSlotCapsV1> &caps, void *stream, const std::vector<int> &group_waves = {});
    int lower(int s, const PrismMdTopologyV1 *t, const PrismMdParamsV1 &p);
    int upload_slot(int s);
    template <class T> int upload_range(T *d, const std::vector<T> &h, size_t begin, size_t count);
    void launch_step(const float *noise_n, const float *noise_np1, bool sample = false);
    void launch_tail(bool seam, bool sample = false);
    void launch_unit();
    int run_steps(uint64_t n_steps, int graph_steps, bool sample);
    bool seam_ = false;           // fused step seams (launch_seam) usable for this arena
    void full_force();
    void harvest_timing();

    struct HostSys {
        int atom_begin, n_atoms, n_slots;
        int block_begin, n_blocks;
        int wave_begin, n_waves;
        std::vector<float> x0;    // creation-time positions (reference grid)
        RefGrid grid;
        int grid_ok;
        double dof;
        uint64_t topo_digest, params_digest;  // identity checked by checkpoint_restore
    };
    std::vector<HostSys> sys_;
    std::vector<SysDesc> desc_;
    // index ranges of one slot in every arena array
    struct Slot {
        int atom_base, atom_cap, wave_base, wave_cap;
        int bond_base, bond_cap, angle_base, angle_cap, dih_base, dih_cap, exc_base, exc_cap;
        int skip_base, skip_cap, ex_base, ex_cap, contrib_base, contrib_cap;
        PrismMdSlotCapsV1 caps;
    };
    std::vector<Slot> slots_;
    // host mirror of the arena arrays written by lower()
    struct HostArena {
        std::vector<float4> x, v;
        std::vector<float> mass, q, sig, eps;
        std::vector<int> atom_sys, block_sys, wave_sys;
        std::vector<int> skip_off, skip_idx, ex_off, ex_idx, inc_off;
        std::vector<Group> groups;
        std::vector<GroupParam> gp;
        std::vector<Bond> bonds;
        std::vector<Angle> angles;
        std::vector<Dihedral> dih;
        std::vector<Exception14> exc;
        std::vector<int> dst_b, dst_a, dst_d, dst_e, inc_idx;
    } h_;
    float *d_mass_ = nullptr, *d_q_ = nullptr, *d_sig_ = nullptr, *d_eps_ = nullptr;
    int *d_atom_sys_ = nullptr, *d_block_sys_ = nullptr, *d_wave_sys_ = nullptr;
    int *d_skip_off_ = nullptr, *d_skip_idx_ = nullptr, *d_ex_off_ = nullptr, *d_ex_idx_ = nullptr, *d_inc_off_ = nullptr;
    int *d_dst_b_ = nullptr, *d_dst_a_ = nullptr, *d_dst_d_ = nullptr, *d_dst_e_ = nullptr, *d_inc_idx_ = nullptr;
    unsigned char *d_term_live_ = nullptr;
    std::vector<unsigned char> term_live_;  // host copy of P_.term_live
    int upload_term_live();
    Group *d_groups_ = nullptr;
    GroupParam *d_gp_ = nullptr;
    Bond *d_bonds_ = nullptr;
    Angle *d_angles_ = nullptr;
    Dihedral *d_dih_ = nullptr;
    Exception14 *d_exc_ = nullptr;
    size_t bytes_ = 0;
    DevPtrs P_{};
    hipStream_t stream_ = nullptr;
    bool own_stream_ = false;
    // nonbonded, bonded, gather; or the step's direct nonbonded only; sample:
    // bracket the nonbonded launch with the next pair of timing events
    void launch_forces(bool step, bool sample);
    std::vector<void *> allocs_;
    int timing_every_ = 0;        // sampled nonbonded timing (kernel_timing)
    std::vector<hipEvent_t> ev_pool_;  // pairs: [2k] before, [2k+1] after
    size_t ev_used_ = 0;
    uint64_t timed_n_ = 0;
    double timed_ms_ = 0.0;
    bool alloc_failed_ = false;
    std::string err_;
    // host-side pair list of the last build_pairs
    std::vector<int32_t> pairs_i_, pairs_j_;
    // device scratch for diagnostics
    unsigned long long *d_count_ = nullptr;
    double *d_red_ = nullptr;
    float *d_noise_ = nullptr;    // explicit noise staging [2][A][3]
    float4 *d_vmin_ = nullptr;    // FIRE velocities
    double *d_fire_ = nullptr;    // FIRE partials
    SysDesc *d_desc_ = nullptr;
    void *static_staging_ = nullptr;  // this device's address of the static pair staging
    // captured step graphs by length (at most MAX_GRAPHS kept)
    std::vector<std::pair<int, hipGraphExec_t>> graphs_;
    static constexpr size_t MAX_GRAPHS = 4;
    static constexpr int SINGLE_GRAPH_STEPS = 20;  // graph length of prism_md_step
    int replay(uint64_t n_steps, int graph_steps);
    uint64_t host_steps_ = 0;
    uint64_t force_evals_ = 0;

    template <class T> T *dalloc(size_t n);
    template <class T> int upload(T *d, const std::vector<T> &h);
};

}  // namespace prism
