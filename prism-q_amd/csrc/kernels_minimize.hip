// kernels_minimize.hip -- FIRE energy minimisation (Bitzek et al., PRL 97,
// 170201 (2006)) on the engine's own force kernels, and the final projection
// of every X-H bond onto its constraint length.
//
// No direct counterpart in the reference, which relaxes unminimised inputs
// with an equilibration friction ramp instead (protocol_director.cu:108-122).
#include "device_common.hpp"

namespace prism {

namespace {

constexpr int MT = 256;

// Per-workgroup FP64 partials of F.v, v.v, F.F and max |F|^2 over real atoms,
// summed afterwards in workgroup order on the host (deterministic).
__global__ __launch_bounds__(MT) void k_fire_partials(DevPtrs P, const float4 *vmin, double *part) {
    const int a = blockIdx.x * MT + threadIdx.x;
    double q[4] = {0.0, 0.0, 0.0, 0.0};
    if (a < P.n_atoms_arena && P.atom_sys[a] >= 0 && P.sys[P.atom_sys[a]].active) {
        const long long *fa = P.f + (size_t)a * 4;
        const double s = 1.0 / (double)(1ll << FORCE_FIXED_BITS);
        const double fx = fa[0] * s, fy = fa[1] * s, fz = fa[2] * s;
        const float4 v = vmin[a];
        q[0] = fx * v.x + fy * v.y + fz * v.z;
        q[1] = (double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z;
        q[2] = fx * fx + fy * fy + fz * fz;
        q[3] = q[2];
    }
    for (int c = 0; c < 3; ++c) q[c] = wave_sum(q[c]);
    for (int o = 32; o > 0; o >>= 1) q[3] = fmax(q[3], __shfl_xor(q[3], o));
    __shared__ double red[MT / 64][4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0)
        for (int c = 0; c < 4; ++c) red[w][c] = q[c];
    __syncthreads();
    if (threadIdx.x == 0) {
        double r[4] = {0.0, 0.0, 0.0, 0.0};
        for (int k = 0; k < MT / 64; ++k) {
            for (int c = 0; c < 3; ++c) r[c] += red[k][c];
            r[3] = fmax(r[3], red[k][3]);
        }
        for (int c = 0; c < 4; ++c) part[blockIdx.x * 4 + c] = r[c];
    }
}

// v <- cv v + cf F (the FIRE mixing; cv = cf = 0 stops), v += dt F/m,
// x += dt v limited to max_step per atom; flags a rebuild when an atom leaves
// its skin sphere.
__global__ __launch_bounds__(MT) void k_fire_update(DevPtrs P, float4 *vmin, float cv, float cf, float dt,
                                                     float max_step) {
    const int a = blockIdx.x * MT + threadIdx.x;
    if (a >= P.n_atoms_arena) return;
    const int s = P.atom_sys[a];
    if (s < 0 || !P.sys[s].active) return;
    const long long *fa = P.f + (size_t)a * 4;
    const float fx = from_fixed(fa[0], FORCE_FIXED_BITS), fy = from_fixed(fa[1], FORCE_FIXED_BITS),
                fz = from_fixed(fa[2], FORCE_FIXED_BITS);
    float4 v = vmin[a];
    const float im = dt / P.mass[a];
    v.x = cv * v.x + cf * fx + im * fx;
    v.y = cv * v.y + cf * fy + im * fy;
    v.z = cv * v.z + cf * fz + im * fz;
    float sx = dt * v.x, sy = dt * v.y, sz = dt * v.z;
    const float sn = sqrtf(sx * sx + sy * sy + sz * sz);
    if (sn > max_step) {
        const float k = max_step / sn;
        sx *= k; sy *= k; sz *= k;
    }
    vmin[a] = v;
    float4 x = P.x[a];
    x.x += sx; x.y += sy; x.z += sz;
    P.x[a] = x;
    *(float3 *)&P.xq[P.slot_of[a]] = make_float3(x.x, x.y, x.z);
    const float4 b = P.x_build[a];
    const float dx = x.x - b.x, dy = x.y - b.y, dz = x.z - b.z;
    if (dx * dx + dy * dy + dz * dz > P.sys[s].half_skin2) {
        P.sst[s].rebuild_flag = 1;
        P.counters->any_rebuild = 1;
    }
}

// Put every hydrogen at exactly its constraint length from its heavy atom,
// along the current bond direction (FP64).
__global__ __launch_bounds__(MT) void k_snap_hydrogens(DevPtrs P) {
    const int g = blockIdx.x * MT + threadIdx.x;
    if (g >= P.n_waves_arena * 64) return;
    const Group G = P.groups[g];
    if (G.atom[0] < 0 || !P.sys[P.atom_sys[G.atom[0]]].active) return;
    const GroupParam gp = P.gparams[g];
    const float4 c = P.x[G.atom[0]];
    for (int h = 0; h < gp.n_h; ++h) {
        const int a = G.atom[1 + h];
        const float4 p = P.x[a];
        const double rx = (double)p.x - c.x, ry = (double)p.y - c.y, rz = (double)p.z - c.z;
        const double k = gp.d[h] / sqrt(rx * rx + ry * ry + rz * rz);
        P.x[a] = make_float4((float)(c.x + k * rx), (float)(c.y + k * ry), (float)(c.z + k * rz), 0.0f);
    }
}

}  // namespace

int fire_partial_blocks(const DevPtrs &P) { return (P.n_atoms_arena + MT - 1) / MT; }

void launch_fire_partials(const DevPtrs &P, const float4 *vmin, double *part, hipStream_t st) {
    hipLaunchKernelGGL(k_fire_partials, dim3(fire_partial_blocks(P)), dim3(MT), 0, st, P, vmin, part);
}

void launch_fire_update(const DevPtrs &P, float4 *vmin, float cv, float cf, float dt, float max_step, hipStream_t st) {
    hipLaunchKernelGGL(k_fire_update, dim3(fire_partial_blocks(P)), dim3(MT), 0, st, P, vmin, cv, cf, dt, max_step);
}

void launch_snap_hydrogens(const DevPtrs &P, hipStream_t st) {
    const int n = P.n_waves_arena * 64;
    hipLaunchKernelGGL(k_snap_hydrogens, dim3((n + MT - 1) / MT), dim3(MT), 0, st, P);
}

}  // namespace prism
