// Kick-drift-kick leapfrog of test particles through a time-dependent
// potential, written for capture in a HIP graph.
//
// The time dependence (Evolving snapshot weights, Shifted centres, Scaled
// factors) is tabulated per step by the host into a schedule: for step s the
// potential at time t0 + (s+1) dt is the list of leaves d_sched[s][0..count[s])
// -- a leaf index into d_leaves and the weight, scale and centre of that leaf
// at that time.  The step counter lives in device memory: the j-th step
// kernel of a call runs step counter + j, and one one-thread kernel after the
// call's last step advances the counter by the call's step count, so a graph
// captured once replays step after step with the time moving on and no host
// work.  (A counter kernel after every step would put a dependency gap between
// launches, about 4.7 us, into every step.)
//
// One launch per step does, per particle, v += dt/2 a; x += dt v;
// a = F(x, t_{s+1}); v += dt/2 a -- one read and one write of (x, v, a).
// No reference counterpart: the reference provides evaluation only and
// leaves integration to callers (gpu_potential.py:826-935).
#include "leaf_dev.h"

namespace agx {

__global__ __launch_bounds__(kBlock) void leapfrog_kernel(
    const agx_component *__restrict__ leaves, int n_leaves, const agx_sched_entry *__restrict__ sched,
    const int *__restrict__ count, int max_active, int64_t n_sched, const int *__restrict__ d_step,
    int j, double *__restrict__ xs, double *__restrict__ vs, double *__restrict__ as, int64_t n,
    double dt)
{
    const int64_t s = (int64_t)__builtin_amdgcn_readfirstlane(cldi(d_step, 0)) + j;
    if (s < 0 || s >= n_sched) return;          // past the schedule: particles stay put
    int cnt = cldi(count, (int)s);          // s < n_sched: a schedule row
    if (cnt > max_active) cnt = max_active;
    const double hdt = 0.5 * dt;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        double x[3], v[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            v[k] = fma(hdt, as[3 * i + k], vs[3 * i + k]);
            x[k] = fma(dt, v[k], xs[3 * i + k]);
        }
        double P = 0.0, g[3] = {0.0, 0.0, 0.0}, h[6];
        for (int e = 0; e < cnt; ++e) {
            const agx_sched_entry E = load_desc<agx_sched_entry>(sched + (int64_t)s * max_active + e);
            if (E.leaf < 0 || E.leaf >= n_leaves) continue;
            const agx_component C = load_desc<agx_component>(leaves + E.leaf);
            add_leaf<ORD_GRAD>(C.type, C.p, C.table, E.ampl, E.scale, E.center, x[0], x[1], x[2], P,
                               g, h);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            xs[3 * i + k] = x[k];
            vs[3 * i + k] = fma(-hdt, g[k], v[k]);
            as[3 * i + k] = -g[k];
        }
    }
}

__global__ void leapfrog_advance(int *d_step, int by)
{
    if (threadIdx.x == 0) d_step[0] = d_step[0] + by;
}

}  // namespace agx

using namespace agx;

extern "C" int agx_leapfrog(const agx_component *d_leaves, int32_t n_leaves,
                            const agx_sched_entry *d_sched, const int32_t *d_count,
                            int32_t max_active, int64_t n_sched, int32_t *d_step, double *d_x,
                            double *d_v, double *d_a, int64_t n, double dt, int32_t n_steps,
                            void *stream)
{
    if (n < 0 || n_steps < 0 || n_leaves < 0 || max_active < 1 || n_sched < 0) return AGX_E_SIZE;
    if (n_steps == 0 || n == 0) return AGX_OK;
    if (!d_leaves || !d_sched || !d_count || !d_step || !d_x || !d_v || !d_a) return AGX_E_NULL;
    hipStream_t s = (hipStream_t)stream;
    const unsigned grid = grid_for(n);
    for (int32_t k = 0; k < n_steps; ++k) {
        hipLaunchKernelGGL(leapfrog_kernel, dim3(grid), dim3(kBlock), 0, s, d_leaves,
                           (int)n_leaves, d_sched, (const int *)d_count, (int)max_active, n_sched,
                           (const int *)d_step, (int)k, d_x, d_v, d_a, n, dt);
    }
    hipLaunchKernelGGL(leapfrog_advance, dim3(1), dim3(64), 0, s, (int *)d_step, (int)n_steps);
    return hipGetLastError() == hipSuccess ? AGX_OK : AGX_E_LAUNCH;
}
