// gfx950 kernels of the reference's per-interface and boundary helpers
// (fluid.py:105-331) as stand-alone elementwise launches; the fused step
// evaluates the same device arithmetic (tf_device.h) inside its sweeps.
#include "tf_internal.h"

namespace tf {

// Maximum of v over a 256-thread block, valid in thread 0.
__device__ __forceinline__ uint32_t helper_block_umax(uint32_t v) {
    __shared__ uint32_t part[4];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    return max(max(part[0], part[1]), max(part[2], part[3]));
}

static inline dim3 grid1(int64_t n, int block = 256) { return dim3((unsigned)((n + block - 1) / block)); }

// ---------------------------------------------------------------- per-interface helpers
// The reference's helpers of one sweep as stand-alone elementwise kernels on
// [5][n] primitive arrays (left and right states of n interfaces, or n
// cells), with the arithmetic of the fused step: fluid.py:105-117 (sound
// speed), :119-160 (Davis wave speeds), :162-195 (their maximum), :206-259
// (Euler flux), :261-305 (HLL flux). D is the flux direction 1..3.
template <int D>
__device__ __forceinline__ Q q_from_prim(const float *__restrict__ w, int64_t n, int64_t i, const Gas &g) {
    float v[5], U[5], rinv;
#pragma unroll
    for (int f = 0; f < 5; ++f) v[f] = w[f * n + i];
    prim_to_cons(v, g, U, rinv);
    const float vel[3] = {v[1], v[2], v[3]};
    return make_q<D>(U, vel, v[4], frcp_nr(v[0]), g);
}

__global__ __launch_bounds__(256) void k_sound_speed(const float *__restrict__ rho, const float *__restrict__ p,
                                                     float *__restrict__ c, int64_t n, Gas g) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) c[i] = sound(p[i], frcp_nr(rho[i]), g);
}

template <int D>
__global__ __launch_bounds__(256) void k_wave_speeds(const float *__restrict__ L, const float *__restrict__ R,
                                                     float *__restrict__ sl, float *__restrict__ sr, int64_t n, Gas g,
                                                     uint32_t *__restrict__ bits) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    uint32_t v = 0;
    if (i < n) {
        const Q ql = q_from_prim<D>(L, n, i, g), qr = q_from_prim<D>(R, n, i, g);
        // the reference's jnp.maximum / jnp.minimum propagate NaN; v_max_f32
        // and v_min_f32 return the other operand
        const bool nan = ql.up != ql.up || qr.up != qr.up || ql.um != ql.um || qr.um != qr.um;
        const float r = nan ? __builtin_nanf("") : fmax3(ql.up, qr.up, 0.0f);
        const float l = nan ? __builtin_nanf("") : fmin3(ql.um, qr.um, 0.0f);
        if (sl) sl[i] = l;
        if (sr) sr[i] = r;
        // |S_R| = S_R, |S_L| = -S_L; the bits of a NaN exceed those of any number
        v = nan ? 0x7fc00000u : __float_as_uint(fmaxf(r, -l));
    }
    if (bits) {
        v = helper_block_umax(v);
        if (threadIdx.x == 0) atomicMax(bits, v);
    }
}

template <int D>
__global__ __launch_bounds__(256) void k_euler_flux(const float *__restrict__ w, float *__restrict__ F, int64_t n,
                                                    Gas g) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const Q q = q_from_prim<D>(w, n, i, g);
#pragma unroll
    for (int f = 0; f < 5; ++f) F[f * n + i] = q.F[f];
}

template <int D>
__global__ __launch_bounds__(256) void k_hll_flux(const float *__restrict__ L, const float *__restrict__ R,
                                                  float *__restrict__ F, int64_t n, Gas g) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float out[5];
    hll(q_from_prim<D>(L, n, i, g), q_from_prim<D>(R, n, i, g), out);
#pragma unroll
    for (int f = 0; f < 5; ++f) F[f * n + i] = out[f];
}

// out = in with the slice set_index along array axis `axis` (0 = variable,
// 1..3 = x, y, z) replaced by the slice get_index (fluid.py:317-331).
__global__ __launch_bounds__(256) void k_set_along_axis(const float *__restrict__ in, float *__restrict__ out, int nx,
                                                        int ny, int nz, int axis, int set_index, int get_index) {
    const int64_t nc = (int64_t)nx * ny * nz;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= 5 * nc) return;
    int64_t c[4] = {i / nc, (i % nc) / ((int64_t)ny * nz), (i % ((int64_t)ny * nz)) / nz, i % nz};
    if (c[axis] == set_index) c[axis] = get_index;
    out[i] = in[((c[0] * nx + c[1]) * ny + c[2]) * nz + c[3]];
}

hipError_t launch_sound_speed(const float *rho, const float *p, float *c, int64_t n, const Gas &g, hipStream_t st) {
    hipLaunchKernelGGL(k_sound_speed, grid1(n), dim3(256), 0, st, rho, p, c, n, g);
    return hipGetLastError();
}

hipError_t launch_wave_speeds(const float *L, const float *R, float *sl, float *sr, int64_t n, const Gas &g, int dir,
                              uint32_t *bits, hipStream_t st) {
    switch (dir) {
        case 1: hipLaunchKernelGGL(k_wave_speeds<1>, grid1(n), dim3(256), 0, st, L, R, sl, sr, n, g, bits); break;
        case 2: hipLaunchKernelGGL(k_wave_speeds<2>, grid1(n), dim3(256), 0, st, L, R, sl, sr, n, g, bits); break;
        case 3: hipLaunchKernelGGL(k_wave_speeds<3>, grid1(n), dim3(256), 0, st, L, R, sl, sr, n, g, bits); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_euler_flux(const float *w, float *F, int64_t n, const Gas &g, int dir, hipStream_t st) {
    switch (dir) {
        case 1: hipLaunchKernelGGL(k_euler_flux<1>, grid1(n), dim3(256), 0, st, w, F, n, g); break;
        case 2: hipLaunchKernelGGL(k_euler_flux<2>, grid1(n), dim3(256), 0, st, w, F, n, g); break;
        case 3: hipLaunchKernelGGL(k_euler_flux<3>, grid1(n), dim3(256), 0, st, w, F, n, g); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_hll_flux(const float *L, const float *R, float *F, int64_t n, const Gas &g, int dir,
                           hipStream_t st) {
    switch (dir) {
        case 1: hipLaunchKernelGGL(k_hll_flux<1>, grid1(n), dim3(256), 0, st, L, R, F, n, g); break;
        case 2: hipLaunchKernelGGL(k_hll_flux<2>, grid1(n), dim3(256), 0, st, L, R, F, n, g); break;
        case 3: hipLaunchKernelGGL(k_hll_flux<3>, grid1(n), dim3(256), 0, st, L, R, F, n, g); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_set_along_axis(const float *in, float *out, int nx, int ny, int nz, int axis, int set_index,
                                 int get_index, hipStream_t st) {
    hipLaunchKernelGGL(k_set_along_axis, grid1(5 * (int64_t)nx * ny * nz), dim3(256), 0, st, in, out, nx, ny, nz, axis,
                       set_index, get_index);
    return hipGetLastError();
}

}  // namespace tf
