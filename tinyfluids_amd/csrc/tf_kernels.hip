// gfx950 kernels of the jax_tinyfluids time step.
//
// The hot kernel is the fused split step (reference fluid.py:560-578): the X,
// Y and Z sweeps of one time step in a single pass over HBM.
//
//   Work unit: one wave64 owns a tile of TX x-columns by 62 z-columns and
//   streams through a chunk of `ly` y-rows (plus one halo row on each side).
//   Lanes run along z, the contiguous axis, so every load and store is a
//   coalesced 256-byte row segment; lanes 0 and 63 are z-halo columns.
//   - X sweep: each lane loads the TX+2 x-neighbours of its column for the
//     current row and does the whole x-stencil in registers (x is the
//     slowest axis, so its neighbours are other rows of the same lane).
//   - Y sweep: the wave streams along y, so the y-neighbour of a cell is the
//     previous row of the same lane; its sweep quantities and the flux at the
//     lower interface stay in registers between rows (Lie splitting: the Y
//     sweep consumes the X-swept cells, fluid.py:574-575).
//   - Z sweep: the z-neighbours are the adjacent lanes; sweep quantities move
//     one lane with DPP wave_shr / wave_shl (fluid.py:576).
//   - Epilogue: each output cell contributes max_d |v_d| + c to the next CFL
//     time step (fluid.py:425-504) through an integer max on float bits,
//     which also propagates NaN, so the next dt needs no second read.
//   Waves never exchange data, so the kernel has no barriers and no LDS.
//
// Algorithmic traffic is 40 bytes per cell-update (5 float32 read, 5 written);
// the halo rows and columns are re-read from L2 / MALL.
#include "tf_internal.h"

namespace tf {

__device__ __forceinline__ int wave_id_uniform() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ uint32_t wave_umax(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}

template <int TX>
__device__ __forceinline__ void load_row(const float *__restrict__ in, size_t nc, const size_t (&xoff)[TX + 2],
                                         size_t yoff, uint32_t zg, float (&dst)[TX + 2][5]) {
#pragma unroll
    for (int j = 0; j < TX + 2; ++j) {
        const float *row = in + xoff[j] + yoff;
#pragma unroll
        for (int f = 0; f < 5; ++f) dst[j][f] = __builtin_nontemporal_load(row + f * nc + zg);
    }
}

// Sweep quantities along x of a loaded primitive cell.
__device__ __forceinline__ Q q_from_prim_x(const float (&w)[5], const Gas &g) {
    float U[5], rinv;
    prim_to_cons(w, g, U, rinv);
    const float v[3] = {w[1], w[2], w[3]};
    return make_q<1>(U, v, w[4], rinv, g);
}

// One wave: the fused X->Y->Z step on its tile. Returns the wave-local max
// of the signal-speed bits of the cells it wrote (0 if none).
template <int TX>
__device__ __forceinline__ uint32_t fused_tile(const float *__restrict__ in, float *__restrict__ out, int nx, int ny,
                                               int nz, int open_bc, const Gas &g, int xt, int zt, int yc, int ly) {
    const int lane = lane_id();
    const int x0 = xt * TX, z0 = zt * kZT, y0 = yc * ly;
    const int nzo = min(kZT, nz - z0);
    const int nyo = min(ly, ny - y0);
    const int nxo = min(TX, nx - x0);
    const size_t plane = (size_t)ny * nz;
    const size_t nc = plane * nx;
    const uint32_t zg = (uint32_t)cell_index(z0 - 1 + lane, nz, open_bc);
    const bool zout = lane >= 1 && lane <= nzo;

    size_t xoff[TX + 2];
#pragma unroll
    for (int j = 0; j < TX + 2; ++j) xoff[j] = (size_t)cell_index(x0 - 1 + j, nx, open_bc) * plane;

    float cur[TX + 2][5];
    load_row<TX>(in, nc, xoff, (size_t)cell_index(y0 - 1, ny, open_bc) * nz, zg, cur);

    Q qy[TX];         // Y-sweep quantities of the previous row
    float fy[TX][5];  // Y flux at the lower interface of the previous row
    uint32_t vmax = 0;
    const int nrows = nyo + 2;

    for (int r = 0; r < nrows; ++r) {
        float nxt[TX + 2][5];
        if (r + 1 < nrows) {
            const int yn = y0 + r;  // row y0 - 1 + (r + 1)
            const int ygn = open_bc ? cell_index(yn, ny, 1) : wrap1(yn, ny);
            load_row<TX>(in, nc, xoff, (size_t)ygn * nz, zg, nxt);
        }
        // ---- X sweep of this row: cells j = 1..TX from x-neighbours j-1, j+1
        Q qb = q_from_prim_x(cur[0], g);
        Q qc = q_from_prim_x(cur[1], g);
        float fa[5];
        hll(qb, qc, fa);
#pragma unroll
        for (int s = 0; s < TX; ++s) {
            const Q qn = q_from_prim_x(cur[s + 2], g);
            float fb[5];
            hll(qc, qn, fb);
            float U1[5];
            godunov(qc.U, fa, fb, g.k, U1);
            const Prim w1 = cons_to_prim(U1, g);
            const Q q1 = make_q<2>(U1, w1.v, w1.p, w1.rinv, g);
            // ---- Y sweep: interface between the previous row and this one
            if (r == 0) {
                qy[s] = q1;
            } else if (r == 1) {
                hll(qy[s], q1, fy[s]);
                qy[s] = q1;
            } else {
                float fn[5];
                hll(qy[s], q1, fn);
                float U2[5];
                godunov(qy[s].U, fy[s], fn, g.k, U2);  // the previous row is now Y-swept
#pragma unroll
                for (int i = 0; i < 5; ++i) fy[s][i] = fn[i];
                qy[s] = q1;
                // ---- Z sweep across lanes
                const Prim w2 = cons_to_prim(U2, g);
                const Q q2 = make_q<3>(U2, w2.v, w2.p, w2.rinv, g);
                Q ql;
#pragma unroll
                for (int i = 0; i < 5; ++i) {
                    ql.U[i] = from_left(q2.U[i]);
                    ql.F[i] = from_left(q2.F[i]);
                }
                ql.up = from_left(q2.up);
                ql.um = from_left(q2.um);
                float fz[5], fzr[5];
                hll(ql, q2, fz);
#pragma unroll
                for (int i = 0; i < 5; ++i) fzr[i] = from_right(fz[i]);
                float U3[5];
                godunov(q2.U, fz, fzr, g.k, U3);
                const Prim w3 = cons_to_prim(U3, g);
                if (zout && s < nxo) {
                    float *o = out + (size_t)(x0 + s) * plane + (size_t)(y0 + r - 2) * nz + (uint32_t)(z0 - 1 + lane);
                    __builtin_nontemporal_store(w3.r, o);
                    __builtin_nontemporal_store(w3.v[0], o + nc);
                    __builtin_nontemporal_store(w3.v[1], o + 2 * nc);
                    __builtin_nontemporal_store(w3.v[2], o + 3 * nc);
                    __builtin_nontemporal_store(w3.p, o + 4 * nc);
                    const float sp = signal_speed(w3.v, sound(w3.p, w3.rinv, g));
                    vmax = max(vmax, __float_as_uint(sp));
                }
            }
            qc = qn;
#pragma unroll
            for (int i = 0; i < 5; ++i) fa[i] = fb[i];
        }
#pragma unroll
        for (int j = 0; j < TX + 2; ++j)
#pragma unroll
            for (int f = 0; f < 5; ++f) cur[j][f] = nxt[j][f];
    }
    return vmax;
}

__device__ __forceinline__ void decode_tile(int local, int ntz, int ntx, int &xt, int &zt, int &yc) {
    zt = local % ntz;
    const int rest = local / ntz;
    xt = rest % ntx;
    yc = rest / ntx;
}

// Fused step of one grid with an explicit dt (the _evolve_state entry point).
template <int TX>
__global__ __launch_bounds__(kBlock) void k_step_direct(SimDev S, const float *__restrict__ in,
                                                        float *__restrict__ out, const double *__restrict__ dt,
                                                        double dx, uint32_t *__restrict__ bits_out) {
    const int w = blockIdx.x * (kBlock / kWave) + wave_id_uniform();
    if (w >= S.waves) return;
    Gas g{S.gamma, S.gm1, S.inv_gm1, (float)(*dt / dx)};
    int xt, zt, yc;
    decode_tile(w, S.ntz, S.ntx, xt, zt, yc);
    uint32_t v = fused_tile<TX>(in, out, S.nx, S.ny, S.nz, S.open_bc, g, xt, zt, yc, S.ly);
    if (bits_out) {
        v = wave_umax(v);
        if (lane_id() == 0) atomicMax(bits_out, v);
    }
}

// Fused step of every active simulation of a table; simulation s owns waves
// [prefix[s], prefix[s+1]). Input/output buffers follow the step parity.
template <int TX>
__global__ __launch_bounds__(kBlock) void k_step_sims(const SimDev *__restrict__ sims, const int *__restrict__ prefix,
                                                      int nsims) {
    const int w = blockIdx.x * (kBlock / kWave) + wave_id_uniform();
    if (w >= prefix[nsims]) return;
    int lo = 0, hi = nsims;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= w) lo = mid; else hi = mid;
    }
    const SimDev &S = sims[lo];
    Ctrl *C = S.ctrl;
    if (!C->active) return;
    const int par = (int)(C->n_iter & 1);
    Gas g{S.gamma, S.gm1, S.inv_gm1, C->k};
    int xt, zt, yc;
    decode_tile(w - prefix[lo], S.ntz, S.ntx, xt, zt, yc);
    uint32_t v = fused_tile<TX>(S.buf[par], S.buf[par ^ 1], S.nx, S.ny, S.nz, S.open_bc, g, xt, zt, yc, S.ly);
    v = wave_umax(v);
    if (lane_id() == 0) atomicMax(&C->max_bits[par ^ 1], v);
}

// ---------------------------------------------------------------- dt control
__device__ __forceinline__ void finalize_ctrl(Ctrl *C, int first) {
    if (!first) {
        if (!C->active) return;
        C->t += C->dt;
        C->n_iter += 1;
    }
    const int par = (int)(C->n_iter & 1);
    const uint32_t bits = C->max_bits[par];
    C->max_bits[par ^ 1] = 0u;
    const float m = __uint_as_float(bits);
    C->max_speed = m;
    if (bits >= 0x7f800000u) C->nonfinite = 1;
    const double dt = C->c_cfl * C->dx / (double)m;  // fluid.py:502
    C->dt = dt;
    C->k = (float)(dt / C->dx);
    C->active = (C->t < C->t_final) && (C->max_steps < 0 || C->n_iter < C->max_steps);  // time_integration.py:85-87
}

__global__ void k_finalize_sims(const SimDev *__restrict__ sims, int nsims) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nsims) finalize_ctrl(sims[i].ctrl, 0);
}

__global__ void k_finalize_one(Ctrl *C, int first) {
    if (threadIdx.x == 0 && blockIdx.x == 0) finalize_ctrl(C, first);
}

__global__ void k_ctrl_reset(Ctrl *C, double dx, double c_cfl, double t_final, int64_t max_steps) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    C->t = 0.0;
    C->dt = 0.0;
    C->n_iter = 0;
    C->active = 1;
    C->nonfinite = 0;
    C->max_speed = 0.0f;
    C->max_bits[0] = 0u;
    C->max_bits[1] = 0u;
    C->k = 0.0f;
    C->t_final = t_final;
    C->dx = dx;
    C->c_cfl = c_cfl;
    C->max_steps = max_steps;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Max over cells of max_d |v_d| + c (the CFL signal speed, fluid.py:425-500).
__global__ __launch_bounds__(256) void k_speed(const float *__restrict__ prim, int64_t nc, float gamma,
                                               uint32_t *__restrict__ bits_out) {
    uint32_t v = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += stride) {
        const float r = prim[i], vx = prim[nc + i], vy = prim[2 * nc + i], vz = prim[3 * nc + i],
                    p = prim[4 * nc + i];
        const float c = fsqrt(gamma * p * frcp(r));
        const float w[3] = {vx, vy, vz};
        v = max(v, __float_as_uint(signal_speed(w, c)));
    }
    v = wave_umax(v);
    if ((threadIdx.x & 63) == 0) atomicMax(bits_out, v);
}

__global__ void k_dt_from_bits(const uint32_t *bits, double dx, double c_cfl, double *dt_out, float *max_out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const float m = __uint_as_float(*bits);
    if (dt_out) *dt_out = c_cfl * dx / (double)m;
    if (max_out) *max_out = m;
}

// ---------------------------------------------------------------- per-axis sweep
// One split sweep (fluid.py:506-557), one thread per cell; both interface
// fluxes of a cell are evaluated from the same inputs by the same code as the
// neighbour evaluates them, so they agree bit for bit.
template <int D>
__global__ __launch_bounds__(256) void k_sweep(const float *__restrict__ in, float *__restrict__ out, int nx, int ny,
                                               int nz, Gas g, const double *__restrict__ dt, double dx) {
    const int64_t nc = (int64_t)nx * ny * nz;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    g.k = (float)(*dt / dx);
    const int z = (int)(i % nz);
    const int64_t xy = i / nz;
    const int y = (int)(xy % ny);
    const int x = (int)(xy / ny);
    int64_t im, ip;
    if (D == 1) {
        im = i + (int64_t)(wrap1(x - 1, nx) - x) * ny * nz;
        ip = i + (int64_t)(wrap1(x + 1, nx) - x) * ny * nz;
    } else if (D == 2) {
        im = i + (int64_t)(wrap1(y - 1, ny) - y) * nz;
        ip = i + (int64_t)(wrap1(y + 1, ny) - y) * nz;
    } else {
        im = i + (wrap1(z - 1, nz) - z);
        ip = i + (wrap1(z + 1, nz) - z);
    }
    Q q[3];
    const int64_t idx[3] = {im, i, ip};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float w[5];
#pragma unroll
        for (int f = 0; f < 5; ++f) w[f] = in[f * nc + idx[c]];
        float U[5], rinv;
        prim_to_cons(w, g, U, rinv);
        const float v[3] = {w[1], w[2], w[3]};
        q[c] = make_q<D>(U, v, w[4], rinv, g);
    }
    float fl[5], fr[5], U1[5];
    hll(q[0], q[1], fl);
    hll(q[1], q[2], fr);
    godunov(q[1].U, fl, fr, g.k, U1);
    const Prim w1 = cons_to_prim(U1, g);
    out[i] = w1.r;
    out[nc + i] = w1.v[0];
    out[2 * nc + i] = w1.v[1];
    out[3 * nc + i] = w1.v[2];
    out[4 * nc + i] = w1.p;
}

// ---------------------------------------------------------------- elementwise
__global__ __launch_bounds__(256) void k_open_bc(const float *__restrict__ in, float *__restrict__ out, int nx, int ny,
                                                 int nz) {
    const int64_t nc = (int64_t)nx * ny * nz;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const int z = (int)(i % nz);
    const int64_t xy = i / nz;
    const int y = (int)(xy % ny);
    const int x = (int)(xy / ny);
    const int64_t src = ((int64_t)open_clamp(x, nx) * ny + open_clamp(y, ny)) * nz + open_clamp(z, nz);
#pragma unroll
    for (int f = 0; f < 5; ++f) out[f * nc + i] = in[f * nc + src];
}

__global__ __launch_bounds__(256) void k_p2c(const float *in, float *out, int64_t nc, Gas g) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += stride) {
        float w[5], U[5], rinv;
#pragma unroll
        for (int f = 0; f < 5; ++f) w[f] = in[f * nc + i];
        prim_to_cons(w, g, U, rinv);
#pragma unroll
        for (int f = 0; f < 5; ++f) out[f * nc + i] = U[f];
    }
}

__global__ __launch_bounds__(256) void k_c2p(const float *in, float *out, int64_t nc, Gas g) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += stride) {
        float U[5];
#pragma unroll
        for (int f = 0; f < 5; ++f) U[f] = in[f * nc + i];
        const Prim w = cons_to_prim(U, g);
        out[i] = w.r;
        out[nc + i] = w.v[0];
        out[2 * nc + i] = w.v[1];
        out[3 * nc + i] = w.v[2];
        out[4 * nc + i] = w.p;
    }
}

// ---------------------------------------------------------------- float64 totals
// Block b sums the cells [b*chunk, (b+1)*chunk) in a fixed order; the final
// kernel adds the block partials in block order: the result does not depend
// on scheduling.
__global__ __launch_bounds__(256) void k_totals_partial(const float *__restrict__ prim, int64_t nc, double gm1,
                                                        double *__restrict__ partials) {
    __shared__ double sh[5][256];
    const int64_t chunk = (nc + gridDim.x - 1) / gridDim.x;
    const int64_t beg = (int64_t)blockIdx.x * chunk;
    const int64_t end = min(nc, beg + chunk);
    double acc[5] = {0, 0, 0, 0, 0};
    for (int64_t i = beg + threadIdx.x; i < end; i += blockDim.x) {
        const double r = prim[i], vx = prim[nc + i], vy = prim[2 * nc + i], vz = prim[3 * nc + i],
                     p = prim[4 * nc + i];
        acc[0] += r;
        acc[1] += r * vx;
        acc[2] += r * vy;
        acc[3] += r * vz;
        acc[4] += p / gm1 + 0.5 * r * (vx * vx + vy * vy + vz * vz);
    }
#pragma unroll
    for (int f = 0; f < 5; ++f) sh[f][threadIdx.x] = acc[f];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
#pragma unroll
            for (int f = 0; f < 5; ++f) sh[f][threadIdx.x] += sh[f][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 5) partials[blockIdx.x * 5 + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void k_totals_final(const double *__restrict__ partials, int nblocks, double *__restrict__ out5) {
    if (threadIdx.x >= 5 || blockIdx.x != 0) return;
    double a = 0.0;
    for (int b = 0; b < nblocks; ++b) a += partials[b * 5 + threadIdx.x];
    out5[threadIdx.x] = a;
}

__global__ void k_gather_status(const SimDev *__restrict__ sims, int nsims, Ctrl *__restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nsims) dst[i] = *sims[i].ctrl;
}

// ================================================================ host launchers
static inline dim3 grid1(int64_t n, int block = 256) { return dim3((unsigned)((n + block - 1) / block)); }

hipError_t launch_step_direct(int tx, const SimDev &S, const float *in, float *out, const double *dt, double dx,
                              uint32_t *bits, hipStream_t st) {
    const dim3 grid((unsigned)((S.waves + 3) / 4)), block(kBlock);
    switch (tx) {
        case 2: hipLaunchKernelGGL(k_step_direct<2>, grid, block, 0, st, S, in, out, dt, dx, bits); break;
        case 4: hipLaunchKernelGGL(k_step_direct<4>, grid, block, 0, st, S, in, out, dt, dx, bits); break;
        case 8: hipLaunchKernelGGL(k_step_direct<8>, grid, block, 0, st, S, in, out, dt, dx, bits); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_step_sims(int tx, const SimDev *sims, const int *prefix, int nsims, int total_waves,
                            hipStream_t st) {
    if (total_waves <= 0) return hipSuccess;
    const dim3 grid((unsigned)((total_waves + 3) / 4)), block(kBlock);
    switch (tx) {
        case 2: hipLaunchKernelGGL(k_step_sims<2>, grid, block, 0, st, sims, prefix, nsims); break;
        case 4: hipLaunchKernelGGL(k_step_sims<4>, grid, block, 0, st, sims, prefix, nsims); break;
        case 8: hipLaunchKernelGGL(k_step_sims<8>, grid, block, 0, st, sims, prefix, nsims); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_finalize_sims(const SimDev *sims, int nsims, hipStream_t st) {
    hipLaunchKernelGGL(k_finalize_sims, grid1(nsims, 64), dim3(64), 0, st, sims, nsims);
    return hipGetLastError();
}

hipError_t launch_finalize_one(Ctrl *C, int first, hipStream_t st) {
    hipLaunchKernelGGL(k_finalize_one, dim3(1), dim3(64), 0, st, C, first);
    return hipGetLastError();
}

hipError_t launch_ctrl_reset(Ctrl *C, double dx, double c_cfl, double t_final, int64_t max_steps, hipStream_t st) {
    hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(64), 0, st, C, dx, c_cfl, t_final, max_steps);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t v, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_u32, grid1(n), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_speed(const float *prim, int64_t nc, float gamma, uint32_t *bits, hipStream_t st) {
    int64_t blocks = (nc + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_speed, dim3((unsigned)blocks), dim3(256), 0, st, prim, nc, gamma, bits);
    return hipGetLastError();
}

hipError_t launch_dt_from_bits(const uint32_t *bits, double dx, double c_cfl, double *dt_out, float *max_out,
                               hipStream_t st) {
    hipLaunchKernelGGL(k_dt_from_bits, dim3(1), dim3(64), 0, st, bits, dx, c_cfl, dt_out, max_out);
    return hipGetLastError();
}

hipError_t launch_sweep(int axis, const float *in, float *out, int nx, int ny, int nz, const Gas &g, const double *dt,
                        double dx, hipStream_t st) {
    const int64_t nc = (int64_t)nx * ny * nz;
    switch (axis) {
        case 1: hipLaunchKernelGGL(k_sweep<1>, grid1(nc), dim3(256), 0, st, in, out, nx, ny, nz, g, dt, dx); break;
        case 2: hipLaunchKernelGGL(k_sweep<2>, grid1(nc), dim3(256), 0, st, in, out, nx, ny, nz, g, dt, dx); break;
        case 3: hipLaunchKernelGGL(k_sweep<3>, grid1(nc), dim3(256), 0, st, in, out, nx, ny, nz, g, dt, dx); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_open_bc(const float *in, float *out, int nx, int ny, int nz, hipStream_t st) {
    const int64_t nc = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(k_open_bc, grid1(nc), dim3(256), 0, st, in, out, nx, ny, nz);
    return hipGetLastError();
}

hipError_t launch_p2c(const float *in, float *out, int64_t nc, const Gas &g, hipStream_t st) {
    int64_t blocks = (nc + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_p2c, dim3((unsigned)blocks), dim3(256), 0, st, in, out, nc, g);
    return hipGetLastError();
}

hipError_t launch_c2p(const float *in, float *out, int64_t nc, const Gas &g, hipStream_t st) {
    int64_t blocks = (nc + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_c2p, dim3((unsigned)blocks), dim3(256), 0, st, in, out, nc, g);
    return hipGetLastError();
}

hipError_t launch_totals(const float *prim, int64_t nc, double gamma, double *partials, double *out5,
                         hipStream_t st) {
    hipLaunchKernelGGL(k_totals_partial, dim3(kTotalsBlocks), dim3(256), 0, st, prim, nc, gamma - 1.0, partials);
    hipLaunchKernelGGL(k_totals_final, dim3(1), dim3(64), 0, st, partials, kTotalsBlocks, out5);
    return hipGetLastError();
}

hipError_t launch_gather_status(const SimDev *sims, int nsims, Ctrl *dst, hipStream_t st) {
    hipLaunchKernelGGL(k_gather_status, grid1(nsims, 64), dim3(64), 0, st, sims, nsims, dst);
    return hipGetLastError();
}

}  // namespace tf
