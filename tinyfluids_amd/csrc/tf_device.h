// Device-side arithmetic of the HLL / Godunov step for gfx950.
//
// One cell of a sweep along axis D (1 = x, 2 = y, 3 = z) is described by its
// conserved vector U, its physical flux F along D, and its two characteristic
// speeds u_D + c and u_D - c. The Davis wave speeds, HLL flux and Godunov
// update below are the reference's fluid.py:119-160, :206-305 and :506-557,
// rearranged so that every per-cell quantity is computed once per sweep:
//   S_R = max(u_L + c_L, u_R + c_R, 0),  S_L = min(u_L - c_L, u_R - c_R, 0)
//   F   = (S_R F_L - S_L F_R + S_L S_R (U_R - U_L)) / (S_R - S_L)
//   U' = U + dt/dx (F[i-1/2] - F[i+1/2])
// with one v_rcp_f32 per interface and per cell in place of IEEE division
// (1 ulp), c = sqrt(gamma p / rho) via v_sqrt_f32, and the kinetic energy as
// m.v (rho v.v, without the square root, where a stored state's p is formed).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tf {

struct Gas {
    float gamma;    // adiabatic index
    float gm1;      // gamma - 1
    float inv_gm1;  // 1 / (gamma - 1)
    float k;        // dt / dx of the current step
};

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
// 1/x with one Newton step on v_rcp_f32. Measured exhaustively over a binade
// on gfx950, v_rcp_f32 is 1 ulp below the rounded reciprocal for 9.1% of
// inputs and 1 ulp above for 1.6% (mean -0.075 ulp). The velocities of a
// stored state, m * (1/rho), inherit that bias once per step, and over 1742
// steps of a shear layer it grew the x-momentum error to 7x the reference
// algorithm's float32 error; with the step (r + r (1 - x r), the residual
// exact in one FMA) in the stored conversion it is 0.5x. Inside a step the
// plain reciprocal measured as accurate.
__device__ __forceinline__ float frcp_nr(float x) {
    const float r = frcp(x);
    return fmaf(r, fmaf(-x, r, 1.0f), r);
}
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fmax3(float a, float b, float c) { return fmaxf(fmaxf(a, b), c); }
__device__ __forceinline__ float fmin3(float a, float b, float c) { return fminf(fminf(a, b), c); }

// Primitive cell with 1/rho carried along (rho, vx, vy, vz, p, rinv).
struct Prim {
    float r, v[3], p, rinv;
};

// One cell of a sweep: conserved U, flux F along the sweep axis, u_D +- c.
struct Q {
    float U[5];
    float F[5];
    float up, um;
};

// p / (gamma - 1): p * f32(1/(gamma-1)) and one Newton correction of the
// residual p - (gamma-1) e, which gives the rounded quotient. The product
// alone is not: f32(gamma-1) * f32(1/(gamma-1)) = 1 + 3.6e-8 for gamma 1.4
// and 5/3, so every primitive -> conserved -> primitive round trip (one per
// step) would add 3.6e-8 of the internal energy, a drift of 2e-5 after 600
// steps.
__device__ __forceinline__ float div_gm1(float p, const Gas &g) {
    const float e = p * g.inv_gm1;
    return fmaf(fmaf(-g.gm1, e, p), g.inv_gm1, e);
}

// Primitive -> conserved for a loaded cell (fluid.py:28-67), also 1/rho.
__device__ __forceinline__ void prim_to_cons(const float w[5], const Gas &g, float U[5], float &rinv) {
    const float r = w[0];
    U[0] = r;
    U[1] = r * w[1];
    U[2] = r * w[2];
    U[3] = r * w[3];
    const float mv = fmaf(U[3], w[3], fmaf(U[2], w[2], U[1] * w[1]));
    U[4] = fmaf(0.5f, mv, div_gm1(w[4], g));
    rinv = frcp(r);
}

// Conserved -> primitive (fluid.py:69-103) inside a step, where only the
// fluxes of the next sweep use the result.
__device__ __forceinline__ Prim cons_to_prim(const float U[5], const Gas &g) {
    Prim w;
    w.r = U[0];
    w.rinv = frcp(U[0]);
    w.v[0] = U[1] * w.rinv;
    w.v[1] = U[2] * w.rinv;
    w.v[2] = U[3] * w.rinv;
    const float mv = fmaf(U[3], w.v[2], fmaf(U[2], w.v[1], U[1] * w.v[0]));
    w.p = g.gm1 * fmaf(-0.5f, mv, U[4]);
    return w;
}

// Conserved -> primitive for a state that is stored (the end of a step, a
// sweep or a conversion): the pressure in the reference's form
// p = (gamma-1) rho (E/rho - |v|^2/2) (fluid.py:95), with E/rho rounded as a
// quotient (E * 1/rho plus one Newton correction of the residual). The
// stored state is converted back to conserved variables at the next step, so
// this rounding recurs once per step: measured over 50-600 steps of blasts
// and shear layers, it holds the float32 error at 0.3-0.8x of the reference
// algorithm's own float32 error on every conserved field, where
// (gamma-1)(E - m.v/2) reached 3x on the energy and E * 1/rho alone 2.8x on
// small momentum fields.
__device__ __forceinline__ Prim cons_to_prim_out(const float U[5], const Gas &g) {
    Prim w;
    w.r = U[0];
    w.rinv = frcp_nr(U[0]);
    w.v[0] = U[1] * w.rinv;
    w.v[1] = U[2] * w.rinv;
    w.v[2] = U[3] * w.rinv;
    float e = U[4] * w.rinv;
    e = fmaf(fmaf(-w.r, e, U[4]), w.rinv, e);
    const float vv = fmaf(w.v[2], w.v[2], fmaf(w.v[1], w.v[1], w.v[0] * w.v[0]));
    w.p = g.gm1 * w.r * fmaf(-0.5f, vv, e);
    return w;
}

// Sound speed (fluid.py:105-117) from p and 1/rho. The signal-speed maximum
// of a stored state is formed with this and frcp_nr both in the step's
// epilogue (from the values it stores) and in the CFL kernels (from the
// loaded state), so that the two agree bit for bit.
__device__ __forceinline__ float sound(float p, float rinv, const Gas &g) { return fsqrt(g.gamma * p * rinv); }

// Sweep quantities along axis D (1..3) of a cell with conserved U and
// primitive velocity/pressure (fluid.py:206-259 for F).
template <int D>
__device__ __forceinline__ Q make_q(const float U[5], const float v[3], float p, float rinv, const Gas &g) {
    Q q;
#pragma unroll
    for (int i = 0; i < 5; ++i) q.U[i] = U[i];
    const float u = v[D - 1];
    q.F[0] = U[D];
    q.F[1] = U[1] * u;
    q.F[2] = U[2] * u;
    q.F[3] = U[3] * u;
    q.F[D] = fmaf(U[D], u, p);
    q.F[4] = (U[4] + p) * u;
    const float c = sound(p, rinv, g);
    q.up = u + c;
    q.um = u - c;
    return q;
}

// Compact form of Q kept between rows of a stream: U, the velocity along the
// sweep axis, p and c; F and u +- c are rebuilt when needed.
struct QS {
    float U[5];
    float u, p, c;
};

template <int D>
__device__ __forceinline__ QS make_qs(const float U[5], const float v[3], float p, float rinv, const Gas &g) {
    QS q;
#pragma unroll
    for (int i = 0; i < 5; ++i) q.U[i] = U[i];
    q.u = v[D - 1];
    q.p = p;
    q.c = sound(p, rinv, g);
    return q;
}

template <int D>
__device__ __forceinline__ Q expand_qs(const QS &s) {
    Q q;
#pragma unroll
    for (int i = 0; i < 5; ++i) q.U[i] = s.U[i];
    q.F[0] = s.U[D];
    q.F[1] = s.U[1] * s.u;
    q.F[2] = s.U[2] * s.u;
    q.F[3] = s.U[3] * s.u;
    q.F[D] = fmaf(s.U[D], s.u, s.p);
    q.F[4] = (s.U[4] + s.p) * s.u;
    q.up = s.u + s.c;
    q.um = s.u - s.c;
    return q;
}

// Coefficients of the HLL flux for Davis speeds S_R = sr >= 0 >= S_L = sl:
// a = S_R/(S_R-S_L), b = S_L/(S_R-S_L), e = a S_L.
struct HllCoef {
    float a, b, e;
};
__device__ __forceinline__ HllCoef hll_coef(float sr, float sl) {
    const float inv = frcp(sr - sl);
    HllCoef c;
    c.a = sr * inv;
    c.b = sl * inv;
    c.e = c.a * sl;
    return c;
}
// One component: F = e (U_R - U_L) + (a F_L - b F_R), given dU = U_R - U_L
// and aFL = a F_L. Every HLL flux of the step (x, y and z, halo and line
// tilings) evaluates this same expression, so results do not depend on
// where the left state came from (registers, a lane shift or LDS).
__device__ __forceinline__ float hll_comp(const HllCoef &c, float dU, float aFL, float FR) {
    return fmaf(c.e, dU, fmaf(-c.b, FR, aFL));
}

// HLL flux at the interface between L and R (fluid.py:261-305).
__device__ __forceinline__ void hll(const Q &L, const Q &R, float F[5]) {
    const HllCoef c = hll_coef(fmax3(L.up, R.up, 0.0f), fmin3(L.um, R.um, 0.0f));
#pragma unroll
    for (int i = 0; i < 5; ++i) F[i] = hll_comp(c, R.U[i] - L.U[i], c.a * L.F[i], R.F[i]);
}

// Godunov update U + k (F_left - F_right) (fluid.py:547-551).
__device__ __forceinline__ void godunov(const float U[5], const float Fl[5], const float Fr[5], float k,
                                        float out[5]) {
#pragma unroll
    for (int i = 0; i < 5; ++i) out[i] = fmaf(k, Fl[i] - Fr[i], U[i]);
}

// max_d |v_d| + c: the cell's contribution to the CFL signal speed. It equals
// the max over the cell's interfaces of max(|S_R|, |S_L|) of fluid.py:162-195
// because rounding is monotone and sign-symmetric.
__device__ __forceinline__ float signal_speed(const float v[3], float c) {
    return fmax3(fabsf(v[0]), fabsf(v[1]), fabsf(v[2])) + c;
}

// Periodic wrap of i in [-n, 2n) and the open-boundary index clamp of
// fluid.py:382-413 (copies from the inner neighbour, applied after the wrap).
__device__ __forceinline__ int wrap1(int i, int n) {
    i = i < 0 ? i + n : i;
    return i >= n ? i - n : i;
}
__device__ __forceinline__ int open_clamp(int i, int n) { return n == 2 ? 1 : min(max(i, 1), n - 2); }
__device__ __forceinline__ int cell_index(int i, int n, int open_bc) {
    int j = i % n;
    j = j < 0 ? j + n : j;
    return open_bc ? open_clamp(j, n) : j;
}

// Lane shifts inside a wave64 (DPP wave_shr:1 / wave_shl:1): from_left gives
// lane l the value of lane l-1, from_right the value of lane l+1; the edge
// lane with no source reads 0 (bound_ctrl). Written right at their use, the
// compiler folds them into the consuming VOP2 instruction (v_subrev_f32_dpp,
// v_mul_f32_dpp) instead of issuing a separate v_mov_b32_dpp; it only looks
// a few instructions ahead of the shift for that.
__device__ __forceinline__ float from_left(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float from_right(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x130, 0xf, 0xf, true));
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

}  // namespace tf
