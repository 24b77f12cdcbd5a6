// gfx950 kernels of the jax_tinyfluids time step.
//
// The hot kernels are the fused split step (reference fluid.py:560-578): the
// X, Y and Z sweeps of one time step in a single pass over HBM.
//
//   Lanes run along z, the contiguous axis; a wave owns TX x-columns and
//   streams along y through a chunk of rows (plus one halo row each side).
//   - X sweep: each lane loads the TX+2 x-neighbours of its column for the
//     current row and does the whole x-stencil in registers (x is the
//     slowest axis, so its neighbours are other rows of the same lane).
//   - Y sweep: the wave streams along y, so the y-neighbour of a cell is the
//     previous row of the same lane; its sweep quantities and the flux at the
//     lower interface stay in registers between rows (Lie splitting: the Y
//     sweep consumes the X-swept cells, fluid.py:574-575).
//   - Z sweep: the z-neighbours are the adjacent lanes (fluid.py:576):
//     through LDS when a workgroup holds whole z-lines (LineCtx, nz <= 1024),
//     through DPP lane shifts when one wave holds a 62-cell z-tile with two
//     halo lanes (HaloCtx).
//   - Epilogue: each output cell contributes max_d |v_d| + c to the next CFL
//     time step (fluid.py:425-504) through an integer max on float bits,
//     which also propagates NaN, so the next dt needs no second read.
//   - Memory issue: raised wave priorities around stores and x-halo loads
//     (MemPriority, line_load_priority) let a wave's memory instructions
//     pass the other resident waves' arithmetic.
//
// Algorithmic traffic is 40 bytes per cell-update (5 float32 read, 5 written);
// the halo rows and columns are re-read from L2.
#include "tf_internal.h"

// Cache policy of the state stores: plain write-back stores, so that L2
// merges the partial 128-byte lines that two neighbouring 62-cell z-tiles
// write; streaming (NT) stores write those halves to HBM separately and
// measured ~25% slower. Whole-line tiles write full lines, yet NT stores
// measured 4% slower there too at 1024^3 (and 0.5% faster at 1536x512x512).
constexpr int kStoreAux = 0;

namespace tf {

__device__ __forceinline__ int wave_id_uniform() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ uint32_t wave_umax(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}

// A wave-uniform address pinned to SGPRs (readfirstlane of both halves), so
// that everything derived from it stays scalar.
template <typename T>
__device__ __forceinline__ T *sgpr_ptr(T *p) {
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (T *)(((uint64_t)hi << 32) | lo);
}

// Raw buffer access: byte offset = voffset + soffset. Loads use the default
// (cached) policy; streaming loads of any of the columns measured slower.
__device__ __forceinline__ float buf_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, (int)voff, (int)soff, kStoreAux);
}
__device__ __forceinline__ __amdgpu_buffer_rsrc_t field_rsrc(const float *field) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)field, (short)0, -1, 0x00020000);
}
// Structured (indexed, stride 4) buffer access. Its range check is on the
// element index only (index < num_records), so every dword of a field of up
// to 2^30 cells is reachable, the last one of an exactly 4 GiB field
// included; the scalar byte offset is added unchecked. The LLVM intrinsics
// have no clang builtin, so they are bound by name.
extern "C" __device__ float tf_struct_load_f32(__amdgpu_buffer_rsrc_t rsrc, int vindex, int voffset, int soffset,
                                               int aux) __asm("llvm.amdgcn.struct.ptr.buffer.load.f32");
extern "C" __device__ void tf_struct_store_f32(float v, __amdgpu_buffer_rsrc_t rsrc, int vindex, int voffset,
                                               int soffset, int aux) __asm("llvm.amdgcn.struct.ptr.buffer.store.f32");
__device__ __forceinline__ __amdgpu_buffer_rsrc_t field_rsrc_indexed(const float *field) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)field, (short)4, -1, 0x00020000);
}

// An empty asm that takes v in a VGPR: v is computed before this point on
// the lanes active here, and its computation cannot be sunk into a later
// lane-masked block. Emits no instruction.
__device__ __forceinline__ void pin_vgpr(float &v) { asm volatile("" : "+v"(v)); }

// Wave priority while a wave issues memory instructions: the SIMD's arbiter
// issues a higher-priority wave's instructions ahead of the other resident
// waves' arithmetic. Stores go at 2 (MemPriority). In the whole-line tiles
// the loads of the two x-halo columns (the neighbouring tiles' columns,
// mostly L2 hits) go at 1 and those of the tile's own columns at 0, each
// column's loads between two s_setprio, which also keeps the compiler from
// moving them across (line_load_priority); the 62-cell tiles load at 2.
// Measured on the 1024^3 step: 9.56 ms without priorities, 9.08 ms with
// every load and store at 2, 8.79 ms with this order (own loads at or above
// the stores' level, or without the s_setprio around them, were slower);
// 256^3 and 1536x512x512 gain another 1-2%, the 62-cell tiles nothing.
struct MemPriority {
    __device__ __forceinline__ MemPriority() { __builtin_amdgcn_s_setprio(2); }
    __device__ __forceinline__ ~MemPriority() { __builtin_amdgcn_s_setprio(0); }
};
// Loaded column j of a whole-line tile of TX columns (0 and TX + 1 are its
// x-halo): priority before (on) and after (!on) its loads.
template <int TX>
__device__ __forceinline__ void line_load_priority(int j, bool on) {
    if (on && (j == 0 || j == TX + 1)) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
}

__device__ __forceinline__ float sgpr_f(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

// ---------------------------------------------------------------------------
// Tile access policies. A policy gives the stream (stream_row) its loads,
// its z-neighbour exchange and its stores; the arithmetic is shared.
//
// HaloCtx: one wave per tile of TX x-columns x 62 z-columns (lanes 0 and 63
// are z-halo columns recomputed by the wave), neighbours by DPP lane shifts.
// Each of the 5 fields of input and output has a buffer resource in SGPRs; a
// cell row is a 32-bit scalar byte offset (x/y position) plus the lane's
// byte offset along z, so loads and stores need no per-lane address
// arithmetic. A raw buffer access is in range only if offset + 4 <=
// num_records <= 0xFFFFFFFF, so the last dword of a field of exactly 4 GiB
// is out of reach of its resource: the tiles that touch that row run in a
// separate kernel with FAR = true (64-bit global addresses).
// ---------------------------------------------------------------------------
template <int TX, bool FAR>
struct HaloCtx {
    __amdgpu_buffer_rsrc_t rin[5], rout[5];
    const float *in_ptr;
    float *out_ptr;
    size_t nc;
    uint32_t xoff[TX + 2];  // byte offsets of the loaded x-rows (uniform)
    uint32_t zb;            // this lane's load byte offset along z
    uint32_t zbo;           // this lane's store byte offset (z - z0)
    int ny, y0, x0, z0, nxo, nrows, open_bc;
    uint32_t plane_b, row_b;  // bytes per x-plane and per y-row of one field
    bool zout;

    __device__ __forceinline__ void load(int j, uint32_t yoff, float (&w)[5]) const {
        const uint32_t soff = xoff[j] + yoff;
        const MemPriority prio;
        if constexpr (FAR) {
#pragma unroll
            for (int f = 0; f < 5; ++f) w[f] = in_ptr[f * nc + ((size_t)soff + zb) / 4];
        } else {
#pragma unroll
            for (int f = 0; f < 5; ++f) w[f] = buf_load(rin[f], zb, soff);
        }
    }
    // Z flux at the lower interface of this lane's cell (left state = lane - 1).
    __device__ __forceinline__ void zflux(const Q &q, float (&F)[5], int) const {
        const HllCoef c = hll_coef(fmax3(from_left(q.up), q.up, 0.0f), fmin3(from_left(q.um), q.um, 0.0f));
#pragma unroll
        for (int i = 0; i < 5; ++i) F[i] = hll_comp(c, q.U[i] - from_left(q.U[i]), c.a * from_left(q.F[i]), q.F[i]);
    }
    // Godunov update with the upper interface's flux taken from lane + 1.
    __device__ __forceinline__ void zupdate(const float (&U)[5], const float (&F)[5], float k, float (&out)[5],
                                            int) const {
#pragma unroll
        for (int i = 0; i < 5; ++i) out[i] = fmaf(k, F[i] - from_right(F[i]), U[i]);
    }
    // Z sweep of the lane's cell: out = U + k (F[z-1/2] - F[z+1/2]).
    __device__ __forceinline__ void zsweep(const Q &q, float k, float (&out)[5], int s) const {
        float F[5];
        zflux(q, F, s);
        zupdate(q.U, F, k, out, s);
    }
    __device__ __forceinline__ bool out_ok(int s) const { return zout && s < nxo; }
    __device__ __forceinline__ void store(int s, int y, const float (&w)[5]) const {
        const uint32_t so = (uint32_t)(x0 + s) * plane_b + (uint32_t)y * row_b + 4u * (uint32_t)z0;
        const MemPriority prio;
        if constexpr (FAR) {
#pragma unroll
            for (int f = 0; f < 5; ++f) out_ptr[f * nc + ((size_t)so + zbo) / 4] = w[f];
        } else {
#pragma unroll
            for (int f = 0; f < 5; ++f) buf_store(rout[f], zbo, so, w[f]);
        }
    }
};

// ---------------------------------------------------------------------------
// LineCtx: a workgroup of W waves (W <= 16) holds whole periodic z-lines
// (nz <= 64 W), so there are no z-halo columns and every row segment a wave
// loads or stores starts at its z-line's start. With W = 1 and nz <= 32 the
// wave packs G = 64/nz groups of lanes, each on its own TX x-columns. Lane
// gl = (group g, z); z-neighbours are exchanged through LDS with periodic
// indices, two barriers per cell of a row (all waves of the group run the
// same rows). Loads and stores are indexed buffer accesses: a per-lane cell index
// (x-row and z, fixed per tile) plus a scalar byte offset for the y-row.
// ---------------------------------------------------------------------------
template <int TX>
struct LineCtx {
    __amdgpu_buffer_rsrc_t rin[5], rout[5];
    uint32_t xi[TX + 2];  // per-lane cell indices of the loaded columns in row 0 (x-row, z; wrapped/clamped)
    uint32_t xs_i;        // per-lane cell index of the lane's own cell in row 0 (unclamped x, z)
    int ny, y0, nrows, open_bc;
    uint32_t row_b, plane_b;
    int lidx, ridx, gl;   // LDS slots: left / right neighbour, own
    int xs0;              // first x-column of this lane's group
    int nx;
    int lanes;            // 64 W: LDS stride of one quantity
    bool valid;
    float *lq;            // LDS: kLds x 64W floats, 16-byte aligned
    static constexpr int kLq = 12;                // sweep quantities per lane
    static constexpr int kLds = kLq + 5;          // ... and the lower-interface flux

    __device__ __forceinline__ void load(int j, uint32_t yoff, float (&w)[5]) const {
        line_load_priority<TX>(j, true);
#pragma unroll
        for (int f = 0; f < 5; ++f) w[f] = tf_struct_load_f32(rin[f], (int)xi[j], 0, (int)yoff, 0);
        line_load_priority<TX>(j, false);
    }
    // The Z sweep with two LDS exchanges per column: every lane publishes its
    // sweep quantities (three consecutive 16-byte vectors), and after a
    // barrier reads its left neighbour's and evaluates the HLL flux at its
    // lower interface, once per interface; it publishes that flux (a 16-byte
    // vector and a scalar, conflict-free by lane) and after a second barrier
    // reads its right neighbour's, the flux at its upper interface. The
    // right neighbour evaluated it with the same operands in the same order
    // as this lane would, so the results equal those of evaluating both
    // fluxes per lane (one barrier, one more HLL per cell), and with one slot
    // of each kind: a lane writes the next column's quantities only after
    // the second barrier, when every lane has read this column's.
    __device__ __forceinline__ void zsweep(const Q &q, float k, float (&out)[5], int) const {
        float4 *own = reinterpret_cast<float4 *>(lq + gl * kLq);
        own[0] = make_float4(q.U[0], q.U[1], q.U[2], q.U[3]);
        own[1] = make_float4(q.U[4], q.F[0], q.F[1], q.F[2]);
        own[2] = make_float4(q.F[3], q.F[4], q.up, q.um);
        __syncthreads();
        const float4 *L4 = reinterpret_cast<const float4 *>(lq + lidx * kLq);
        const float4 la = L4[0], lb = L4[1], lc = L4[2];
        const float Lu[5] = {la.x, la.y, la.z, la.w, lb.x};
        const float Lf[5] = {lb.y, lb.z, lb.w, lc.x, lc.y};
        const HllCoef cl = hll_coef(fmax3(lc.z, q.up, 0.0f), fmin3(lc.w, q.um, 0.0f));
        float fl[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) fl[i] = hll_comp(cl, q.U[i] - Lu[i], cl.a * Lf[i], q.F[i]);
        float4 *f4 = reinterpret_cast<float4 *>(lq + kLq * lanes);  // fluxes 0..3, one vector per lane
        float *f1 = lq + (kLq + 4) * lanes;                           // flux 4
        f4[gl] = make_float4(fl[0], fl[1], fl[2], fl[3]);
        f1[gl] = fl[4];
        __syncthreads();
        const float4 r4 = f4[ridx];
        const float fr[5] = {r4.x, r4.y, r4.z, r4.w, f1[ridx]};
#pragma unroll
        for (int i = 0; i < 5; ++i) out[i] = fmaf(k, fl[i] - fr[i], q.U[i]);
    }
    __device__ __forceinline__ bool out_ok(int s) const { return valid && xs0 + s < nx; }
    __device__ __forceinline__ void store(int s, int y, const float (&w)[5]) const {
        const uint32_t so = (uint32_t)y * row_b + (uint32_t)s * plane_b;
        const MemPriority prio;
#pragma unroll
        for (int f = 0; f < 5; ++f) tf_struct_store_f32(w[f], rout[f], (int)xs_i, 0, (int)so, kStoreAux);
    }
};

// Sweep quantities along x of a loaded primitive cell.
__device__ __forceinline__ Q q_from_prim_x(const float (&w)[5], const Gas &g) {
    float U[5], rinv;
    prim_to_cons(w, g, U, rinv);
    const float v[3] = {w[1], w[2], w[3]};
    return make_q<1>(U, v, w[4], rinv, g);
}

// One row of the stream: X sweep of row y0-1+r, its Y flux against the
// previous row (qa/fa in), the Y update + Z sweep + store of the previous
// row, and the reload of every loaded column with row r+1. Writes the new
// Y-stream state to qb/fb.
template <int TX, class Ctx>
__device__ __forceinline__ void stream_row(const Ctx &T, const Gas &g, int r, float (&cur)[TX + 2][5],
                                           const QS (&qa)[TX], const float (&fa)[TX][5], QS (&qb)[TX],
                                           float (&fb)[TX][5], uint32_t &vmax) {
    const int yn = r + 1 < T.nrows ? T.y0 + r : T.y0 + r - 1;  // the last row reloads itself
    const int ygn = T.open_bc ? open_clamp(wrap1(yn, T.ny), T.ny) : wrap1(yn, T.ny);
    const uint32_t yoffn = (uint32_t)__builtin_amdgcn_readfirstlane(ygn) * T.row_b;
    Q qx0 = q_from_prim_x(cur[0], g);
    T.load(0, yoffn, cur[0]);
    Q qx1 = q_from_prim_x(cur[1], g);
    T.load(1, yoffn, cur[1]);
    float fx0[5];
    hll(qx0, qx1, fx0);
#pragma unroll
    for (int s = 0; s < TX; ++s) {
        const Q qx2 = q_from_prim_x(cur[s + 2], g);
        T.load(s + 2, yoffn, cur[s + 2]);
        float fx1[5];
        hll(qx1, qx2, fx1);
        float U1[5];
        godunov(qx1.U, fx0, fx1, g.k, U1);
        const Prim w1 = cons_to_prim(U1, g);
        qb[s] = make_qs<2>(U1, w1.v, w1.p, w1.rinv, g);
        if (r >= 1) hll(expand_qs<2>(qa[s]), expand_qs<2>(qb[s]), fb[s]);  // Y flux: previous row | this row
        if (r >= 2) {
            float U2[5];
            godunov(qa[s].U, fa[s], fb[s], g.k, U2);  // the previous row is now Y-swept
            const Prim w2 = cons_to_prim(U2, g);
            const Q q2 = make_q<3>(U2, w2.v, w2.p, w2.rinv, g);
            float U3[5];
            T.zsweep(q2, g.k, U3, s);
            // U3 is complete here on every lane: no exec-mask change separates
            // the lane shifts above from their uses (which lets them fold into
            // the arithmetic); only the output lanes continue below.
#pragma unroll
            for (int i = 0; i < 5; ++i) pin_vgpr(U3[i]);
            if (T.out_ok(s)) {
                const Prim w3 = cons_to_prim_out(U3, g);
                const float wo[5] = {w3.r, w3.v[0], w3.v[1], w3.v[2], w3.p};
                T.store(s, T.y0 + r - 2, wo);
                const float sp = signal_speed(w3.v, sound(w3.p, w3.rinv, g));
                vmax = max(vmax, __float_as_uint(sp));
            }
        }
        qx1 = qx2;
#pragma unroll
        for (int i = 0; i < 5; ++i) fx0[i] = fx1[i];
    }
}

// The stream over the rows of a tile, after the context is set up.
template <int TX, class Ctx>
__device__ __forceinline__ uint32_t stream_tile(const Ctx &T, const Gas &g, uint32_t yoff0) {
    // The loaded row: each column is reloaded with the next row right after
    // its last use, so a full row of work hides every load's latency.
    float cur[TX + 2][5];
#pragma unroll
    for (int j = 0; j < TX + 2; ++j) T.load(j, yoff0, cur[j]);
    // Y-stream state of the previous row, in two alternating sets so that
    // rows hand it over without register copies.
    QS qA[TX], qB[TX];
    float fA[TX][5], fB[TX][5];
    uint32_t vmax = 0;
    int r = 0;
    for (; r + 1 < T.nrows; r += 2) {
        stream_row<TX>(T, g, r, cur, qA, fA, qB, fB, vmax);
        stream_row<TX>(T, g, r + 1, cur, qB, fB, qA, fA, vmax);
    }
    if (r < T.nrows) stream_row<TX>(T, g, r, cur, qA, fA, qB, fB, vmax);
    return vmax;
}

// One wave of the halo tiling: the fused X->Y->Z step on its tile. Returns
// the wave-local max of the signal-speed bits of the cells it wrote.
template <int TX, bool FAR>
__device__ __forceinline__ uint32_t fused_tile(const float *in_, float *out_, int nx, int ny, int nz, int open_bc,
                                               const Gas &g_in, int xt, int zt, int y0, int rows) {
    const Gas g{sgpr_f(g_in.gamma), sgpr_f(g_in.gm1), sgpr_f(g_in.inv_gm1), sgpr_f(g_in.k)};
    // Everything below is wave-uniform; readfirstlane keeps it in SGPRs even
    // where it was read from memory with vector loads.
    nx = __builtin_amdgcn_readfirstlane(nx);
    ny = __builtin_amdgcn_readfirstlane(ny);
    nz = __builtin_amdgcn_readfirstlane(nz);
    open_bc = __builtin_amdgcn_readfirstlane(open_bc);
    xt = __builtin_amdgcn_readfirstlane(xt);
    zt = __builtin_amdgcn_readfirstlane(zt);
    y0 = __builtin_amdgcn_readfirstlane(y0);
    rows = __builtin_amdgcn_readfirstlane(rows);
    HaloCtx<TX, FAR> T;
    const size_t nc = (size_t)nx * ny * nz;
    {
        const float *pin = sgpr_ptr(in_);
        float *pout = sgpr_ptr(out_);
        T.in_ptr = pin;
        T.out_ptr = pout;
        T.nc = nc;
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            T.rin[f] = field_rsrc(pin + f * nc);
            T.rout[f] = field_rsrc(pout + f * nc);
        }
    }
    const int lane = lane_id();
    T.x0 = xt * TX;
    T.z0 = zt * kZT;
    T.y0 = y0;
    T.ny = ny;
    T.open_bc = open_bc;
    const int nzo = min(kZT, nz - T.z0);
    T.nrows = rows + 2;
    T.nxo = min(TX, nx - T.x0);
    T.row_b = 4u * (uint32_t)nz;
    T.plane_b = T.row_b * (uint32_t)ny;
    T.zb = 4u * (uint32_t)cell_index(T.z0 - 1 + lane, nz, open_bc);
    T.zbo = 4u * (uint32_t)(lane > 0 ? lane - 1 : 0);
    T.zout = lane >= 1 && lane <= nzo;
#pragma unroll
    for (int j = 0; j < TX + 2; ++j)
        T.xoff[j] = (uint32_t)__builtin_amdgcn_readfirstlane(cell_index(T.x0 - 1 + j, nx, open_bc)) * T.plane_b;
    const uint32_t yoff0 = (uint32_t)__builtin_amdgcn_readfirstlane(cell_index(T.y0 - 1, ny, open_bc)) * T.row_b;
    return stream_tile<TX>(T, g, yoff0);
}

// One workgroup of the line tiling (W waves, whole z-lines): returns the
// wave-local max of the signal-speed bits of the cells this wave wrote.
// `groups` x-column groups of nz lanes each, lane after lane across the
// workgroup's waves; lanes past groups * nz idle.
template <int TX>
__device__ __forceinline__ uint32_t fused_line(const float *in_, float *out_, int nx, int ny, int nz, int open_bc,
                                               const Gas &g_in, int W, int groups, int xt, int yc, int ly,
                                               float *lq) {
    const Gas g{sgpr_f(g_in.gamma), sgpr_f(g_in.gm1), sgpr_f(g_in.inv_gm1), sgpr_f(g_in.k)};
    nx = __builtin_amdgcn_readfirstlane(nx);
    ny = __builtin_amdgcn_readfirstlane(ny);
    nz = __builtin_amdgcn_readfirstlane(nz);
    open_bc = __builtin_amdgcn_readfirstlane(open_bc);
    groups = __builtin_amdgcn_readfirstlane(groups);
    xt = __builtin_amdgcn_readfirstlane(xt);
    yc = __builtin_amdgcn_readfirstlane(yc);
    ly = __builtin_amdgcn_readfirstlane(ly);
    LineCtx<TX> T;
    T.lanes = 64 * W;
    const size_t nc = (size_t)nx * ny * nz;
    {
        const float *pin = sgpr_ptr(in_);
        float *pout = sgpr_ptr(out_);
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            T.rin[f] = field_rsrc_indexed(pin + f * nc);
            T.rout[f] = field_rsrc_indexed(pout + f * nc);
        }
    }
    const int gl = (int)threadIdx.x;  // 0 .. 64 W - 1
    int grp = gl / nz, z = gl - grp * nz;
    T.valid = grp < groups;
    if (!T.valid) {  // idle lanes mirror lane 0 (valid addresses, never stored)
        grp = 0;
        z = 0;
    }
    T.gl = gl;
    T.lidx = grp * nz + (z == 0 ? nz - 1 : z - 1);
    T.ridx = grp * nz + (z == nz - 1 ? 0 : z + 1);
    T.nx = nx;
    T.ny = ny;
    T.y0 = yc * ly;
    T.open_bc = open_bc;
    T.nrows = min(ly, ny - T.y0) + 2;
    T.row_b = 4u * (uint32_t)nz;
    const uint32_t plane_b = T.row_b * (uint32_t)ny;
    T.plane_b = plane_b;
    T.xs0 = xt * groups * TX + grp * TX;
    const uint32_t plane_c = (uint32_t)ny * (uint32_t)nz;  // cells per x-plane
    T.xs_i = (uint32_t)min(T.xs0, nx - 1) * plane_c + (uint32_t)z;  // stores only where xs0 + s < nx
    const uint32_t zc = (uint32_t)cell_index(z, nz, open_bc);
#pragma unroll
    for (int j = 0; j < TX + 2; ++j) T.xi[j] = (uint32_t)cell_index(T.xs0 - 1 + j, nx, open_bc) * plane_c + zc;
    T.lq = lq;
    const uint32_t yoff0 = (uint32_t)__builtin_amdgcn_readfirstlane(cell_index(T.y0 - 1, ny, open_bc)) * T.row_b;
    return stream_tile<TX>(T, g, yoff0);
}

// Wave index of a halo-tiled launch with the XCD-aware block order
// (xcd_logical_block): -1 for the padding blocks of a grid rounded up to a
// multiple of 8. decode_tile then orders a simulation's waves z-tiles
// fastest, so the four z-tiles of a workgroup share their boundary lines in
// L1/L2 (orders with x-tiles or y-chunks inside a workgroup measured slower).
__device__ __forceinline__ int xcd_wave_index(int total_waves) {
    constexpr int kWavesPerBlock = kBlock / kWave;
    const int nblocks = (total_waves + kWavesPerBlock - 1) / kWavesPerBlock;
    const int lb = xcd_logical_block((int)blockIdx.x, nblocks);
    if (lb < 0) return -1;
    const int w = lb * kWavesPerBlock + wave_id_uniform();
    return w < total_waves ? w : -1;
}

static inline unsigned xcd_grid(int total_waves) {
    const int nblocks = (total_waves + (kBlock / kWave) - 1) / (kBlock / kWave);
    return (unsigned)(((nblocks + 7) / 8) * 8);
}

// Waves per SIMD the register allocation must allow: the stream state grows
// with TX, and below two resident waves nothing hides the load latency.
template <int TX>
constexpr int waves_per_simd() { return TX <= 2 ? 4 : (TX <= 4 ? 3 : (TX <= 6 ? 2 : 1)); }

// Maximum of v over a 256-thread block, valid in thread 0 (the block's
// four waves meet in LDS: one atomic per block instead of per wave).
__device__ __forceinline__ uint32_t block_umax256(uint32_t v) {
    __shared__ uint32_t part[4];
    v = wave_umax(v);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    return max(max(part[0], part[1]), max(part[2], part[3]));
}

// A wave's or workgroup's signal-speed maximum of the state it wrote (the
// step read parity par): into the plan's per-XCD slot of the simulation,
// else into the control block. Device-scope atomics to one address
// serialise across the XCDs (~20 ns each); one line per XCD keeps a step's
// atomics from queueing on a single address.
__device__ __forceinline__ void publish_max(const SimDev &S, Ctrl *C, int par, uint32_t v) {
    if (S.skip_max) return;
    uint32_t *maxp = S.maxp;
    if (maxp)
        atomicMax(maxp + ((par ^ 1) * kMaxSlots + (int)(blockIdx.x & (kMaxSlots - 1))) * kSlotStride, v);
    else
        atomicMax(&C->max_bits[par ^ 1], v);
}

// Fused step of one grid with an explicit dt (the _evolve_state entry point).
template <int TX>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(waves_per_simd<TX>()))) void k_step_direct(
    SimDev S, const float *__restrict__ in, float *__restrict__ out, const double *__restrict__ dt, double dx,
    uint32_t *__restrict__ bits_out) {
    const int w = xcd_wave_index(S.waves);
    if (w < 0) return;
    Gas g{S.gamma, S.gm1, S.inv_gm1, (float)(*dt / dx)};
    int xt, zt, yc;
    decode_tile(w, S.ntz, S.ntx, xt, zt, yc);
    if (S.far && touches_last_row(S, TX, xt, yc)) return;  // k_step_direct_far does it
    const int y0 = yc * S.ly;
    uint32_t v = fused_tile<TX, false>(in, out, S.nx, S.ny, S.nz, S.open_bc, g, xt, zt, y0, min(S.ly, S.ny - y0));
    if (bits_out) {
        v = wave_umax(v);
        if (lane_id() == 0) atomicMax(bits_out, v);
    }
}

// The far tiles of k_step_direct: one wave per (listed tile, z-tile).
template <int TX>
__global__ __launch_bounds__(kBlock) void k_step_direct_far(SimDev S, FarList F, const float *__restrict__ in,
                                                            float *__restrict__ out, const double *__restrict__ dt,
                                                            double dx, uint32_t *__restrict__ bits_out) {
    const int w = blockIdx.x * (kBlock / kWave) + wave_id_uniform();
    if (w >= F.n * S.ntz) return;
    const FarTile &t = F.t[w / S.ntz];
    Gas g{S.gamma, S.gm1, S.inv_gm1, (float)(*dt / dx)};
    uint32_t v = fused_tile<TX, true>(in, out, S.nx, S.ny, S.nz, S.open_bc, g, t.xt, w % S.ntz, t.y0, t.rows);
    if (bits_out) {
        v = wave_umax(v);
        if (lane_id() == 0) atomicMax(bits_out, v);
    }
}

// Fused step of every active simulation of a table; simulation s owns waves
// [prefix[s], prefix[s+1]). Input/output buffers follow the step parity.
template <int TX>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(waves_per_simd<TX>()))) void k_step_sims(
    const SimDev *__restrict__ sims, const int *__restrict__ prefix, int nsims) {
    const int w = xcd_wave_index(prefix[nsims]);
    if (w < 0) return;
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
    if (S.far && touches_last_row(S, TX, xt, yc)) return;  // k_step_sims_far does it
    const int y0 = yc * S.ly;
    uint32_t v = fused_tile<TX, false>(S.buf[par], S.buf[par ^ 1], S.nx, S.ny, S.nz, S.open_bc, g, xt, zt, y0,
                                       min(S.ly, S.ny - y0));
    v = wave_umax(v);
    if (lane_id() == 0) publish_max(S, C, par, v);
}

// The far tiles of k_step_sims: `far` lists (simulation, x-tile, row range);
// one wave per (listed tile, z-tile), max_ntz waves per entry.
template <int TX>
__global__ __launch_bounds__(kBlock) void k_step_sims_far(const SimDev *__restrict__ sims,
                                                          const FarTile *__restrict__ far, int nfar, int max_ntz) {
    const int w = blockIdx.x * (kBlock / kWave) + wave_id_uniform();
    if (w >= nfar * max_ntz) return;
    const FarTile t = far[w / max_ntz];
    const SimDev &S = sims[t.sim];
    const int zt = w % max_ntz;
    if (zt >= S.ntz) return;
    Ctrl *C = S.ctrl;
    if (!C->active) return;
    const int par = (int)(C->n_iter & 1);
    Gas g{S.gamma, S.gm1, S.inv_gm1, C->k};
    uint32_t v = fused_tile<TX, true>(S.buf[par], S.buf[par ^ 1], S.nx, S.ny, S.nz, S.open_bc, g, t.xt, zt, t.y0,
                                      t.rows);
    v = wave_umax(v);
    if (lane_id() == 0) publish_max(S, C, par, v);
}

// Line tiling: block b of a class of simulations with W waves per z-line;
// simulation s owns blocks [prefix[s], prefix[s+1]) (XCD-aware order).
// The line tiling keeps per-lane column offsets (TX + 2 VGPRs) on top of the
// stream state, so it allows one resident wave fewer than the halo tiling.
template <int TX>
constexpr int line_waves_per_simd() { return TX <= 2 ? 4 : (TX <= 4 ? 2 : 1); }

// Workgroup lb of a line class -> (simulation, x-tile, y-chunk) and the step.
template <int TX>
__device__ __forceinline__ void line_block(const SimDev *__restrict__ sims, const int *__restrict__ prefix, int nsims,
                                           int W, float *lq) {
    const int lb = xcd_logical_block((int)blockIdx.x, prefix[nsims]);
    if (lb < 0) return;  // whole workgroup: uniform, before any barrier
    int lo = 0, hi = nsims;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= lb) lo = mid; else hi = mid;
    }
    const SimDev &S = sims[lo];
    Ctrl *C = S.ctrl;
    if (!C->active) return;  // uniform over the workgroup
    const int par = (int)(C->n_iter & 1);
    Gas g{S.gamma, S.gm1, S.inv_gm1, C->k};
    int xt, yc;
    decode_line_block(lb - prefix[lo], S.ntx, xt, yc);
    uint32_t v = fused_line<TX>(S.buf[par], S.buf[par ^ 1], S.nx, S.ny, S.nz, S.open_bc, g, W, S.groups, xt, yc,
                                S.ly, lq);
    // one device-scope atomic per workgroup: atomics to the control line
    // serialise across the XCDs (1% of a 1024^3 step with one per wave)
    v = wave_umax(v);
    __shared__ uint32_t wmax[16];
    if (lane_id() == 0) wmax[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < W; ++i) v = max(v, wmax[i]);
        publish_max(S, C, par, v);
    }
}

// Line tiling, W <= 4 waves per workgroup (W fixed): block b of a class of
// simulations; simulation s owns blocks [prefix[s], prefix[s+1]) (XCD-aware
// order). The line tiling keeps per-lane column offsets (TX + 2 VGPRs) on
// top of the stream state, so it allows one resident wave fewer than the
// halo tiling.
template <int TX, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(line_waves_per_simd<TX>()))) void k_line_sims(
    const SimDev *__restrict__ sims, const int *__restrict__ prefix, int nsims) {
    __shared__ __attribute__((aligned(16))) float lq[LineCtx<TX>::kLds * 64 * W];
    line_block<TX>(sims, prefix, nsims, W, lq);
}

// Line tiling with 5 <= W <= MAXW waves per workgroup (W at run time): long
// z-lines (up to 64 MAXW cells); one workgroup per CU at MAXW = 16, two at
// <2, 8> (LDS and four waves per SIMD allow 16 waves per CU).
template <int TX, int MAXW>
__global__ __launch_bounds__(64 * MAXW) __attribute__((amdgpu_waves_per_eu(line_waves_per_simd<TX>()))) void k_line_wide(
    const SimDev *__restrict__ sims, const int *__restrict__ prefix, int nsims, int W) {
    __shared__ __attribute__((aligned(16))) float lq[LineCtx<TX>::kLds * 64 * MAXW];
    line_block<TX>(sims, prefix, nsims, W, lq);
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
    if (i >= nsims) return;
    const SimDev &S = sims[i];
    Ctrl *C = S.ctrl;
    if (S.maxp && C->active) {
        // the step just run wrote parity p: combine and clear its slots
        const int p = (int)((C->n_iter + 1) & 1);
        uint32_t *slot = S.maxp + p * kMaxSlots * kSlotStride;
        uint32_t bits = 0u;
        for (int k = 0; k < kMaxSlots; ++k) {
            bits = max(bits, slot[k * kSlotStride]);
            slot[k * kSlotStride] = 0u;
        }
        C->max_bits[p] = bits;
    }
    finalize_ctrl(C, 0);
}

// Table and control block of a single explicit-dt step through the
// whole-line kernels: the simulation (buf[0] = input, buf[1] = output) at
// step 0 with k = float(dt / dx), one class of `nblocks` workgroups.
__global__ void k_direct_setup(SimDev S, SimDev *dst, int *prefix, int nblocks, const double *dt, double dx) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    *dst = S;
    prefix[0] = 0;
    prefix[1] = nblocks;
    Ctrl *C = S.ctrl;
    C->n_iter = 0;
    C->active = 1;
    C->nonfinite = 0;
    C->max_bits[0] = 0u;
    C->max_bits[1] = 0u;
    C->k = (float)(*dt / dx);
}

// Max-accumulates the signal speed of the state a direct step wrote.
__global__ void k_direct_bits(const Ctrl *C, uint32_t *bits_out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) atomicMax(bits_out, C->max_bits[1]);
}

__global__ void k_finalize_one(Ctrl *C, int first) {
    if (threadIdx.x == 0 && blockIdx.x == 0) finalize_ctrl(C, first);
}

__device__ __forceinline__ void ctrl_reset(Ctrl *C, double dx, double c_cfl, double t_final, int64_t max_steps) {
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

__global__ void k_ctrl_reset(Ctrl *C, double dx, double c_cfl, double t_final, int64_t max_steps) {
    if (threadIdx.x == 0 && blockIdx.x == 0) ctrl_reset(C, dx, c_cfl, t_final, max_steps);
}

// tf_batch_reset: every simulation's control block, then the signal-speed
// maximum of every buf[0] (simulation s owns workgroups [bprefix[s],
// bprefix[s+1]) of kSpeedCellsPerBlock cells), then the first dt.
__global__ void k_reset_sims(const SimDev *__restrict__ sims, const ResetParams *__restrict__ rp, int nsims) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nsims) ctrl_reset(sims[i].ctrl, rp[i].dx, rp[i].c_cfl, rp[i].t_final, rp[i].max_steps);
}

__global__ __launch_bounds__(256) void k_speed_sims(const SimDev *__restrict__ sims, const int *__restrict__ bprefix,
                                                    int nsims) {
    const int b = (int)blockIdx.x;
    if (b >= bprefix[nsims]) return;
    int lo = 0, hi = nsims;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (bprefix[mid] <= b) lo = mid; else hi = mid;
    }
    const SimDev &S = sims[lo];
    const int64_t nc = (int64_t)S.nx * S.ny * S.nz;
    const float *prim = S.buf[0];
    const int64_t base = (int64_t)(b - bprefix[lo]) * kSpeedCellsPerBlock;
    uint32_t v = 0;
#pragma unroll
    for (int j = 0; j < kSpeedCellsPerBlock / 256; ++j) {
        const int64_t i = base + j * 256 + threadIdx.x;
        if (i < nc) {
            const float r = prim[i], vx = prim[nc + i], vy = prim[2 * nc + i], vz = prim[3 * nc + i],
                        p = prim[4 * nc + i];
            const float c = sound(p, frcp_nr(r), Gas{S.gamma, 0.f, 0.f, 0.f});
            const float w[3] = {vx, vy, vz};
            v = max(v, __float_as_uint(signal_speed(w, c)));
        }
    }
    v = block_umax256(v);
    if (threadIdx.x == 0) atomicMax(&S.ctrl->max_bits[0], v);
}

__global__ void k_finalize_first_sims(const SimDev *__restrict__ sims, int nsims) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nsims) finalize_ctrl(sims[i].ctrl, 1);
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Flat 16-byte-per-thread copy: one read and one write of every byte, the
// access order that reaches the highest HBM throughput measured on gfx950
// (one float4 per thread and workgroups of 256 beat grid-stride loops, wider
// unrolls and larger workgroups). The practical ceiling of a read+write kernel.
__global__ __launch_bounds__(256) void k_copy16(const float4 *__restrict__ src, float4 *__restrict__ dst, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// Max over cells of max_d |v_d| + c (the CFL signal speed, fluid.py:425-500).
__global__ __launch_bounds__(256) void k_speed(const float *__restrict__ prim, int64_t nc, float gamma,
                                               uint32_t *__restrict__ bits_out) {
    uint32_t v = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += stride) {
        const float r = prim[i], vx = prim[nc + i], vy = prim[2 * nc + i], vz = prim[3 * nc + i],
                    p = prim[4 * nc + i];
        const float c = sound(p, frcp_nr(r), Gas{gamma, 0.f, 0.f, 0.f});
        const float w[3] = {vx, vy, vz};
        v = max(v, __float_as_uint(signal_speed(w, c)));
    }
    v = block_umax256(v);
    if (threadIdx.x == 0) atomicMax(bits_out, v);
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
    const Prim w1 = cons_to_prim_out(U1, g);
    out[i] = w1.r;
    out[nc + i] = w1.v[0];
    out[2 * nc + i] = w1.v[1];
    out[3 * nc + i] = w1.v[2];
    out[4 * nc + i] = w1.p;
}

// ---------------------------------------------------------------- elementwise
// Source index along one axis of the open boundaries with g ghost cells
// (fluid.py:334-380): the left copies (cells [0, g) take cell g) run first,
// then the right ones (cells [n-g, n) take cell n-g-1, as the left copies
// left it). Needs n >= g + 1; g = 1 is open_clamp.
__device__ __forceinline__ int open_src(int i, int n, int g) {
    const int j = i >= n - g ? n - g - 1 : i;
    return j < g ? g : j;
}

__global__ __launch_bounds__(256) void k_open_bc(const float *__restrict__ in, float *__restrict__ out, int nx, int ny,
                                                 int nz, int g) {
    const int64_t nc = (int64_t)nx * ny * nz;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const int z = (int)(i % nz);
    const int64_t xy = i / nz;
    const int y = (int)(xy % ny);
    const int x = (int)(xy / ny);
    const int64_t src = ((int64_t)open_src(x, nx, g) * ny + open_src(y, ny, g)) * nz + open_src(z, nz, g);
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
        const Prim w = cons_to_prim_out(U, g);
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
    const dim3 grid(xcd_grid(S.waves)), block(kBlock);
    switch (tx) {
        case 2: hipLaunchKernelGGL(k_step_direct<2>, grid, block, 0, st, S, in, out, dt, dx, bits); break;
        case 4: hipLaunchKernelGGL(k_step_direct<4>, grid, block, 0, st, S, in, out, dt, dx, bits); break;
        case 6: hipLaunchKernelGGL(k_step_direct<6>, grid, block, 0, st, S, in, out, dt, dx, bits); break;
        case 8: hipLaunchKernelGGL(k_step_direct<8>, grid, block, 0, st, S, in, out, dt, dx, bits); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_step_direct_far(int tx, const SimDev &S, const FarList &F, const float *in, float *out,
                                  const double *dt, double dx, uint32_t *bits, hipStream_t st) {
    if (F.n <= 0) return hipSuccess;
    const dim3 grid((unsigned)((F.n * S.ntz + 3) / 4)), block(kBlock);
    switch (tx) {
        case 2: hipLaunchKernelGGL(k_step_direct_far<2>, grid, block, 0, st, S, F, in, out, dt, dx, bits); break;
        case 4: hipLaunchKernelGGL(k_step_direct_far<4>, grid, block, 0, st, S, F, in, out, dt, dx, bits); break;
        case 6: hipLaunchKernelGGL(k_step_direct_far<6>, grid, block, 0, st, S, F, in, out, dt, dx, bits); break;
        case 8: hipLaunchKernelGGL(k_step_direct_far<8>, grid, block, 0, st, S, F, in, out, dt, dx, bits); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_step_sims_far(int tx, const SimDev *sims, const FarTile *far, int nfar, int max_ntz,
                                hipStream_t st) {
    if (nfar <= 0) return hipSuccess;
    const dim3 grid((unsigned)((nfar * max_ntz + 3) / 4)), block(kBlock);
    switch (tx) {
        case 2: hipLaunchKernelGGL(k_step_sims_far<2>, grid, block, 0, st, sims, far, nfar, max_ntz); break;
        case 4: hipLaunchKernelGGL(k_step_sims_far<4>, grid, block, 0, st, sims, far, nfar, max_ntz); break;
        case 6: hipLaunchKernelGGL(k_step_sims_far<6>, grid, block, 0, st, sims, far, nfar, max_ntz); break;
        case 8: hipLaunchKernelGGL(k_step_sims_far<8>, grid, block, 0, st, sims, far, nfar, max_ntz); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int W>
static hipError_t launch_line_w(int tx, const SimDev *sims, const int *prefix, int nsims, int nblocks,
                                hipStream_t st) {
    const dim3 grid((unsigned)(((nblocks + 7) / 8) * 8)), block(64 * W);
    switch (tx) {
        case 2: hipLaunchKernelGGL((k_line_sims<2, W>), grid, block, 0, st, sims, prefix, nsims); break;
        case 4: hipLaunchKernelGGL((k_line_sims<4, W>), grid, block, 0, st, sims, prefix, nsims); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Line classes: W <= 4 with tile width 2 or 4; 5 <= W <= 8 with 4 (two waves
// per SIMD); 9 <= W <= 16 with 2 (four waves per SIMD).
hipError_t launch_line_sims(int tx, int w, const SimDev *sims, const int *prefix, int nsims, int nblocks,
                            hipStream_t st) {
    if (nsims <= 0 || nblocks <= 0) return hipSuccess;
    const dim3 grid((unsigned)(((nblocks + 7) / 8) * 8)), block(64 * w);
    switch (w) {
        case 1: return launch_line_w<1>(tx, sims, prefix, nsims, nblocks, st);
        case 2: return launch_line_w<2>(tx, sims, prefix, nsims, nblocks, st);
        case 3: return launch_line_w<3>(tx, sims, prefix, nsims, nblocks, st);
        case 4: return launch_line_w<4>(tx, sims, prefix, nsims, nblocks, st);
        default: break;
    }
    // 5..8 waves: TX=4 (two waves per SIMD) or TX=2 with two workgroups per
    // CU (four waves per SIMD; 384-cell lines run 23% faster that way)
    if (w <= 8 && tx == 4)
        hipLaunchKernelGGL((k_line_wide<4, 8>), grid, block, 0, st, sims, prefix, nsims, w);
    else if (w <= 8 && tx == 2)
        hipLaunchKernelGGL((k_line_wide<2, 8>), grid, block, 0, st, sims, prefix, nsims, w);
    else if (w <= 16 && tx == 2)
        hipLaunchKernelGGL((k_line_wide<2, 16>), grid, block, 0, st, sims, prefix, nsims, w);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_step_sims(int tx, const SimDev *sims, const int *prefix, int nsims, int total_waves,
                            hipStream_t st) {
    if (total_waves <= 0) return hipSuccess;
    const dim3 grid(xcd_grid(total_waves)), block(kBlock);
    switch (tx) {
        case 2: hipLaunchKernelGGL(k_step_sims<2>, grid, block, 0, st, sims, prefix, nsims); break;
        case 4: hipLaunchKernelGGL(k_step_sims<4>, grid, block, 0, st, sims, prefix, nsims); break;
        case 6: hipLaunchKernelGGL(k_step_sims<6>, grid, block, 0, st, sims, prefix, nsims); break;
        case 8: hipLaunchKernelGGL(k_step_sims<8>, grid, block, 0, st, sims, prefix, nsims); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_direct_setup(const SimDev &S, SimDev *dst, int *prefix, int nblocks, const double *dt, double dx,
                               hipStream_t st) {
    hipLaunchKernelGGL(k_direct_setup, dim3(1), dim3(64), 0, st, S, dst, prefix, nblocks, dt, dx);
    return hipGetLastError();
}

hipError_t launch_direct_bits(const Ctrl *C, uint32_t *bits_out, hipStream_t st) {
    hipLaunchKernelGGL(k_direct_bits, dim3(1), dim3(64), 0, st, C, bits_out);
    return hipGetLastError();
}

hipError_t launch_reset_sims(const SimDev *sims, const ResetParams *rp, const int *bprefix, int nsims,
                             int speed_blocks, hipStream_t st) {
    hipLaunchKernelGGL(k_reset_sims, grid1(nsims, 64), dim3(64), 0, st, sims, rp, nsims);
    if (speed_blocks > 0)
        hipLaunchKernelGGL(k_speed_sims, dim3((unsigned)speed_blocks), dim3(256), 0, st, sims, bprefix, nsims);
    hipLaunchKernelGGL(k_finalize_first_sims, grid1(nsims, 64), dim3(64), 0, st, sims, nsims);
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

hipError_t launch_copy16(const void *src, void *dst, int64_t n16, hipStream_t st) {
    // at most 2^31 threads per launch
    constexpr int64_t kChunk = (int64_t)1 << 31;
    for (int64_t o = 0; o < n16; o += kChunk) {
        const int64_t n = n16 - o < kChunk ? n16 - o : kChunk;
        hipLaunchKernelGGL(k_copy16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                           static_cast<const float4 *>(src) + o, static_cast<float4 *>(dst) + o, n);
    }
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

hipError_t launch_open_bc(const float *in, float *out, int nx, int ny, int nz, int g, hipStream_t st) {
    const int64_t nc = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(k_open_bc, grid1(nc), dim3(256), 0, st, in, out, nx, ny, nz, g);
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
