// Device-memory layout shared by the kernels, the C ABI and the host-side
// plan builder: control blocks, simulation tables, far-row tiles, the
// workspace header, and the tile decoders that map a wave or workgroup
// index to its tile. Plain C++ with no HIP dependency, so that the plan
// builder (tf_plan.cpp) also compiles with the host compiler alone (its
// sanitizer build); under hipcc the decoders are host and device functions.
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __HIP__
#define TF_HD __host__ __device__ inline __attribute__((always_inline))
#else
#define TF_HD inline
#endif

namespace tf {

constexpr int kWave = 64;
constexpr int kZT = kWave - 2;      // output z-columns per wave (lanes 1..62; lanes 0 and 63 are halo)
constexpr int kBlock = 256;         // threads per block of the stencil kernels (4 independent waves)
constexpr int kTotalsBlocks = 1024; // fixed grid of the float64 totals reduction (deterministic order)
constexpr int kXcd = 8;             // XCDs of the device; workgroup b runs on XCD b % 8

// Per-simulation control block in device memory. The first 48 bytes mirror
// tf_sim_status of the public header.
struct Ctrl {
    double t;           // simulated time of the current state
    double dt;          // dt of the next step
    int64_t n_iter;     // completed steps; the state is in buf[n_iter & 1]
    int32_t active;     // the next step runs
    int32_t nonfinite;  // a non-finite signal speed was seen
    float max_speed;    // max |v_d| + c of the current state
    int32_t pad0[3];
    uint32_t max_bits[2];  // signal-speed accumulators (float bits), by parity of the state written
    float k;               // float(dt / dx)
    int32_t pad1;
    double t_final;
    double dx;
    double c_cfl;
    int64_t max_steps;     // < 0: unlimited
    double pad2[4];
};
static_assert(sizeof(Ctrl) == 128, "Ctrl must match TF_CTRL_BYTES");

// Geometry and physics of one simulation as the stencil kernels see it.
struct SimDev {
    float *buf[2];
    Ctrl *ctrl;
    int nx, ny, nz, open_bc;
    float gamma, gm1, inv_gm1, pad;
    int ntx, ntz, nyc, ly;  // tiles along x (of TX), z (of kZT), y-chunks and their length
    int waves;              // ntx * ntz * nyc
    int far;                // fields of exactly 4 GiB: tiles touching the last (x, y) row run in k_step_far
    int line_w;             // 0: halo tiling (waves of 62 z-columns); W >= 1: line tiling, W waves per z-line
    int groups;             // line tiling: whole z-lines (x-column groups) per workgroup
    int skip_max;           // 1: nobody reads this step's signal-speed maximum (single step without max_speed_out)
    // Signal-speed accumulators of a plan's step: per parity of the state
    // written, kMaxSlots slots of one cache line each, one per XCD (block
    // index mod 8); the dt update combines and clears them. nullptr: the
    // step accumulates into ctrl->max_bits directly (single explicit step).
    uint32_t *maxp;
};
constexpr int kMaxSlots = 8;
constexpr int kSlotStride = 32;  // uint32 per slot: one 128-byte line
constexpr size_t kSlotBytesPerSim = 2 * kMaxSlots * kSlotStride * sizeof(uint32_t);

// A tile handled by the far-row kernel: simulation index, x-tile, and the
// rows [y0, y0 + rows) it writes.
struct FarTile {
    int sim, xt, y0, rows;
};
constexpr int kMaxFarPerSim = 9;  // at most 3 x-tiles x 3 y-chunks touch the last row
constexpr int kFarSplit = 8;      // batch plans split each far tile into up to 8 shorter row ranges

struct FarList {
    int n;
    int pad;
    FarTile t[kMaxFarPerSim];
};

// Loop bounds of one simulation of a plan (from its tf_sim_desc), for
// tf_batch_reset.
struct ResetParams {
    double dx, c_cfl, t_final;
    int64_t max_steps;
};
// Cells per workgroup of the batched signal-speed kernel (256 threads x 8).
constexpr int kSpeedCellsPerBlock = 2048;

// Workspace header of the single-grid entry points.
// sim/prefix/ctrl: the single explicit-dt step through the whole-line
// kernels; bits: signal-speed maximum of the CFL entry point; totals[0]: the
// dt the step tuner times with; ctrl also serves the single-grid loop.
struct WsHeader {
    SimDev sim;
    Ctrl ctrl;
    int prefix[2];
    uint32_t bits[4];
    double totals[5];
};

constexpr size_t kWsHeaderBytes = 1024;
static_assert(sizeof(WsHeader) <= kWsHeaderBytes, "workspace header too large");

// ------------------------------------------------------------ tile decoders
// Logical index of workgroup b in a grid of nblocks workgroups rounded up to
// a multiple of 8: workgroups are dispatched to the XCDs round-robin, each
// with its own L2, and XCD k gets the contiguous logical range [k*Q,
// (k+1)*Q), so neighbouring tiles, which share halo rows and cache lines,
// meet in one L2. -1 for the padding workgroups.
TF_HD int xcd_logical_block(int b, int nblocks) {
    const int q = (nblocks + kXcd - 1) / kXcd;
    const int lb = (b % kXcd) * q + b / kXcd;
    return lb < nblocks ? lb : -1;
}

// Halo tiling: wave `local` of a simulation -> (x-tile, z-tile, y-chunk),
// z-tiles fastest, then x-tiles, then y-chunks.
TF_HD void decode_tile(int local, int ntz, int ntx, int &xt, int &zt, int &yc) {
    zt = local % ntz;
    const int rest = local / ntz;
    xt = rest % ntx;
    yc = rest / ntx;
}

// Line tiling: workgroup `local` of a simulation -> (x-tile, y-chunk).
TF_HD void decode_line_block(int local, int ntx, int &xt, int &yc) {
    xt = local % ntx;
    yc = local / ntx;
}

// Does halo tile (xt, yc) of tile width tx load or store the last (x, y)
// row? Loaded x-rows are x0-1 .. x0+tx (the first tile's left halo wraps to
// nx-1), loaded y-rows y0-1 .. min(y0+ly, ny) (the first chunk's lower halo
// wraps to ny-1).
TF_HD bool touches_last_row(const SimDev &S, int tx, int xt, int yc) {
    const bool xr = xt == 0 || xt * tx + tx >= S.nx - 1;
    const bool yr = yc == 0 || yc * S.ly + S.ly >= S.ny - 1;
    return xr && yr;
}

}  // namespace tf
