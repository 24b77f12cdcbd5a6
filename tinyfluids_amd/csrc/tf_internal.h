// Internal host/device structures shared by the kernels and the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "tf_device.h"

namespace tf {

constexpr int kWave = 64;
constexpr int kZT = kWave - 2;      // output z-columns per wave (lanes 1..62; lanes 0 and 63 are halo)
constexpr int kBlock = 256;         // threads per block of the stencil kernels (4 independent waves)
constexpr int kTotalsBlocks = 1024; // fixed grid of the float64 totals reduction (deterministic order)

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

}  // namespace tf
