// Host-side planning of the fused stencil: tilings, the tables of a batch
// plan and where they sit in the caller's workspace. No HIP calls and no
// device memory access (device addresses are only computed), so this part
// also builds with the host compiler alone, under address and undefined-
// behaviour sanitizers.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/tinyfluids.h"
#include "tf_layout.h"

namespace tf {

constexpr int kMaxLineW = 16;            // whole z-lines of up to 1024 cells per workgroup
constexpr int kClasses = kMaxLineW + 1;  // halo tiling + line tiling with W = 1..16

// Argument checks: TF_OK, or an error code with its message in err.
int check_dims(int nx, int ny, int nz, std::string &err);
int check_fused_size(int nx, int ny, int nz, std::string &err);
int check_open_bc(int open_bc, int nx, int ny, int nz, std::string &err);
bool valid_tx(int tx);

// Default tilings.
int default_tile_x(int nx);
int default_tile_y(int nx, int ny, int nz, int tx);
int default_loop_tx(int nx, int ny, int nz);

// Halo-tiled geometry of one simulation (device pointers left null).
SimDev make_simdev(int nx, int ny, int nz, int open_bc, double gamma, int tx, int ly);

// The halo tiles of S that touch the last (x, y) row of a field of exactly
// 2^30 cells (the device-side predicate touches_last_row, evaluated for
// every tile), appended for simulation `sim`. With `split`, each tile's rows
// are cut into up to kFarSplit ranges of at least 4 rows.
void far_tiles(const SimDev &S, int tx, int sim, std::vector<FarTile> &out, bool split);
// The same tiles unsplit, as the single-step kernels take them; an error if
// they do not fit a FarList.
int far_list(const SimDev &S, int tx, FarList &out, std::string &err);

// Line tiling of a z-line of nz cells: w waves per workgroup holding
// `groups` whole lines.
void line_shape(int nz, int &w, int &groups);
double halo_lane_use(int nz);
double line_lane_use(int nz);
// Tile width of a line class of w waves for a requested tile width tx.
int line_tx(int w, int tx);
// Turns S into the line tiling of its z-lines with requested tile width tx.
void make_line(SimDev &S, int tx);

// Workspace bytes of a plan over nsims simulations (tf_batch_workspace_bytes).
size_t plan_workspace_bytes(int nsims);

// One tiling class of a plan: one launch over the tiles of `count`
// simulations starting at table index `begin` (w = 0: halo tiling, prefix
// over waves; w >= 1: line tiling with w waves per workgroup, prefix over
// workgroups). The class's prefix starts at prefix[prefix_at].
struct PlanClass {
    int w = 0, tx = 0, begin = 0, count = 0, total = 0, prefix_at = 0;
};

struct PlanRequest {
    const tf_sim_desc *sims = nullptr;
    int nsims = 0;
    uintptr_t workspace = 0;  // device address of the workspace (never dereferenced)
    size_t workspace_bytes = 0;
    int tile_x = 0, tile_y = 0;            // 0: plan defaults
    const int *tuned_ly = nullptr;         // per simulation: measured y-chunk rows, 0 for none (or null)
    int tuned_tx = 0;                      // measured tile width of the largest simulation, 0 for none
    const unsigned char *line = nullptr;   // per simulation: 1 runs whole z-lines (null: none does)
};

struct PlanTables {
    int nsims = 0, tx = 0;
    long total_waves = 0;
    std::vector<SimDev> sims;  // device table order, with device pointers
    std::vector<int> prefix;   // every class's prefix, concatenated
    std::vector<int> order;    // sims[k] is the caller's simulation order[k]
    PlanClass cls[kClasses];
    std::vector<FarTile> far;  // far-row tiles (table indices)
    int far_max_ntz = 0;
    int speed_blocks = 0;      // workgroups of the reset's signal-speed kernel
    int nlaunch = 0;           // non-empty stencil launches per step (classes + far rows)
    // Byte offsets in the workspace: the tables from off_sims to
    // off_tables_end are uploaded as one host image; the far tiles follow,
    // then the signal-speed slots (cleared on the device).
    size_t off_sims = 0, off_prefix = 0, off_status = 0, off_reset = 0, off_speed = 0, off_tables_end = 0,
           off_far = 0, off_maxp = 0, bytes = 0;
    std::vector<char> image;
};

// Builds every table of a plan and its upload image. TF_OK or an error code
// (TF_ERR_ARG, TF_ERR_UNSUPPORTED) with its message in err.
int build_plan(const PlanRequest &rq, PlanTables &out, std::string &err);

}  // namespace tf
