// C ABI of libtfhip.so (declared in include/tinyfluids.h): argument checks,
// tiling choice, the device-resident time loop and the batched engine.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/tinyfluids.h"
#include "tf_internal.h"
#include "tf_plan.h"

namespace tf {
hipError_t launch_step_direct(int tx, const SimDev &S, const float *in, float *out, const double *dt, double dx,
                              uint32_t *bits, hipStream_t st);
hipError_t launch_step_sims(int tx, const SimDev *sims, const int *prefix, int nsims, int total_waves,
                            hipStream_t st);
hipError_t launch_step_direct_far(int tx, const SimDev &S, const FarList &F, const float *in, float *out,
                                  const double *dt, double dx, uint32_t *bits, hipStream_t st);
hipError_t launch_step_sims_far(int tx, const SimDev *sims, const FarTile *far, int nfar, int max_ntz,
                                hipStream_t st);
hipError_t launch_line_sims(int tx, int w, const SimDev *sims, const int *prefix, int nsims, int nblocks,
                            hipStream_t st);
hipError_t launch_finalize_sims(const SimDev *sims, int nsims, hipStream_t st);
hipError_t launch_reset_sims(const SimDev *sims, const ResetParams *rp, const int *bprefix, int nsims,
                             int speed_blocks, hipStream_t st);
hipError_t launch_direct_setup(const SimDev &S, SimDev *dst, int *prefix, int nblocks, const double *dt, double dx,
                               hipStream_t st);
hipError_t launch_direct_bits(const Ctrl *C, uint32_t *bits_out, hipStream_t st);
hipError_t launch_finalize_one(Ctrl *C, int first, hipStream_t st);
hipError_t launch_ctrl_reset(Ctrl *C, double dx, double c_cfl, double t_final, int64_t max_steps, hipStream_t st);
hipError_t launch_fill_u32(uint32_t *p, uint32_t v, int n, hipStream_t st);
hipError_t launch_copy16(const void *src, void *dst, int64_t n16, hipStream_t st);
hipError_t launch_speed(const float *prim, int64_t nc, float gamma, uint32_t *bits, hipStream_t st);
hipError_t launch_dt_from_bits(const uint32_t *bits, double dx, double c_cfl, double *dt_out, float *max_out,
                               hipStream_t st);
hipError_t launch_sweep(int axis, const float *in, float *out, int nx, int ny, int nz, const Gas &g, const double *dt,
                        double dx, hipStream_t st);
hipError_t launch_open_bc(const float *in, float *out, int nx, int ny, int nz, int g, hipStream_t st);
hipError_t launch_p2c(const float *in, float *out, int64_t nc, const Gas &g, hipStream_t st);
hipError_t launch_c2p(const float *in, float *out, int64_t nc, const Gas &g, hipStream_t st);
hipError_t launch_totals(const float *prim, int64_t nc, double gamma, double *partials, double *out5,
                         hipStream_t st);
hipError_t launch_gather_status(const SimDev *sims, int nsims, Ctrl *dst, hipStream_t st);
hipError_t launch_sound_speed(const float *rho, const float *p, float *c, int64_t n, const Gas &g, hipStream_t st);
hipError_t launch_wave_speeds(const float *L, const float *R, float *sl, float *sr, int64_t n, const Gas &g, int dir,
                              uint32_t *bits, hipStream_t st);
hipError_t launch_euler_flux(const float *w, float *F, int64_t n, const Gas &g, int dir, hipStream_t st);
hipError_t launch_hll_flux(const float *L, const float *R, float *F, int64_t n, const Gas &g, int dir,
                           hipStream_t st);
hipError_t launch_set_along_axis(const float *in, float *out, int nx, int ny, int nz, int axis, int set_index,
                                 int get_index, hipStream_t st);
}  // namespace tf

using namespace tf;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define TF_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) return fail(TF_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_));         \
    } while (0)

#define TF_REQUIRE(cond, ...)                                   \
    do {                                                        \
        if (!(cond)) return fail(TF_ERR_ARG, __VA_ARGS__);      \
    } while (0)

inline hipStream_t S_(void *s) { return reinterpret_cast<hipStream_t>(s); }

// The two timing events of a tuner, destroyed on every return path.
struct EventPair {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t create() {
        const hipError_t e = hipEventCreate(&e0);
        return e == hipSuccess ? hipEventCreate(&e1) : e;
    }
    ~EventPair() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

// Argument checks of tf_plan.cpp, with the message kept for tf_last_error.
int req(int rc, const std::string &msg) { return rc == TF_OK ? TF_OK : fail(rc, "%s", msg.c_str()); }
int req_dims(int nx, int ny, int nz) {
    std::string e;
    return req(check_dims(nx, ny, nz, e), e);
}
int req_fused_size(int nx, int ny, int nz) {
    std::string e;
    return req(check_fused_size(nx, ny, nz, e), e);
}
int req_open_bc(int open_bc, int nx, int ny, int nz) {
    std::string e;
    return req(check_open_bc(open_bc, nx, ny, nz, e), e);
}

Gas make_gas(double gamma, float k = 0.0f) {
    Gas g;
    g.gamma = (float)gamma;
    g.gm1 = (float)gamma - 1.0f;  // gamma is float32 first, as in the reference's float32 runs
    g.inv_gm1 = 1.0f / g.gm1;
    g.k = k;
    return g;
}

// ------------------------------------------------------------- tiling registry
// Tilings measured or set per grid shape:
// - g_tiling: the single explicit-dt step (tf_autotune_evolve, tf_set_tiling);
// - g_loop_tiling: the plan that loops and batches run (tf_autotune_loop,
//   tf_set_tiling), tile_y 0 letting the plan choose.
std::mutex g_tiling_mu;
std::map<std::tuple<int, int, int>, std::pair<int, int>> g_tiling;

struct LoopTiling {
    int tx, ly;
    int line;  // 1: whole z-lines where the shape allows them; 0: 62-cell z-tiles
};
std::map<std::tuple<int, int, int>, LoopTiling> g_loop_tiling;
// Set by the loop tuner while it builds a candidate plan: -1 none, else 0/1.
thread_local int tl_line_force = -1;

void loop_tiling_for(int nx, int ny, int nz, int &tx, int &ly) {
    {
        std::lock_guard<std::mutex> lk(g_tiling_mu);
        auto it = g_loop_tiling.find({nx, ny, nz});
        if (it != g_loop_tiling.end()) {
            tx = it->second.tx;
            ly = it->second.ly;
            return;
        }
        auto jt = g_tiling.find({nx, ny, nz});
        if (jt != g_tiling.end()) {
            tx = jt->second.first;
            ly = jt->second.second;
            return;
        }
    }
    tx = default_loop_tx(nx, ny, nz);
    ly = 0;
}

// The tiling of tf_evolve_state with 62-cell z-tiles: the single-step
// setting of the shape (tf_autotune_evolve, tf_set_tiling), else defaults.
void tiling_for(int nx, int ny, int nz, int &tx, int &ly) {
    {
        std::lock_guard<std::mutex> lk(g_tiling_mu);
        auto it = g_tiling.find({nx, ny, nz});
        if (it != g_tiling.end()) {
            tx = it->second.first;
            ly = it->second.second;
            return;
        }
    }
    tx = default_tile_x(nx);
    ly = default_tile_y(nx, ny, nz, tx);
}

// ... and with whole z-lines: the single-step setting, else the loop tuner's
// measurement of the shape (the same kernels), else defaults.
void line_tiling_for(int nx, int ny, int nz, int &tx, int &ly) {
    {
        std::lock_guard<std::mutex> lk(g_tiling_mu);
        auto it = g_tiling.find({nx, ny, nz});
        if (it != g_tiling.end()) {
            tx = it->second.first;
            ly = it->second.second;
            return;
        }
    }
    loop_tiling_for(nx, ny, nz, tx, ly);
    if (ly <= 0) ly = default_tile_y(nx, ny, nz, tx);
}

std::atomic<int> g_line_mode{1};

// Whole z-lines are possible where they fit a workgroup and use the lanes at
// least as well as 62-cell z-tiles (every row segment then starts at its
// line's start instead of one cell before a 62-cell tile).
bool line_allowed(int nz) {
    return g_line_mode.load() != 0 && nz <= 64 * kMaxLineW && line_lane_use(nz) >= halo_lane_use(nz) - 1e-9;
}
// ... and used unless the loop tuner measured 62-cell tiles faster for this
// shape (few waves per line leave too few waves per CU; 384^3 is one such
// shape).
bool use_line_tiling(int nx, int ny, int nz) {
    if (!line_allowed(nz)) return false;
    if (tl_line_force >= 0) return tl_line_force == 1;
    std::lock_guard<std::mutex> lk(g_tiling_mu);
    auto it = g_loop_tiling.find({nx, ny, nz});
    return it == g_loop_tiling.end() || it->second.line != 0;
}
}  // namespace

extern "C" {

int tf_abi_version(void) { return TF_ABI_VERSION; }

const char *tf_last_error(void) { return g_err.c_str(); }

// Workspace of the single-grid entry points: the header, the float64 totals
// partials, then the one-simulation plan of tf_time_integration.
constexpr size_t kLoopPlanOffset = kWsHeaderBytes + (size_t)kTotalsBlocks * 5 * sizeof(double);
static_assert(kLoopPlanOffset % 256 == 0, "plan workspace alignment");

size_t tf_workspace_bytes(int nx, int ny, int nz) {
    (void)nx;
    (void)ny;
    (void)nz;
    return kLoopPlanOffset + tf_batch_workspace_bytes(1);
}

int tf_conserved_state_from_primitive(const float *prim, float *cons, int64_t ncells, double gamma, void *stream) {
    TF_REQUIRE(prim && cons, "null state pointer");
    TF_REQUIRE(ncells >= 0, "ncells must be >= 0");
    if (ncells == 0) return TF_OK;
    TF_HIP(launch_p2c(prim, cons, ncells, make_gas(gamma), S_(stream)));
    return TF_OK;
}

int tf_primitive_state_from_conserved(const float *cons, float *prim, int64_t ncells, double gamma, void *stream) {
    TF_REQUIRE(prim && cons, "null state pointer");
    TF_REQUIRE(ncells >= 0, "ncells must be >= 0");
    if (ncells == 0) return TF_OK;
    TF_HIP(launch_c2p(cons, prim, ncells, make_gas(gamma), S_(stream)));
    return TF_OK;
}

int tf_cfl_time_step(const float *prim, int nx, int ny, int nz, double grid_spacing, double gamma, double c_cfl,
                     double *dt_out, float *max_speed_out, void *workspace, void *stream) {
    if (int e = req_dims(nx, ny, nz)) return e;
    TF_REQUIRE(prim && workspace, "null pointer");
    TF_REQUIRE(dt_out || max_speed_out, "nothing to write: dt_out and max_speed_out are both NULL");
    auto *H = reinterpret_cast<WsHeader *>(workspace);
    hipStream_t st = S_(stream);
    TF_HIP(launch_fill_u32(H->bits, 0u, 1, st));
    TF_HIP(launch_speed(prim, (int64_t)nx * ny * nz, (float)gamma, H->bits, st));
    TF_HIP(launch_dt_from_bits(H->bits, grid_spacing, c_cfl, dt_out, max_speed_out, st));
    return TF_OK;
}

int tf_open_boundaries(const float *in, float *out, int nx, int ny, int nz, int num_ghost_cells, void *stream) {
    if (int e = req_dims(nx, ny, nz)) return e;
    TF_REQUIRE(in && out, "null state pointer");
    TF_REQUIRE(in != out, "in and out must not alias");
    TF_REQUIRE(num_ghost_cells >= 0, "num_ghost_cells must be >= 0 (got %d)", num_ghost_cells);
    TF_REQUIRE(nx > num_ghost_cells && ny > num_ghost_cells && nz > num_ghost_cells,
               "open boundaries with %d ghost cells need every dimension >= %d", num_ghost_cells, num_ghost_cells + 1);
    TF_HIP(launch_open_bc(in, out, nx, ny, nz, num_ghost_cells, S_(stream)));
    return TF_OK;
}

int tf_evolve_state_along_axis(const float *in, float *out, int nx, int ny, int nz, double grid_spacing,
                               const double *dt, double gamma, int axis, void *stream) {
    if (int e = req_dims(nx, ny, nz)) return e;
    TF_REQUIRE(in && out && dt, "null pointer");
    TF_REQUIRE(in != out, "in and out must not alias");
    TF_REQUIRE(axis >= 1 && axis <= 3, "axis must be 1 (x), 2 (y) or 3 (z), got %d", axis);
    TF_HIP(launch_sweep(axis, in, out, nx, ny, nz, make_gas(gamma), dt, grid_spacing, S_(stream)));
    return TF_OK;
}

}  // extern "C"

namespace {
// tf_evolve_state with whole z-lines: the table and the control block (step
// 0, k from dt) are written into the workspace header on the stream by a
// setup kernel, so the call stays asynchronous.
int evolve_line(const float *in, float *out, int nx, int ny, int nz, double grid_spacing, const double *dt,
                double gamma, int handle_boundaries, unsigned int *max_speed_out, WsHeader *H, int tx, int ly,
                hipStream_t st) {
    SimDev S = make_simdev(nx, ny, nz, handle_boundaries, gamma, tx, ly);
    make_line(S, tx);
    S.buf[0] = const_cast<float *>(in);
    S.buf[1] = out;
    S.ctrl = &H->ctrl;
    S.skip_max = max_speed_out ? 0 : 1;
    const int nblocks = S.ntx * S.nyc;
    TF_HIP(launch_direct_setup(S, &H->sim, H->prefix, nblocks, dt, grid_spacing, st));
    TF_HIP(launch_line_sims(line_tx(S.line_w, tx), S.line_w, &H->sim, H->prefix, 1, nblocks, st));
    if (max_speed_out) TF_HIP(launch_direct_bits(&H->ctrl, max_speed_out, st));
    return TF_OK;
}

// tf_evolve_state with 62-cell z-tiles (no workspace needed).
int evolve_halo(const float *in, float *out, int nx, int ny, int nz, double grid_spacing, const double *dt,
                double gamma, int handle_boundaries, unsigned int *max_speed_out, int tx, int ly, hipStream_t st) {
    const SimDev S = make_simdev(nx, ny, nz, handle_boundaries, gamma, tx, ly);
    FarList F;
    std::string e;
    if (int rc = far_list(S, tx, F, e)) return fail(rc, "%s", e.c_str());
    TF_HIP(launch_step_direct(tx, S, in, out, dt, grid_spacing, max_speed_out, st));
    TF_HIP(launch_step_direct_far(tx, S, F, in, out, dt, grid_spacing, max_speed_out, st));
    return TF_OK;
}
}  // namespace

extern "C" {

int tf_evolve_state(const float *in, float *out, int nx, int ny, int nz, double grid_spacing, const double *dt,
                    double gamma, int handle_boundaries, unsigned int *max_speed_out, void *workspace,
                    void *stream) {
    if (int e = req_dims(nx, ny, nz)) return e;
    if (int e = req_fused_size(nx, ny, nz)) return e;
    TF_REQUIRE(in && out && dt, "null pointer");
    TF_REQUIRE(in != out, "in and out must not alias");
    if (int e = req_open_bc(handle_boundaries, nx, ny, nz)) return e;
    int tx, ly;
    if (workspace && use_line_tiling(nx, ny, nz)) {
        line_tiling_for(nx, ny, nz, tx, ly);
        return evolve_line(in, out, nx, ny, nz, grid_spacing, dt, gamma, handle_boundaries, max_speed_out,
                           reinterpret_cast<WsHeader *>(workspace), tx, ly, S_(stream));
    }
    tiling_for(nx, ny, nz, tx, ly);
    return evolve_halo(in, out, nx, ny, nz, grid_spacing, dt, gamma, handle_boundaries, max_speed_out, tx, ly,
                       S_(stream));
}


int tf_conserved_totals(const float *prim, int nx, int ny, int nz, double gamma, double *out5, void *workspace,
                        void *stream) {
    if (int e = req_dims(nx, ny, nz)) return e;
    TF_REQUIRE(prim && out5 && workspace, "null pointer");
    double *partials = reinterpret_cast<double *>(static_cast<char *>(workspace) + kWsHeaderBytes);
    TF_HIP(launch_totals(prim, (int64_t)nx * ny * nz, gamma, partials, out5, S_(stream)));
    return TF_OK;
}

// ------------------------------------------------------------------ per-interface helpers
int tf_speed_of_sound(const float *rho, const float *p, float *c, int64_t n, double gamma, void *stream) {
    TF_REQUIRE(rho && p && c, "null pointer");
    TF_REQUIRE(n >= 0, "n must be >= 0");
    if (n == 0) return TF_OK;
    TF_HIP(launch_sound_speed(rho, p, c, n, make_gas(gamma), S_(stream)));
    return TF_OK;
}

int tf_get_wave_speeds(const float *primitives_left, const float *primitives_right, int64_t ncells, double gamma,
                       int flux_direction_index, float *wave_speeds_left_minus, float *wave_speeds_right_plus,
                       void *stream) {
    TF_REQUIRE(primitives_left && primitives_right && wave_speeds_left_minus && wave_speeds_right_plus,
               "null pointer");
    TF_REQUIRE(ncells >= 0, "ncells must be >= 0");
    TF_REQUIRE(flux_direction_index >= 1 && flux_direction_index <= 3,
               "flux_direction_index must be 1, 2 or 3 (got %d)", flux_direction_index);
    if (ncells == 0) return TF_OK;
    TF_HIP(launch_wave_speeds(primitives_left, primitives_right, wave_speeds_left_minus, wave_speeds_right_plus,
                              ncells, make_gas(gamma), flux_direction_index, nullptr, S_(stream)));
    return TF_OK;
}

int tf_get_max_wave_speeds(const float *primitives_left, const float *primitives_right, int64_t ncells, double gamma,
                           int flux_direction_index, float *max_out, void *workspace, void *stream) {
    TF_REQUIRE(primitives_left && primitives_right && max_out && workspace, "null pointer");
    TF_REQUIRE(ncells >= 1, "ncells must be >= 1 (the maximum of nothing is undefined)");
    TF_REQUIRE(flux_direction_index >= 1 && flux_direction_index <= 3,
               "flux_direction_index must be 1, 2 or 3 (got %d)", flux_direction_index);
    auto *H = reinterpret_cast<WsHeader *>(workspace);
    hipStream_t st = S_(stream);
    TF_HIP(launch_fill_u32(&H->bits[1], 0u, 1, st));
    TF_HIP(launch_wave_speeds(primitives_left, primitives_right, nullptr, nullptr, ncells, make_gas(gamma),
                              flux_direction_index, &H->bits[1], st));
    TF_HIP(launch_dt_from_bits(&H->bits[1], 1.0, 0.0, nullptr, max_out, st));
    return TF_OK;
}

int tf_euler_flux(const float *primitive_state, float *flux, int64_t ncells, double gamma, int flux_direction_index,
                  void *stream) {
    TF_REQUIRE(primitive_state && flux, "null pointer");
    TF_REQUIRE(primitive_state != flux, "primitive_state and flux must not alias");
    TF_REQUIRE(ncells >= 0, "ncells must be >= 0");
    TF_REQUIRE(flux_direction_index >= 1 && flux_direction_index <= 3,
               "flux_direction_index must be 1, 2 or 3 (got %d)", flux_direction_index);
    if (ncells == 0) return TF_OK;
    TF_HIP(launch_euler_flux(primitive_state, flux, ncells, make_gas(gamma), flux_direction_index, S_(stream)));
    return TF_OK;
}

int tf_hll_solver(const float *primitives_left, const float *primitives_right, float *flux, int64_t ncells,
                  double gamma, int flux_direction_index, void *stream) {
    TF_REQUIRE(primitives_left && primitives_right && flux, "null pointer");
    TF_REQUIRE(flux != primitives_left && flux != primitives_right, "flux must not alias the states");
    TF_REQUIRE(ncells >= 0, "ncells must be >= 0");
    TF_REQUIRE(flux_direction_index >= 1 && flux_direction_index <= 3,
               "flux_direction_index must be 1, 2 or 3 (got %d)", flux_direction_index);
    if (ncells == 0) return TF_OK;
    TF_HIP(launch_hll_flux(primitives_left, primitives_right, flux, ncells, make_gas(gamma), flux_direction_index,
                           S_(stream)));
    return TF_OK;
}

int tf_set_along_axis(const float *in, float *out, int nx, int ny, int nz, int axis, int set_index, int get_index,
                      void *stream) {
    if (int e = req_dims(nx, ny, nz)) return e;
    TF_REQUIRE(in && out, "null pointer");
    TF_REQUIRE(in != out, "in and out must not alias");
    TF_REQUIRE(axis >= 0 && axis <= 3, "axis must be 0 (variables), 1, 2 or 3 (got %d)", axis);
    const int n = axis == 0 ? 5 : (axis == 1 ? nx : (axis == 2 ? ny : nz));
    const int s = set_index < 0 ? set_index + n : set_index, g = get_index < 0 ? get_index + n : get_index;
    TF_REQUIRE(s >= 0 && s < n, "set_index %d out of range for axis %d of size %d", set_index, axis, n);
    TF_REQUIRE(g >= 0 && g < n, "get_index %d out of range for axis %d of size %d", get_index, axis, n);
    TF_HIP(launch_set_along_axis(in, out, nx, ny, nz, axis, s, g, S_(stream)));
    return TF_OK;
}

// ------------------------------------------------------------------ HBM copy
int tf_hbm_copy(const void *src, void *dst, size_t bytes, void *stream) {
    TF_REQUIRE(src && dst, "null pointer");
    TF_REQUIRE(bytes % 16 == 0 && reinterpret_cast<uintptr_t>(src) % 16 == 0 &&
                   reinterpret_cast<uintptr_t>(dst) % 16 == 0,
               "tf_hbm_copy needs 16-byte aligned pointers and a size in multiples of 16 bytes");
    const uintptr_t s0 = reinterpret_cast<uintptr_t>(src), d0 = reinterpret_cast<uintptr_t>(dst);
    TF_REQUIRE(s0 + bytes <= d0 || d0 + bytes <= s0, "tf_hbm_copy buffers overlap");
    if (bytes == 0) return TF_OK;
    TF_HIP(launch_copy16(src, dst, (int64_t)(bytes / 16), S_(stream)));
    return TF_OK;
}

// ------------------------------------------------------------------ tiling
int tf_get_tiling(int nx, int ny, int nz, int *tile_x, int *tile_y) {
    if (int e = req_dims(nx, ny, nz)) return e;
    TF_REQUIRE(tile_x && tile_y, "null pointer");
    tiling_for(nx, ny, nz, *tile_x, *tile_y);
    return TF_OK;
}

int tf_set_tiling(int nx, int ny, int nz, int tile_x, int tile_y) {
    if (int e = req_dims(nx, ny, nz)) return e;
    TF_REQUIRE(valid_tx(tile_x), "tile_x must be 2, 4, 6 or 8 (got %d)", tile_x);
    TF_REQUIRE(tile_y >= 1, "tile_y must be >= 1");
    std::lock_guard<std::mutex> lk(g_tiling_mu);
    g_tiling[{nx, ny, nz}] = {tile_x, std::min(tile_y, ny)};
    g_loop_tiling[{nx, ny, nz}] = {tile_x, std::min(tile_y, ny), 1};
    return TF_OK;
}

int tf_autotune_evolve(const float *in, float *out, int nx, int ny, int nz, double grid_spacing, double gamma,
                       void *workspace, int *tile_x, int *tile_y, float *ms_per_step, void *stream) {
    if (int e = req_dims(nx, ny, nz)) return e;
    if (int e = req_fused_size(nx, ny, nz)) return e;
    TF_REQUIRE(in && out && workspace && tile_x && tile_y, "null pointer");
    TF_REQUIRE(in != out, "in and out must not alias");
    hipStream_t st = S_(stream);
    auto *H = reinterpret_cast<WsHeader *>(workspace);
    double *dt = &H->totals[0];
    // a realistic dt: the CFL step of the input
    TF_HIP(launch_fill_u32(H->bits, 0u, 1, st));
    TF_HIP(launch_speed(in, (int64_t)nx * ny * nz, (float)gamma, H->bits, st));
    TF_HIP(launch_dt_from_bits(H->bits, grid_spacing, 0.4, dt, nullptr, st));
    std::vector<int> lys;
    for (int v : {8, 16, 32, 64, 128, 256, 512}) if (v < ny) lys.push_back(v);
    lys.push_back(ny);
    // The candidates of the kernels tf_evolve_state runs for this shape with
    // a workspace: whole z-lines (tile width 2 or 4) or 62-cell z-tiles.
    const bool line = use_line_tiling(nx, ny, nz);
    const std::vector<int> txs = line ? std::vector<int>{2, 4} : std::vector<int>{2, 4, 6, 8};
    EventPair ev;
    TF_HIP(ev.create());
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    float best = 1e30f;
    int btx = 4, bly = ny;
    int rc = TF_OK;
    std::vector<std::pair<int, int>> seen;  // resolved (tile width, rows)
    for (int tx : txs) {
        if (tx > 2 && tx > nx) continue;
        for (int ly : lys) {
            SimDev S = make_simdev(nx, ny, nz, 0, gamma, tx, ly);
            if (line) make_line(S, tx);
            const std::pair<int, int> key{line ? line_tx(S.line_w, tx) : tx, S.ly};
            if (std::find(seen.begin(), seen.end(), key) != seen.end()) continue;
            seen.push_back(key);
            auto step = [&]() {
                return line ? evolve_line(in, out, nx, ny, nz, grid_spacing, dt, gamma, 0, nullptr, H, tx, ly, st)
                            : evolve_halo(in, out, nx, ny, nz, grid_spacing, dt, gamma, 0, nullptr, tx, ly, st);
            };
            const int reps = 3;
            rc = step();
            if (rc == TF_OK && hipEventRecord(e0, st) != hipSuccess) rc = fail(TF_ERR_HIP, "autotune event");
            for (int r = 0; r < reps && rc == TF_OK; ++r) rc = step();
            if (rc != TF_OK) break;
            hipError_t e = hipEventRecord(e1, st);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e != hipSuccess) {
                rc = fail(TF_ERR_HIP, "autotune launch: %s", hipGetErrorString(e));
                break;
            }
            ms /= reps;
            if (ms < best) {
                best = ms;
                btx = tx;
                bly = S.ly;
            }
        }
        if (rc != TF_OK) break;
    }
    if (rc != TF_OK) return rc;
    {
        std::lock_guard<std::mutex> lk(g_tiling_mu);
        g_tiling[{nx, ny, nz}] = {btx, bly};
    }
    *tile_x = btx;
    *tile_y = bly;
    if (ms_per_step) *ms_per_step = best;
    return TF_OK;
}

// ------------------------------------------------------------------ batched engine
// A class of simulations launched together: halo tiling (w == 0, prefix over
// waves) or line tiling with w waves per workgroup (prefix over workgroups).

struct SimClass {
    int w = 0;
    int tx = 0;  // tile width of the class's kernel
    int begin = 0, count = 0, total = 0;
    int *d_prefix = nullptr;
};

struct tf_batch {
    int nsims = 0;
    int tx = 4;
    int total_waves = 0;
    SimDev *d_sims = nullptr;
    int *d_prefix = nullptr;
    Ctrl *d_status = nullptr;
    SimClass cls[kClasses];
    std::vector<SimDev> h_sims;  // device table order
    std::vector<int> h_prefix;
    std::vector<int> order;      // h_sims[k] is the caller's simulation order[k]
    FarTile *d_far = nullptr;    // far-row tiles (table indices), see far_tiles
    int nfar = 0;
    ResetParams *d_reset = nullptr;  // loop bounds per simulation (table order), for tf_batch_reset
    int *d_speed_prefix = nullptr;   // signal-speed workgroups per simulation (prefix, table order)
    int speed_blocks = 0;
    int far_max_ntz = 0;
    // Captured graphs of kGraphChunk steps (slot 0) and of a shorter
    // remainder (slot 1), replayed by tf_batch_step_graph; slot 2 holds a
    // whole run (tf_batch_run_graph).
    static constexpr int kGraphChunk = 16;
    static constexpr int kGraphSlots = 3;
    static constexpr int kWholeRunMax = 1024;
    hipGraph_t graph[kGraphSlots] = {};
    hipGraphExec_t exec[kGraphSlots] = {};
    int graph_steps[kGraphSlots] = {};
    hipStream_t capture_stream = nullptr;
    // Concurrent launches of the tiling classes of one step: side streams
    // (default priority) forked from and joined back into the caller's stream
    // with events; created on first use when a plan has several classes.
    static constexpr int kSide = 3;
    hipStream_t side[kSide] = {};
    hipEvent_t fork_ev = nullptr;
    hipEvent_t join_ev[kSide] = {};
    int nlaunch = 0;  // non-empty stencil launches per step (classes + far rows)
};

size_t tf_batch_workspace_bytes(int nsims) { return plan_workspace_bytes(nsims); }

int tf_set_line_mode(int enabled) {
    g_line_mode.store(enabled ? 1 : 0);
    return TF_OK;
}

int tf_sim_reset(const tf_sim_desc *sim, void *stream) {
    TF_REQUIRE(sim, "null sim");
    if (int e = req_dims(sim->nx, sim->ny, sim->nz)) return e;
    TF_REQUIRE(sim->buf0 && sim->buf1 && sim->ctrl, "null device pointer in sim");
    if (int e = req_open_bc(sim->handle_boundaries, sim->nx, sim->ny, sim->nz)) return e;
    hipStream_t st = S_(stream);
    Ctrl *C = reinterpret_cast<Ctrl *>(sim->ctrl);
    TF_HIP(launch_ctrl_reset(C, sim->grid_spacing, sim->c_cfl, sim->t_final, sim->max_steps, st));
    TF_HIP(launch_speed(sim->buf0, (int64_t)sim->nx * sim->ny * sim->nz, (float)sim->gamma, &C->max_bits[0], st));
    TF_HIP(launch_finalize_one(C, 1, st));
    return TF_OK;
}

}  // extern "C"

namespace {
// Page-locked staging blocks for plan-table uploads that do not wait
// (tf_batch_create_async). A block is reused once the event recorded after
// its last upload has completed; blocks are kept until the process ends, so
// no page-locked free (which may synchronise the device) is ever needed.
struct Staging {
    unsigned char *host = nullptr;
    size_t bytes = 0;
    int device = -1;
    bool in_use = false;
    hipEvent_t done = nullptr;
};
std::mutex g_staging_mu;
std::vector<Staging> g_staging;

// Index of a free block of at least `bytes` on `device`, allocated if none is.
int staging_acquire(size_t bytes, int device, hipError_t &err) {
    std::lock_guard<std::mutex> lk(g_staging_mu);
    for (size_t i = 0; i < g_staging.size(); ++i) {
        Staging &s = g_staging[i];
        if (s.in_use || s.device != device || s.bytes < bytes) continue;
        if (s.done && hipEventQuery(s.done) != hipSuccess) continue;
        s.in_use = true;
        return (int)i;
    }
    Staging s;
    s.bytes = std::max<size_t>(bytes, 64 << 10);
    s.device = device;
    err = hipHostMalloc((void **)&s.host, s.bytes, hipHostMallocDefault);
    if (err != hipSuccess) return -1;
    err = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
    if (err != hipSuccess) {
        (void)hipHostFree(s.host);  // error path only
        return -1;
    }
    s.in_use = true;
    g_staging.push_back(s);
    return (int)g_staging.size() - 1;
}

// Records the block's event on `st` after its uploads and releases it. If the
// record fails, the block stays in use for good: an event that was never
// recorded would report its uploads complete.
hipError_t staging_release(int i, hipStream_t st) {
    std::lock_guard<std::mutex> lk(g_staging_mu);
    Staging &s = g_staging[i];
    const hipError_t e = hipEventRecord(s.done, st);
    if (e == hipSuccess) s.in_use = false;
    return e;
}

unsigned char *staging_host(int i) {
    std::lock_guard<std::mutex> lk(g_staging_mu);
    return g_staging[i].host;
}

int batch_create(tf_batch **out, const tf_sim_desc *sims, int nsims, void *workspace, size_t workspace_bytes,
                 int tile_x, int tile_y, void *stream, bool wait);
}  // namespace

extern "C" {

int tf_batch_create(tf_batch **out, const tf_sim_desc *sims, int nsims, void *workspace, size_t workspace_bytes,
                    int tile_x, int tile_y, void *stream) {
    return batch_create(out, sims, nsims, workspace, workspace_bytes, tile_x, tile_y, stream, true);
}

int tf_batch_create_async(tf_batch **out, const tf_sim_desc *sims, int nsims, void *workspace,
                          size_t workspace_bytes, int tile_x, int tile_y, void *stream) {
    return batch_create(out, sims, nsims, workspace, workspace_bytes, tile_x, tile_y, stream, false);
}

}  // extern "C"

namespace {
int batch_create(tf_batch **out, const tf_sim_desc *sims, int nsims, void *workspace, size_t workspace_bytes,
                 int tile_x, int tile_y, void *stream, bool wait) {
    TF_REQUIRE(out, "null output handle");
    *out = nullptr;
    TF_REQUIRE(nsims >= 1, "nsims must be >= 1");
    TF_REQUIRE(sims && workspace, "null pointer");
    for (int i = 0; i < nsims; ++i)
        if (int e = req_dims(sims[i].nx, sims[i].ny, sims[i].nz)) return e;
    // The registry's part of the plan: whole z-lines or not per shape, and,
    // when neither tile_x nor tile_y is given, the measured loop tilings
    // (tf_autotune_loop / tf_set_tiling): the tile width of the largest
    // simulation and each shape's own y-chunks.
    std::vector<unsigned char> line(nsims);
    std::vector<int> tuned_ly(nsims, 0);
    int tuned_tx = 0;
    {
        int largest = 0;
        double max_cells = 0.0;
        for (int i = 0; i < nsims; ++i) {
            line[i] = use_line_tiling(sims[i].nx, sims[i].ny, sims[i].nz) ? 1 : 0;
            const double c = (double)sims[i].nx * sims[i].ny * sims[i].nz;
            if (c > max_cells) {
                max_cells = c;
                largest = i;
            }
        }
        if (tile_x == 0 && tile_y == 0) {
            std::lock_guard<std::mutex> lk(g_tiling_mu);
            for (int i = 0; i < nsims; ++i) {
                auto it = g_loop_tiling.find({sims[i].nx, sims[i].ny, sims[i].nz});
                if (it == g_loop_tiling.end()) continue;
                tuned_ly[i] = it->second.ly;
                if (i == largest) tuned_tx = it->second.tx;
            }
        }
    }
    PlanRequest rq;
    rq.sims = sims;
    rq.nsims = nsims;
    rq.workspace = reinterpret_cast<uintptr_t>(workspace);
    rq.workspace_bytes = workspace_bytes;
    rq.tile_x = tile_x;
    rq.tile_y = tile_y;
    rq.tuned_ly = tuned_ly.data();
    rq.tuned_tx = tuned_tx;
    rq.line = line.data();
    PlanTables P;
    std::string err;
    if (int rc = build_plan(rq, P, err)) return fail(rc, "%s", err.c_str());
    auto *b = new tf_batch();
    b->nsims = nsims;
    b->tx = P.tx;
    b->total_waves = (int)P.total_waves;
    char *const base = static_cast<char *>(workspace);
    b->d_sims = reinterpret_cast<SimDev *>(base + P.off_sims);
    b->d_prefix = reinterpret_cast<int *>(base + P.off_prefix);
    b->d_status = reinterpret_cast<Ctrl *>(base + P.off_status);
    b->d_reset = reinterpret_cast<ResetParams *>(base + P.off_reset);
    b->d_speed_prefix = reinterpret_cast<int *>(base + P.off_speed);
    b->d_far = reinterpret_cast<FarTile *>(base + P.off_far);
    for (int c = 0; c < kClasses; ++c) {
        const PlanClass &pc = P.cls[c];
        SimClass &C = b->cls[c];
        C.w = pc.w;
        C.tx = pc.tx;
        C.begin = pc.begin;
        C.count = pc.count;
        C.total = pc.total;
        C.d_prefix = b->d_prefix + pc.prefix_at;
    }
    b->h_sims = P.sims;
    b->h_prefix = P.prefix;
    b->order = P.order;
    b->nfar = (int)P.far.size();
    b->far_max_ntz = P.far_max_ntz;
    b->speed_blocks = P.speed_blocks;
    b->nlaunch = P.nlaunch;
    // The tables in one upload of their host image, the far tiles (when there
    // are any) in a second; the signal-speed slots are cleared on the device.
    // Waiting: from the host vectors, then the stream is synchronised. Not
    // waiting: from a page-locked staging block, guarded by an event.
    hipStream_t st = S_(stream);
    const size_t far_bytes = P.far.size() * sizeof(FarTile);
    const void *image = P.image.data();
    const void *far = P.far.data();
    int slot = -1;
    hipError_t e = hipSuccess;
    if (!wait) {
        int dev = 0;
        e = hipGetDevice(&dev);
        if (e == hipSuccess) slot = staging_acquire(P.image.size() + far_bytes, dev, e);
        if (slot >= 0) {
            unsigned char *h = staging_host(slot);
            std::memcpy(h, image, P.image.size());
            if (far_bytes) std::memcpy(h + P.image.size(), far, far_bytes);
            image = h;
            far = h + P.image.size();
        }
    }
    if (e == hipSuccess) e = hipMemsetAsync(base + P.off_maxp, 0, (size_t)nsims * kSlotBytesPerSim, st);  // not in a graph
    if (e == hipSuccess) e = hipMemcpyAsync(base, image, P.image.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && b->nfar > 0) e = hipMemcpyAsync(b->d_far, far, far_bytes, hipMemcpyHostToDevice, st);
    if (slot >= 0) {
        const hipError_t r = staging_release(slot, st);
        if (e == hipSuccess) e = r;
    }
    if (e == hipSuccess && wait) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        delete b;
        return fail(TF_ERR_HIP, "batch table upload: %s", hipGetErrorString(e));
    }
    *out = b;
    return TF_OK;
}
}  // namespace

extern "C" {

int tf_batch_sim_tiling(const tf_batch *b, int sim, int *tile_x, int *tile_y, int *line_waves) {
    TF_REQUIRE(b && tile_x && tile_y && line_waves, "null pointer");
    TF_REQUIRE(sim >= 0 && sim < b->nsims, "simulation index %d out of range", sim);
    for (int k = 0; k < b->nsims; ++k) {
        if (b->order[k] != sim) continue;
        const SimDev &S = b->h_sims[k];
        *line_waves = S.line_w;
        *tile_x = b->cls[S.line_w].tx;
        *tile_y = S.ly;
        return TF_OK;
    }
    return fail(TF_ERR_ARG, "simulation %d not in the plan", sim);
}

int tf_batch_destroy(tf_batch *b) {
    if (!b) return TF_OK;
    for (int i = 0; i < tf_batch::kGraphSlots; ++i) {
        if (b->exec[i]) (void)hipGraphExecDestroy(b->exec[i]);
        if (b->graph[i]) (void)hipGraphDestroy(b->graph[i]);
    }
    if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
    for (int i = 0; i < tf_batch::kSide; ++i) {
        if (b->side[i]) (void)hipStreamDestroy(b->side[i]);
        if (b->join_ev[i]) (void)hipEventDestroy(b->join_ev[i]);
    }
    if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
    delete b;
    return TF_OK;
}

// The k-th non-empty stencil launch of a step (the halo-tiled class, each
// line-tiled class in order, then the far rows) on stream st.
static hipError_t launch_stencil_part(tf_batch *b, int k, hipStream_t st) {
    int i = 0;
    for (int c = 0; c < kClasses; ++c) {
        const SimClass &C = b->cls[c];
        if (C.count <= 0) continue;
        if (i++ != k) continue;
        if (c == 0) return launch_step_sims(C.tx, b->d_sims + C.begin, C.d_prefix, C.count, C.total, st);
        return launch_line_sims(C.tx, c, b->d_sims + C.begin, C.d_prefix, C.count, C.total, st);
    }
    if (b->nfar > 0 && i == k) return launch_step_sims_far(b->tx, b->d_sims, b->d_far, b->nfar, b->far_max_ntz, st);
    return hipSuccess;
}

// One batched stencil step: every tiling class and the far-row tiles, in
// order on st, or, with `concurrent` and more than one launch, launch 0 on st
// and the others on side streams forked from st and joined back (each
// launch's tail then overlaps the others instead of draining the device
// between them). Captured graphs use the concurrent form, where the fork and
// join cost nothing at replay; eager steps use the ordered form, whose host
// cost per step is lower.
static hipError_t launch_stencils(tf_batch *b, hipStream_t st, bool concurrent) {
    if (!concurrent || b->nlaunch <= 1) {
        hipError_t e = hipSuccess;
        for (int k = 0; k < std::max(1, b->nlaunch) && e == hipSuccess; ++k) e = launch_stencil_part(b, k, st);
        return e;
    }
    hipError_t e = hipSuccess;
    if (!b->fork_ev) {
        e = hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming);
        for (int i = 0; i < tf_batch::kSide && e == hipSuccess; ++i) {
            e = hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join_ev[i], hipEventDisableTiming);
        }
        if (e != hipSuccess) return e;
    }
    const int nside = std::min(b->nlaunch - 1, (int)tf_batch::kSide);
    e = hipEventRecord(b->fork_ev, st);
    for (int i = 0; i < nside && e == hipSuccess; ++i) e = hipStreamWaitEvent(b->side[i], b->fork_ev, 0);
    // launch k > 0 goes to side stream (k - 1) % nside; launch 0 to st
    for (int k = 1; k < b->nlaunch && e == hipSuccess; ++k) e = launch_stencil_part(b, k, b->side[(k - 1) % nside]);
    if (e == hipSuccess) e = launch_stencil_part(b, 0, st);
    for (int i = 0; i < nside && e == hipSuccess; ++i) {
        e = hipEventRecord(b->join_ev[i], b->side[i]);
        if (e == hipSuccess) e = hipStreamWaitEvent(st, b->join_ev[i], 0);
    }
    return e;
}

static int enqueue_steps(tf_batch *b, int steps, hipStream_t st, bool concurrent = false) {
    for (int i = 0; i < steps; ++i) {
        TF_HIP(launch_stencils(b, st, concurrent));
        TF_HIP(launch_finalize_sims(b->d_sims, b->nsims, st));
    }
    return TF_OK;
}

int tf_batch_step(tf_batch *b, int steps, void *stream) {
    TF_REQUIRE(b, "null batch");
    TF_REQUIRE(steps >= 0, "steps must be >= 0");
    return enqueue_steps(b, steps, S_(stream));
}

// Graph slot `slot` holding `steps` captured steps (captured on first use
// and when the step count changes).
static int ensure_graph(tf_batch *b, int slot, int steps) {
    if (b->exec[slot] && b->graph_steps[slot] == steps) return TF_OK;
    if (b->exec[slot]) {
        (void)hipGraphExecDestroy(b->exec[slot]);
        b->exec[slot] = nullptr;
    }
    if (b->graph[slot]) {
        (void)hipGraphDestroy(b->graph[slot]);
        b->graph[slot] = nullptr;
    }
    if (!b->capture_stream) TF_HIP(hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking));
    TF_HIP(hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_steps(b, steps, b->capture_stream, true);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->capture_stream, &g);
    if (rc != TF_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(TF_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
    b->graph[slot] = g;
    TF_HIP(hipGraphInstantiate(&b->exec[slot], g, nullptr, nullptr, 0));
    b->graph_steps[slot] = steps;
    return TF_OK;
}

// Steps run as replays of a graph of kGraphChunk steps plus one of the
// remainder: instantiation cost stays that of at most 2 x kGraphChunk steps
// whatever the step count, and one replay of 16 steps costs a single launch.
int tf_batch_step_graph(tf_batch *b, int steps, void *stream) {
    TF_REQUIRE(b, "null batch");
    TF_REQUIRE(steps >= 1, "steps must be >= 1");
    const int chunk = std::min(steps, (int)tf_batch::kGraphChunk);
    if (int e = ensure_graph(b, 0, chunk)) return e;
    for (int i = 0; i < steps / chunk; ++i) TF_HIP(hipGraphLaunch(b->exec[0], S_(stream)));
    if (const int rem = steps % chunk) {
        if (int e = ensure_graph(b, 1, rem)) return e;
        TF_HIP(hipGraphLaunch(b->exec[1], S_(stream)));
    }
    return TF_OK;
}

// A whole run as one graph: every step of it captured once per step count
// and launched as a single graph; simulations that reach their end return at
// once in the remaining steps.
int tf_batch_run_graph(tf_batch *b, int steps, void *stream) {
    TF_REQUIRE(b, "null batch");
    TF_REQUIRE(steps >= 1 && steps <= tf_batch::kWholeRunMax,
               "steps must be 1..%d for one graph (tf_batch_step_graph replays chunks of any length)",
               tf_batch::kWholeRunMax);
    if (int e = ensure_graph(b, 2, steps)) return e;
    TF_HIP(hipGraphLaunch(b->exec[2], S_(stream)));
    return TF_OK;
}

int tf_batch_step_timed(tf_batch *b, int steps, float *stencil_ms, void *stream) {
    TF_REQUIRE(b && stencil_ms, "null pointer");
    TF_REQUIRE(steps >= 1, "steps must be >= 1");
    hipStream_t st = S_(stream);
    std::vector<hipEvent_t> ev(2 * (size_t)steps, nullptr);
    int rc = TF_OK;
    for (auto &e : ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            rc = fail(TF_ERR_HIP, "hipEventCreate failed");
            break;
        }
    }
    for (int i = 0; i < steps && rc == TF_OK; ++i) {
        hipError_t e = hipEventRecord(ev[2 * i], st);
        if (e == hipSuccess) e = launch_stencils(b, st, false);
        if (e == hipSuccess) e = hipEventRecord(ev[2 * i + 1], st);
        if (e == hipSuccess) e = launch_finalize_sims(b->d_sims, b->nsims, st);
        if (e != hipSuccess) rc = fail(TF_ERR_HIP, "timed step: %s", hipGetErrorString(e));
    }
    float total = 0.f;
    if (rc == TF_OK) {
        hipError_t e = hipStreamSynchronize(st);
        for (int i = 0; i < steps && e == hipSuccess; ++i) {
            float ms = 0.f;
            e = hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]);
            total += ms;
        }
        if (e != hipSuccess) rc = fail(TF_ERR_HIP, "timed step events: %s", hipGetErrorString(e));
    }
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    if (rc == TF_OK) *stencil_ms = total;
    return rc;
}

int tf_batch_reset(tf_batch *b, void *stream) {
    TF_REQUIRE(b, "null batch");
    TF_HIP(launch_reset_sims(b->d_sims, b->d_reset, b->d_speed_prefix, b->nsims, b->speed_blocks, S_(stream)));
    return TF_OK;
}

// tf_batch_status copies the first sizeof(tf_sim_status) bytes of each Ctrl.
static_assert(sizeof(tf_sim_status) == 48 && offsetof(Ctrl, t) == offsetof(tf_sim_status, t) &&
                  offsetof(Ctrl, dt) == offsetof(tf_sim_status, dt) &&
                  offsetof(Ctrl, n_iter) == offsetof(tf_sim_status, n_iter) &&
                  offsetof(Ctrl, active) == offsetof(tf_sim_status, active) &&
                  offsetof(Ctrl, nonfinite) == offsetof(tf_sim_status, nonfinite) &&
                  offsetof(Ctrl, max_speed) == offsetof(tf_sim_status, max_speed),
              "Ctrl must begin with the layout of tf_sim_status");

size_t tf_batch_status_bytes(const tf_batch *b) { return b ? (size_t)b->nsims * sizeof(Ctrl) : 0; }

// The control blocks in plan order are gathered into the plan's status
// buffer and copied to the caller's host memory, both on `stream`.
int tf_batch_status_enqueue(tf_batch *b, void *host_raw, void *stream) {
    TF_REQUIRE(b && host_raw, "null pointer");
    hipStream_t st = S_(stream);
    TF_HIP(launch_gather_status(b->d_sims, b->nsims, b->d_status, st));
    TF_HIP(hipMemcpyAsync(host_raw, b->d_status, b->nsims * sizeof(Ctrl), hipMemcpyDeviceToHost, st));
    return TF_OK;
}

int tf_batch_status_unpack(const tf_batch *b, const void *host_raw, tf_sim_status *host_out) {
    TF_REQUIRE(b && host_raw && host_out, "null pointer");
    const Ctrl *c = static_cast<const Ctrl *>(host_raw);
    for (int k = 0; k < b->nsims; ++k) std::memcpy(&host_out[b->order[k]], &c[k], sizeof(tf_sim_status));
    return TF_OK;
}

int tf_batch_status(tf_batch *b, tf_sim_status *host_out, void *stream) {
    TF_REQUIRE(b && host_out, "null pointer");
    std::vector<Ctrl> tmp(b->nsims);
    if (int e = tf_batch_status_enqueue(b, tmp.data(), stream)) return e;
    TF_HIP(hipStreamSynchronize(S_(stream)));
    return tf_batch_status_unpack(b, tmp.data(), host_out);
}

}  // extern "C"

// ------------------------------------------------------------------ single-grid loop
namespace {
struct PlanGuard {
    tf_batch *b = nullptr;
    ~PlanGuard() { tf_batch_destroy(b); }
};

int loop_args(float *state, float *scratch, int nx, int ny, int nz, int handle_boundaries, void *workspace) {
    if (int e = req_dims(nx, ny, nz)) return e;
    if (int e = req_fused_size(nx, ny, nz)) return e;
    TF_REQUIRE(state && scratch && workspace, "null device pointer");
    TF_REQUIRE(state != scratch, "state and scratch must be different buffers");
    if (int e = req_open_bc(handle_boundaries, nx, ny, nz)) return e;
    return TF_OK;
}

tf_sim_desc loop_desc(float *state, float *scratch, int nx, int ny, int nz, double grid_spacing, double t_final,
                      double gamma, int handle_boundaries, int64_t max_steps, double c_cfl, void *workspace) {
    tf_sim_desc d;
    std::memset(&d, 0, sizeof d);
    d.buf0 = state;
    d.buf1 = scratch;
    d.ctrl = &reinterpret_cast<WsHeader *>(workspace)->ctrl;
    d.nx = nx;
    d.ny = ny;
    d.nz = nz;
    d.handle_boundaries = handle_boundaries;
    d.grid_spacing = grid_spacing;
    d.gamma = gamma;
    d.c_cfl = c_cfl;
    d.t_final = t_final;
    d.max_steps = max_steps;
    return d;
}

void *plan_workspace(void *workspace) { return static_cast<char *>(workspace) + kLoopPlanOffset; }
}  // namespace

int tf_time_integration(float *state, float *scratch, int nx, int ny, int nz, double grid_spacing, double t_final,
                        double gamma, int handle_boundaries, int64_t max_steps, double c_cfl, void *workspace,
                        int64_t *num_iterations, double *t_out, int *result_in_scratch, int *nonfinite,
                        void *stream) {
    if (int e = loop_args(state, scratch, nx, ny, nz, handle_boundaries, workspace)) return e;
    TF_REQUIRE(num_iterations && t_out && result_in_scratch, "null host result pointer");
    hipStream_t st = S_(stream);
    const tf_sim_desc d = loop_desc(state, scratch, nx, ny, nz, grid_spacing, t_final, gamma, handle_boundaries,
                                    max_steps, c_cfl, workspace);
    int tx, ly;
    loop_tiling_for(nx, ny, nz, tx, ly);
    PlanGuard plan;
    if (int e = tf_batch_create(&plan.b, &d, 1, plan_workspace(workspace), tf_batch_workspace_bytes(1), tx, ly,
                                stream))
        return e;
    if (int e = tf_sim_reset(&d, stream)) return e;
    // Steps run in chunks; a step whose control block says inactive returns
    // at once. The host reads the control block after each chunk and sizes
    // the next one from the remaining steps or the remaining time at the
    // current dt, so a run needs few synchronisations.
    const Ctrl *C = reinterpret_cast<const Ctrl *>(d.ctrl);
    Ctrl host;
    int64_t chunk = max_steps >= 0 && std::isinf(t_final) ? max_steps : 4;
    int64_t last_n = -1;
    for (;;) {
        chunk = std::max<int64_t>(1, std::min<int64_t>(chunk, 4096));
        if (max_steps != 0)
            if (int e = enqueue_steps(plan.b, (int)chunk, st)) return e;
        TF_HIP(hipMemcpyAsync(&host, C, sizeof host, hipMemcpyDeviceToHost, st));
        TF_HIP(hipStreamSynchronize(st));
        if (!host.active) break;
        // an active simulation advances by at least one step per chunk
        if (host.n_iter == last_n)
            return fail(TF_ERR_HIP, "time loop made no progress at step %lld", (long long)host.n_iter);
        last_n = host.n_iter;
        int64_t est = max_steps >= 0 ? max_steps - host.n_iter : (int64_t)1 << 40;
        if (std::isfinite(t_final) && host.dt > 0.0)
            est = std::min<int64_t>(est, (int64_t)std::ceil(std::min((t_final - host.t) / host.dt, 1e9)) + 1);
        chunk = est;
    }
    *num_iterations = host.n_iter;
    *t_out = host.t;
    *result_in_scratch = (int)(host.n_iter & 1);
    if (nonfinite) *nonfinite = host.nonfinite;
    return TF_OK;
}

int tf_autotune_loop(float *state, float *scratch, int nx, int ny, int nz, double grid_spacing, double gamma,
                     int handle_boundaries, void *workspace, int *tile_x, int *tile_y, float *ms_per_step,
                     void *stream) {
    if (int e = loop_args(state, scratch, nx, ny, nz, handle_boundaries, workspace)) return e;
    TF_REQUIRE(tile_x && tile_y, "null pointer");
    hipStream_t st = S_(stream);
    const tf_sim_desc d = loop_desc(state, scratch, nx, ny, nz, grid_spacing, INFINITY, gamma, handle_boundaries,
                                    -1, 0.4, workspace);
    // dt from the state; the control block stays at step 0 (no dt update is
    // launched), so every timed launch reads state and writes scratch.
    if (int e = tf_sim_reset(&d, stream)) return e;
    const double cells = (double)nx * ny * nz;
    const bool big = cells > 67108864.0;
    const std::vector<int> txs{2, 4, 8};
    std::vector<int> lys = big ? std::vector<int>{16, 32, 64, 128, 256}
                               : std::vector<int>{0, 1, 2, 4, 8, 16, 32, 64, 128};
    EventPair ev;
    TF_HIP(ev.create());
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    float best = 1e30f;
    int btx = default_loop_tx(nx, ny, nz), bly = 0, bline = 1;
    std::vector<std::tuple<int, int, int>> seen;  // (tx, resolved ly, line waves)
    int rc = TF_OK;
    const std::vector<int> lines = line_allowed(nz) ? std::vector<int>{1, 0} : std::vector<int>{0};
    for (int line : lines) {
        for (int tx : txs) {
            if (tx > 2 && tx > nx) continue;
            for (int ly : lys) {
                if (ly > ny) continue;
                PlanGuard plan;
                tl_line_force = line;
                rc = tf_batch_create(&plan.b, &d, 1, plan_workspace(workspace), tf_batch_workspace_bytes(1), tx, ly,
                                     stream);
                tl_line_force = -1;
                if (rc != TF_OK) break;
                const SimDev &S = plan.b->h_sims[0];
                const auto key = std::make_tuple(S.line_w > 0 ? line_tx(S.line_w, tx) : tx, S.ly, S.line_w);
                if (std::find(seen.begin(), seen.end(), key) != seen.end()) continue;
                seen.push_back(key);
                const int reps = big ? 2 : 5;
                hipError_t e = launch_stencils(plan.b, st, false);
                if (e == hipSuccess) e = hipEventRecord(e0, st);
                for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_stencils(plan.b, st, false);
                if (e == hipSuccess) e = hipEventRecord(e1, st);
                if (e == hipSuccess) e = hipEventSynchronize(e1);
                float ms = 0.f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                if (e != hipSuccess) {
                    rc = fail(TF_ERR_HIP, "autotune launch: %s", hipGetErrorString(e));
                    break;
                }
                ms /= reps;
                if (ms < best) {
                    best = ms;
                    btx = tx;
                    bly = S.ly;
                    bline = S.line_w > 0 ? 1 : 0;
                }
            }
            if (rc != TF_OK) break;
        }
        if (rc != TF_OK) break;
    }
    if (rc != TF_OK) return rc;
    {
        std::lock_guard<std::mutex> lk(g_tiling_mu);
        g_loop_tiling[{nx, ny, nz}] = {btx, bly, bline};
    }
    *tile_x = btx;
    *tile_y = bly;
    if (ms_per_step) *ms_per_step = best;
    return TF_OK;
}

// ------------------------------------------------------------------ allocation-free loop
namespace {
// Device memory of tf_time_integration_noalloc: the workspace of one
// single-grid loop (header, totals partials, one-simulation plan tables),
// one copy per device, held until the process ends. Host code uses only the
// address hipGetSymbolAddress returns for it (per device), never the name.
constexpr size_t kStaticScratchBytes = 48 * 1024;
__device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_init_mu;
void *g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch`, per device
std::mutex g_scratch_call_mu[kMaxDevices];  // one call at a time per device
}  // namespace

int tf_time_integration_noalloc(float *state, float *buf1, int nx, int ny, int nz, double grid_spacing,
                                double t_final, double gamma, int handle_boundaries, int64_t max_steps, double c_cfl,
                                int64_t *num_iterations, double *t_out, int *result_in_scratch, int *nonfinite,
                                void *stream) {
    const size_t need = tf_workspace_bytes(nx, ny, nz);
    if (need > kStaticScratchBytes)
        return fail(TF_ERR_UNSUPPORTED, "tf_time_integration_noalloc: a loop needs %zu bytes of scratch, the static "
                                        "array holds %zu", need, kStaticScratchBytes);
    int dev = -1;
    TF_HIP(hipGetDevice(&dev));
    TF_REQUIRE(dev >= 0 && dev < kMaxDevices, "tf_time_integration_noalloc: device %d beyond the %d it serves", dev,
               kMaxDevices);
    void *ws = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_init_mu);
        if (!g_scratch_ptr[dev]) {
            void *p = nullptr;
            TF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            g_scratch_ptr[dev] = p;
        }
        ws = g_scratch_ptr[dev];
    }
    // The loop blocks until its work is done, so calls on one device take
    // turns on the array; after an error the stream is drained before the
    // next call may reuse it.
    std::lock_guard<std::mutex> lk(g_scratch_call_mu[dev]);
    const int rc = tf_time_integration(state, buf1, nx, ny, nz, grid_spacing, t_final, gamma, handle_boundaries,
                                       max_steps, c_cfl, ws, num_iterations, t_out, result_in_scratch, nonfinite,
                                       stream);
    if (rc != TF_OK) {
        const std::string msg = g_err;
        (void)hipStreamSynchronize(S_(stream));
        return fail(rc, "%s", msg.c_str());
    }
    return TF_OK;
}
