/*
 * tinyfluids.h -- C ABI of libtfhip.so, the gfx950 (MI355X) implementation of
 * the jax_tinyfluids time step: first-order Godunov, HLL fluxes, periodic
 * neighbours, dimensionally split X -> Y -> Z sweeps and the CFL time step.
 *
 * Conventions shared by every entry point:
 *  - State buffers are device pointers to float32 arrays laid out as the
 *    reference's STATE_TYPE, state[5][nx][ny][nz] in C order (z fastest),
 *    holding PRIMITIVE variables rho, vx, vy, vz, p at indices 0..4
 *    (reference tinyfluids/jax_tinyfluids/fluid.py:10-22). Conversions take
 *    and return the conserved layout rho, mx, my, mz, E.
 *  - `stream` is a hipStream_t passed as void*; NULL is the default stream.
 *    Work is enqueued asynchronously unless an entry point says otherwise.
 *  - `workspace` is device memory of at least tf_workspace_bytes(...) bytes,
 *    owned by the caller. The library never allocates or frees device memory.
 *  - Return value: TF_OK (0) or a negative TF_ERR_* code; tf_last_error()
 *    gives a message (thread-local).
 *  - Arithmetic: float32 per cell; the signal-speed maximum is exact (a max of
 *    float32 values), dt = C_CFL * dx / max, t and dt are float64.
 *
 * Reference paths below are relative to the reference repository root.
 */
#ifndef TINYFLUIDS_H
#define TINYFLUIDS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TF_OK 0
#define TF_ERR_ARG (-1)         /* bad argument: null pointer, bad size, bad axis */
#define TF_ERR_HIP (-2)         /* a HIP runtime call failed */
#define TF_ERR_NONFINITE (-3)   /* non-finite signal speed seen (negative p, NaN) */
#define TF_ERR_UNSUPPORTED (-4) /* a shape or option this build does not run */

#define TF_ABI_VERSION 1

/* Library ABI version (TF_ABI_VERSION). No reference counterpart. */
int tf_abi_version(void);

/* Message of the last error on this thread ("" if none). No reference counterpart. */
const char *tf_last_error(void);

/* Bytes of workspace the entry points below need for an nx*ny*nz grid.
 * No reference counterpart (JAX allocates internally). */
size_t tf_workspace_bytes(int nx, int ny, int nz);

/* Replaces conserved_state_from_primitive (fluid.py:28-67):
 * cons = (rho, rho*v, p/(gamma-1) + rho*|v|^2/2) for ncells cells.
 * prim and cons are [5][ncells]; they may alias. */
int tf_conserved_state_from_primitive(const float *prim, float *cons, int64_t ncells, double gamma, void *stream);

/* Replaces primitive_state_from_conserved (fluid.py:69-103):
 * prim = (rho, m/rho, (gamma-1)*(E - |m|^2/(2 rho))). May alias. */
int tf_primitive_state_from_conserved(const float *cons, float *prim, int64_t ncells, double gamma, void *stream);

/* Replaces _cfl_time_step (fluid.py:425-504): writes the float64
 * dt = C_CFL * grid_spacing / max_signal_speed to the device pointer dt_out,
 * where max_signal_speed is the maximum over periodic interfaces and axes of
 * max(|S_R|, |S_L|) with the Davis wave speeds of fluid.py:119-160; it equals
 * the per-cell maximum of |v_d| + c exactly, which is what is evaluated.
 * max_speed_out (device, may be NULL) receives that float maximum. */
int tf_cfl_time_step(const float *prim, int nx, int ny, int nz, double grid_spacing, double gamma, double c_cfl,
                     double *dt_out, float *max_speed_out, void *workspace, void *stream);

/* Replaces speed_of_sound (fluid.py:105-117): c = sqrt(gamma p / rho) for n
 * cells (rho, p, c: n floats each). */
int tf_speed_of_sound(const float *rho, const float *p, float *c, int64_t n, double gamma, void *stream);

/* Replaces get_wave_speeds (fluid.py:119-160): the Davis speeds of ncells
 * interfaces between primitive states primitives_left and primitives_right
 * ([5][ncells] each), S_R = max(u_L + c_L, u_R + c_R, 0) and
 * S_L = min(u_L - c_L, u_R - c_R, 0) with u the velocity component
 * flux_direction_index (1..3); NaN propagates as in the reference. */
int tf_get_wave_speeds(const float *primitives_left, const float *primitives_right, int64_t ncells, double gamma,
                       int flux_direction_index, float *wave_speeds_left_minus, float *wave_speeds_right_plus,
                       void *stream);

/* Replaces get_max_wave_speeds (fluid.py:162-195): max(max |S_R|, max |S_L|)
 * over the ncells interfaces, written to the device float max_out. */
int tf_get_max_wave_speeds(const float *primitives_left, const float *primitives_right, int64_t ncells, double gamma,
                           int flux_direction_index, float *max_out, void *workspace, void *stream);

/* Replaces _euler_flux (fluid.py:206-259): the conserved flux along
 * flux_direction_index (1..3) of ncells primitive cells ([5][ncells]). */
int tf_euler_flux(const float *primitive_state, float *flux, int64_t ncells, double gamma, int flux_direction_index,
                  void *stream);

/* Replaces _hll_solver (fluid.py:261-305): the HLL flux
 * (S_R F_L - S_L F_R + S_L S_R (U_R - U_L)) / (S_R - S_L) at ncells
 * interfaces. flux must not alias the states. */
int tf_hll_solver(const float *primitives_left, const float *primitives_right, float *flux, int64_t ncells,
                  double gamma, int flux_direction_index, void *stream);

/* Replaces _set_along_axis (fluid.py:317-331): out = in with the slice
 * set_index along array axis `axis` of the [5][nx][ny][nz] state (0 = the
 * variables, 1..3 = x, y, z) replaced by the slice get_index; negative
 * indices count from the end, as in Python. in and out must not alias. */
int tf_set_along_axis(const float *in, float *out, int nx, int ny, int nz, int axis, int set_index, int get_index,
                      void *stream);

/* Replaces _open_boundaries(state, num_ghost_cells) (fluid.py:317-413): along
 * x, then y, then z, the first g cells take the value of cell g and then the
 * last g cells that of cell n-g-1, so out[v][x][y][z] =
 * in[v][s(x)][s(y)][s(z)] with s(i) = L(i >= n-g ? n-g-1 : i),
 * L(j) = max(j, g); for g = 1 (what _evolve_state uses, fluid.py:572) s is
 * clamp(., 1, n-2). Requires g >= 0 and every n >= g + 1. in and out must
 * not alias. */
int tf_open_boundaries(const float *in, float *out, int nx, int ny, int nz, int num_ghost_cells, void *stream);

/* Replaces _evolve_state_along_axis (fluid.py:506-557): one split sweep along
 * axis (1 = x, 2 = y, 3 = z) with the HLL flux at every periodic interface:
 * U += dt/dx * (F[i-1/2] - F[i+1/2]), then back to primitives. dt is a
 * device pointer to a float64. in and out must not alias. */
int tf_evolve_state_along_axis(const float *in, float *out, int nx, int ny, int nz, double grid_spacing,
                               const double *dt, double gamma, int axis, void *stream);

/* Replaces _evolve_state (fluid.py:560-578): [open boundaries if
 * handle_boundaries] then the X, Y and Z sweeps, fused into one pass over
 * HBM. dt is a device pointer to a float64. in and out must not alias.
 * If max_speed_out is not NULL, the maximum per-cell |v_d| + c of the new
 * state is max-accumulated into it (as float bits; clear it first).
 * With a workspace and nz <= 1024 the step runs whole z-lines per workgroup
 * and keeps its table and control block in the workspace header: calls that
 * share a workspace must be ordered on one stream. Without a workspace it
 * runs 62-cell z-tiles. */
int tf_evolve_state(const float *in, float *out, int nx, int ny, int nz, double grid_spacing, const double *dt,
                    double gamma, int handle_boundaries, unsigned int *max_speed_out, void *workspace, void *stream);

/* Replaces time_integration(..., shard_mapped=False)
 * (time_integration.py:94-128, loop at :63-92): while t < t_final (and fewer
 * than max_steps steps when max_steps >= 0): dt = _cfl_time_step(state);
 * state = _evolve_state(state, dt); t += dt; n += 1. No dt clipping; a NaN dt
 * ends the loop after the step that used it, as in the reference.
 * state and scratch are two device buffers of the state's size; the result
 * ends in state if *result_in_scratch == 0, else in scratch. Blocks until the
 * run is done (the reference returns after block_until_ready()).
 * num_iterations, t_out, result_in_scratch, nonfinite are host pointers
 * (nonfinite may be NULL). */
int tf_time_integration(float *state, float *scratch, int nx, int ny, int nz, double grid_spacing, double t_final,
                        double gamma, int handle_boundaries, int64_t max_steps, double c_cfl, void *workspace,
                        int64_t *num_iterations, double *t_out, int *result_in_scratch, int *nonfinite,
                        void *stream);

/* tf_time_integration without a workspace argument: allocates nothing. No
 * reference counterpart; for callers that cannot allocate device memory
 * (serving loops with a fixed footprint, harnesses that forbid allocation).
 * The workspace lives in a 48 KiB array of device memory that the library
 * holds on each device until the process ends; a loop of any supported grid
 * needs tf_workspace_bytes() <= 48 KiB of it (larger needs: TF_ERR_UNSUPPORTED).
 * Calls on one device take turns on that array (the call blocks until its
 * run is done); arguments, results and errors as tf_time_integration. */
int tf_time_integration_noalloc(float *state, float *scratch, int nx, int ny, int nz, double grid_spacing,
                                double t_final, double gamma, int handle_boundaries, int64_t max_steps, double c_cfl,
                                int64_t *num_iterations, double *t_out, int *result_in_scratch, int *nonfinite,
                                void *stream);

/* Totals of mass, momentum x/y/z and energy of a primitive state, each cell
 * converted and summed in float64 in a fixed order (deterministic). out5 is a
 * device pointer to 5 doubles. No reference counterpart (diagnostic). */
int tf_conserved_totals(const float *prim, int nx, int ny, int nz, double gamma, double *out5, void *workspace,
                        void *stream);

/* ------------------------------------------------------------------------
 * Batched engine: many independent simulations of uneven sizes advanced
 * together. No reference counterpart (the reference runs one grid per call).
 * ------------------------------------------------------------------------ */

#define TF_CTRL_BYTES 128

/* One simulation. buf0/buf1 are its two state buffers ([5][nx][ny][nz]
 * float32 each); after tf_sim_reset the state is in buf0, and after n steps
 * in buf(n % 2). ctrl is TF_CTRL_BYTES of device memory owned by the caller
 * that holds the simulation's t, dt, step count and flags on the device. */
typedef struct tf_sim_desc {
    float *buf0;
    float *buf1;
    void *ctrl;
    int nx, ny, nz;
    int handle_boundaries;
    double grid_spacing;
    double gamma;
    double c_cfl;
    double t_final;     /* +inf for a fixed step count */
    int64_t max_steps;  /* < 0: no limit */
} tf_sim_desc;

/* Host view of a simulation's control block (its first 48 bytes). */
typedef struct tf_sim_status {
    double t;          /* simulated time reached */
    double dt;         /* dt of the next step */
    int64_t n_iter;    /* steps taken; the state is in buf(n_iter % 2) */
    int32_t active;    /* 1 while the simulation still steps */
    int32_t nonfinite; /* 1 if a non-finite signal speed was seen */
    float max_speed;   /* max |v_d| + c of the current state */
    int32_t pad[3];
} tf_sim_status;

/* Sets t = 0, n = 0, active, and the first dt from the state in buf0
 * (the loop head of time_integration.py:63-72). Asynchronous. */
int tf_sim_reset(const tf_sim_desc *sim, void *stream);

/* Opaque batch plan: the device tables of a set of simulations. */
typedef struct tf_batch tf_batch;

/* Workspace bytes for a plan over nsims simulations. */
size_t tf_batch_workspace_bytes(int nsims);

/* Builds a plan; its device tables live in the caller's workspace.
 * tile_x (x-cells per wave: 2, 4, 6 or 8) and tile_y (rows per y-chunk) choose
 * the stencil tiling; 0 picks the library default. Grids with nz <= 1024 run
 * whole z-lines per workgroup (1..16 waves) where that uses the lanes at
 * least as well as 62-cell z-tiles, with tile width min(tile_x, 4), or 2
 * above 8 waves. */
int tf_batch_create(tf_batch **out, const tf_sim_desc *sims, int nsims, void *workspace, size_t workspace_bytes,
                    int tile_x, int tile_y, void *stream);

/* tf_batch_create without waiting, for callers that rebuild plans while
 * earlier steps are still queued on the stream: the tables are uploaded on
 * `stream` from page-locked staging memory the library keeps (and reuses
 * once the upload has completed), and the call returns at once. The plan
 * may be stepped on `stream` at once, or on another stream after an event
 * recorded on `stream`. */
int tf_batch_create_async(tf_batch **out, const tf_sim_desc *sims, int nsims, void *workspace,
                          size_t workspace_bytes, int tile_x, int tile_y, void *stream);

int tf_batch_destroy(tf_batch *b);

/* The tiling a plan runs simulation `sim` (caller's order) with: tile width,
 * y-chunk rows, and waves per z-line (0: 62-cell z-tiles of one wave). */
int tf_batch_sim_tiling(const tf_batch *b, int sim, int *tile_x, int *tile_y, int *line_waves);

/* Enqueues `steps` batched steps: each is one fused multi-simulation stencil
 * launch over every simulation's tiles plus one per-simulation dt update.
 * Simulations that are no longer active are skipped on the device. */
int tf_batch_step(tf_batch *b, int steps, void *stream);

/* As tf_batch_step, but the steps run as replays of captured HIP graphs:
 * one of 16 steps (fewer when `steps` < 16) launched steps / 16 times, and
 * one of the remainder. Each graph is captured on first use and again when
 * its step count changes; the stencil launches of a step run concurrently
 * inside it. */
int tf_batch_step_graph(tf_batch *b, int steps, void *stream);

/* As tf_batch_step for 1..1024 steps, as one captured HIP graph of the
 * whole run launched once: captured on first use and again when the step
 * count changes, so repeated runs of a plan replay a single launch. */
int tf_batch_run_graph(tf_batch *b, int steps, void *stream);

/* As tf_batch_step, also recording HIP events around every stencil launch
 * on `stream`; waits for the steps and returns in *stencil_ms the summed
 * duration of the stencil launches alone (without the dt updates). */
int tf_batch_step_timed(tf_batch *b, int steps, float *stencil_ms, void *stream);

/* Process-wide default for plans and single steps built afterwards: 1
 * (default) runs grids with nz <= 1024 in the line tiling, whole periodic
 * z-lines per workgroup, where that keeps at least as many lanes busy as
 * 62-column tiles (and the loop tuner did not measure 62-column tiles faster
 * for the shape); 0 uses the 62-column tiling everywhere. Results are
 * identical. */
int tf_set_line_mode(int enabled);

/* tf_sim_reset of every simulation of the plan, in three launches (control
 * blocks, signal-speed maxima of every buf0, first dt). Asynchronous. */
int tf_batch_reset(tf_batch *b, void *stream);

/* Copies every simulation's status to host memory (synchronous). */
int tf_batch_status(tf_batch *b, tf_sim_status *host_out, void *stream);

/* tf_batch_status in two halves, for callers that keep further steps queued
 * while they read a status. tf_batch_status_enqueue enqueues on `stream`
 * (the stream the batch steps on) a copy of every simulation's status into
 * `host_raw`, tf_batch_status_bytes(b) bytes of page-locked host memory (with
 * pageable memory the copy blocks), and returns. Once the stream has passed
 * the copy (an event recorded after it has completed), tf_batch_status_unpack
 * converts `host_raw` into one tf_sim_status per simulation, in the order
 * the simulations were given to tf_batch_create. */
size_t tf_batch_status_bytes(const tf_batch *b);
int tf_batch_status_enqueue(tf_batch *b, void *host_raw, void *stream);
int tf_batch_status_unpack(const tf_batch *b, const void *host_raw, tf_sim_status *host_out);

/* ------------------------------------------------------------------------
 * Tiling and autotuning of the fused stencil. No reference counterpart.
 * ------------------------------------------------------------------------ */

/* Copies bytes (a multiple of 16, 16-byte aligned, non-overlapping) with the
 * flat one-float4-per-thread kernel: one read and one write of every byte.
 * Its rate is the practical HBM ceiling of a read+write kernel that the
 * fused step's rate is compared with. Asynchronous on the stream. No
 * reference counterpart. */
int tf_hbm_copy(const void *src, void *dst, size_t bytes, void *stream);

/* Tiling used for an nx*ny*nz grid by tf_evolve_state; tf_set_tiling also
 * fixes it for tf_time_integration. */
int tf_get_tiling(int nx, int ny, int nz, int *tile_x, int *tile_y);
int tf_set_tiling(int nx, int ny, int nz, int tile_x, int tile_y);

/* Times the candidate tilings of the fused step on (in -> out), keeps the
 * fastest for this shape (tf_set_tiling) and reports it. Synchronous. */
int tf_autotune_evolve(const float *in, float *out, int nx, int ny, int nz, double grid_spacing, double gamma,
                       void *workspace, int *tile_x, int *tile_y, float *ms_per_step, void *stream);

/* Times the candidate tilings of the step that tf_time_integration runs
 * (tile width 2/4/8 x y-chunk rows, halo or whole-z-line tiling) on this
 * state, keeps the fastest for this shape and reports it (tile_y as
 * resolved). The halo/line choice also holds for tf_evolve_state (with a
 * workspace) and for batch plans of this shape, until tf_set_tiling resets
 * it to whole lines. Reads state, writes scratch; synchronous. No reference
 * counterpart (JAX compiles per shape on the first call instead). */
int tf_autotune_loop(float *state, float *scratch, int nx, int ny, int nz, double grid_spacing, double gamma,
                     int handle_boundaries, void *workspace, int *tile_x, int *tile_y, float *ms_per_step,
                     void *stream);

#ifdef __cplusplus
}
#endif

#endif /* TINYFLUIDS_H */
