// Host-side planning of the fused stencil (declared in tf_plan.h): tilings,
// batch-plan tables and their place in the caller's workspace. No HIP calls.
#include "tf_plan.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

namespace tf {

namespace {
int set_err(std::string &err, int code, const char *fmt, ...) __attribute__((format(printf, 3, 4)));
int set_err(std::string &err, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    err = buf;
    return code;
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }
}  // namespace

int check_dims(int nx, int ny, int nz, std::string &err) {
    if (!(nx >= 1 && ny >= 1 && nz >= 1))
        return set_err(err, TF_ERR_ARG, "grid dimensions must be >= 1 (got %d x %d x %d)", nx, ny, nz);
    const double cells = (double)nx * ny * nz;
    if (!(cells * 5.0 < 9.0e18)) return set_err(err, TF_ERR_ARG, "grid too large");
    if (!((double)nx * ny < 2.0e9 && (double)ny * nz < 2.0e9))
        return set_err(err, TF_ERR_ARG, "grid planes exceed 2^31 cells");
    return TF_OK;
}

// The fused stencil addresses each field with 32-bit byte offsets.
int check_fused_size(int nx, int ny, int nz, std::string &err) {
    if ((double)nx * ny * nz > 1073741824.0)
        return set_err(err, TF_ERR_UNSUPPORTED, "fused step: %d x %d x %d exceeds 2^30 cells per field", nx, ny, nz);
    return TF_OK;
}

int check_open_bc(int open_bc, int nx, int ny, int nz, std::string &err) {
    if (open_bc && !(nx >= 2 && ny >= 2 && nz >= 2))
        return set_err(err, TF_ERR_ARG, "open boundaries need every dimension >= 2");
    return TF_OK;
}

bool valid_tx(int tx) { return tx == 2 || tx == 4 || tx == 6 || tx == 8; }

int default_tile_x(int nx) { return nx >= 8 ? 8 : (nx >= 4 ? 4 : 2); }

// Rows per y-chunk: enough waves to fill the chip (256 CUs, a few waves per
// SIMD) with several rounds so the last round is short, without letting the
// two halo rows per chunk cost much.
int default_tile_y(int nx, int ny, int nz, int tx) {
    const long ntx = (nx + tx - 1) / tx, ntz = (nz + kZT - 1) / kZT;
    const long per_chunk = ntx * ntz;
    const long target = 256L * 4 * 3 * 6;  // ~6 rounds of full occupancy
    int ly = ny;
    while (ly > 32 && per_chunk * ((ny + ly - 1) / ly) < target) ly = (ly + 1) / 2;
    return std::max(1, ly);
}

// Untuned loop tiling: narrow tiles for grids too small to fill the chip,
// so that each wave's serial stream stays short.
int default_loop_tx(int nx, int ny, int nz) {
    const double cells = (double)nx * ny * nz;
    if (cells >= 16777216.0) return default_tile_x(nx);
    if (cells >= 2097152.0) return nx >= 4 ? 4 : 2;
    return 2;
}

SimDev make_simdev(int nx, int ny, int nz, int open_bc, double gamma, int tx, int ly) {
    SimDev S;
    std::memset(&S, 0, sizeof S);
    S.nx = nx;
    S.ny = ny;
    S.nz = nz;
    S.open_bc = open_bc ? 1 : 0;
    // gamma is rounded to float32 first, as in the reference's float32 runs
    S.gamma = (float)gamma;
    S.gm1 = S.gamma - 1.0f;
    S.inv_gm1 = 1.0f / S.gm1;
    S.ly = std::max(1, std::min(ly, ny));
    S.ntx = (nx + tx - 1) / tx;
    S.ntz = (nz + kZT - 1) / kZT;
    S.nyc = (ny + S.ly - 1) / S.ly;
    S.waves = S.ntx * S.ntz * S.nyc;
    // fields of exactly 4 GiB: the last dword is out of reach of a buffer resource
    S.far = ((int64_t)nx * ny * nz == (int64_t)1 << 30) ? 1 : 0;
    return S;
}

void far_tiles(const SimDev &S, int tx, int sim, std::vector<FarTile> &out, bool split) {
    if (!S.far || S.line_w > 0) return;  // line-tiled grids reach the last cell themselves
    for (int yc = 0; yc < S.nyc; ++yc)
        for (int xt = 0; xt < S.ntx; ++xt) {
            if (!touches_last_row(S, tx, xt, yc)) continue;
            const int y0 = yc * S.ly, rows = std::min(S.ly, S.ny - y0);
            const int sub = split ? std::max(4, (rows + kFarSplit - 1) / kFarSplit) : rows;
            for (int r = 0; r < rows; r += sub) out.push_back(FarTile{sim, xt, y0 + r, std::min(sub, rows - r)});
        }
}

int far_list(const SimDev &S, int tx, FarList &F, std::string &err) {
    std::vector<FarTile> v;
    far_tiles(S, tx, 0, v, false);
    std::memset(&F, 0, sizeof F);
    if (v.size() > (size_t)kMaxFarPerSim)
        return set_err(err, TF_ERR_UNSUPPORTED, "%zu far-row tiles exceed the %d a single step takes", v.size(),
                       kMaxFarPerSim);
    F.n = (int)v.size();
    for (int i = 0; i < F.n; ++i) F.t[i] = v[i];
    return TF_OK;
}

// The fewest waves that hold a line, unless a few more (up to 8) hold
// several lines with at least 2% fewer idle lanes: 48-cell lines go four to
// three waves, 96-cell lines two to three waves (measured 6% faster on 48^3
// and 96^3).
void line_shape(int nz, int &w, int &groups) {
    const int w0 = (nz + 63) / 64;
    w = w0;
    groups = std::max(1, 64 * w0 / nz);
    double best = (double)(groups * nz) / (64.0 * w0);
    for (int cw = w0 + 1; cw <= std::min(kMaxLineW, std::max(w0, 8)); ++cw) {
        const int g = 64 * cw / nz;
        const double use = (double)(g * nz) / (64.0 * cw);
        if (use > best + 0.02) {
            best = use;
            w = cw;
            groups = g;
        }
    }
}

double halo_lane_use(int nz) { return (double)nz / (64.0 * ((nz + kZT - 1) / kZT)); }

double line_lane_use(int nz) {
    int w, g;
    line_shape(nz, w, g);
    return (double)(g * nz) / (64.0 * w);
}

// At most 4 (registers for two waves per SIMD), and 2 above 8 waves per
// workgroup (four waves per SIMD).
int line_tx(int w, int tx) { return w > 8 ? 2 : std::min(tx, 4); }

void make_line(SimDev &S, int tx) {
    int w, groups;
    line_shape(S.nz, w, groups);
    const int txl = line_tx(w, tx);
    S.line_w = w;
    S.groups = groups;
    S.ntx = (S.nx + groups * txl - 1) / (groups * txl);
    S.ntz = 1;
    S.waves = S.ntx * S.nyc * w;
}

size_t plan_workspace_bytes(int nsims) {
    if (nsims < 0) nsims = 0;
    return align256((size_t)nsims * sizeof(SimDev)) + align256((size_t)(nsims + kClasses) * sizeof(int)) +
           align256((size_t)nsims * sizeof(Ctrl)) +
           align256((size_t)nsims * kMaxFarPerSim * kFarSplit * sizeof(FarTile)) +
           align256((size_t)nsims * sizeof(ResetParams)) + align256((size_t)(nsims + 1) * sizeof(int)) +
           align256((size_t)nsims * kSlotBytesPerSim);
}

int build_plan(const PlanRequest &rq, PlanTables &P, std::string &err) {
    const int nsims = rq.nsims;
    if (nsims < 1) return set_err(err, TF_ERR_ARG, "nsims must be >= 1");
    if (!rq.sims || !rq.workspace) return set_err(err, TF_ERR_ARG, "null pointer");
    if (rq.workspace % 16 != 0) return set_err(err, TF_ERR_ARG, "workspace must be 16-byte aligned");
    const size_t need = plan_workspace_bytes(nsims);
    if (rq.workspace_bytes < need)
        return set_err(err, TF_ERR_ARG, "workspace too small: %zu < %zu", rq.workspace_bytes, need);
    if (rq.tile_y < 0) return set_err(err, TF_ERR_ARG, "tile_y must be >= 0");
    if (rq.tile_x != 0 && !valid_tx(rq.tile_x))
        return set_err(err, TF_ERR_ARG, "tile_x must be 0, 2, 4, 6 or 8 (got %d)", rq.tile_x);
    if (rq.tuned_tx != 0 && !valid_tx(rq.tuned_tx))
        return set_err(err, TF_ERR_ARG, "tuned tile width must be 2, 4, 6 or 8 (got %d)", rq.tuned_tx);
    double total_cells = 0.0, max_cells = 0.0;
    for (int i = 0; i < nsims; ++i) {
        const tf_sim_desc &d = rq.sims[i];
        if (int e = check_dims(d.nx, d.ny, d.nz, err)) return e;
        if (int e = check_fused_size(d.nx, d.ny, d.nz, err)) return e;
        if (int e = check_open_bc(d.handle_boundaries, d.nx, d.ny, d.nz, err)) return e;
        if (!(d.buf0 && d.buf1 && d.ctrl)) return set_err(err, TF_ERR_ARG, "null device pointer in sim %d", i);
        if (rq.line && rq.line[i] && d.nz > 64 * kMaxLineW)
            return set_err(err, TF_ERR_ARG, "sim %d: whole z-lines need nz <= %d", i, 64 * kMaxLineW);
        if (rq.tuned_ly && rq.tuned_ly[i] < 0) return set_err(err, TF_ERR_ARG, "sim %d: negative tuned tile_y", i);
        const double c = (double)d.nx * d.ny * d.nz;
        total_cells += c;
        max_cells = std::max(max_cells, c);
    }
    // Default tile width: 8 for work that saturates HBM (fewest halo
    // columns); 4 below that, where more resident waves per SIMD hide the
    // latency of many short streams.
    const int tx = rq.tile_x != 0 ? rq.tile_x : (rq.tuned_tx != 0 ? rq.tuned_tx : (total_cells >= 268435456.0 ? 8 : 4));
    auto build = [&](int ly_cap, std::vector<SimDev> &tab) {
        long total = 0;
        for (int i = 0; i < nsims; ++i) {
            const tf_sim_desc &d = rq.sims[i];
            const int tuned = rq.tuned_ly ? rq.tuned_ly[i] : 0;
            const int ly = rq.tile_y > 0 ? rq.tile_y : (tuned > 0 ? tuned : std::max(1, std::min(d.ny, ly_cap)));
            SimDev S = make_simdev(d.nx, d.ny, d.nz, d.handle_boundaries, d.gamma, tx, ly);
            S.buf[0] = d.buf0;
            S.buf[1] = d.buf1;
            S.ctrl = reinterpret_cast<Ctrl *>(d.ctrl);
            if (rq.line && rq.line[i]) make_line(S, tx);
            tab[i] = S;
            total += S.waves;
        }
        return total;
    };
    // A batched step lasts at least as long as its longest wave's serial
    // stream (rows x tile width x ~0.37 us per column-row at gfx950 clocks),
    // and at best as long as its HBM traffic (40 B per cell at ~4.5 TB/s).
    // Without a given tile_y, y-chunks are capped so that the longest stream
    // stays within half the traffic time.
    int ly_cap = 1 << 30;
    if (rq.tile_y == 0) {
        const double step_us = total_cells * 40.0 / 4.5e6;
        const double row_us = tx * 0.37;
        ly_cap = (int)std::max(4.0, std::min(1.0e9, 0.5 * step_us / row_us - 2.0));
    }
    std::vector<SimDev> tab(nsims);
    long total = build(ly_cap, tab);
    if (rq.tile_y == 0) {
        // Whole y-lines per wave unless the batch is too small to fill the
        // chip several times over; then halve every simulation's y-chunks
        // (down to 4 rows: a small batch is bound by the length of each
        // wave's serial stream, not by the halo rows).
        const long target = 256L * 4 * 3 * 6;
        int maxny = 1;
        for (int i = 0; i < nsims; ++i) maxny = std::max(maxny, rq.sims[i].ny);
        for (int cap = ly_cap; total < target && cap > 4;) {
            cap = std::max(4, std::min(cap, maxny) / 2);
            total = build(cap, tab);
        }
    }
    if (total > 2000000000L) return set_err(err, TF_ERR_ARG, "batch too large");
    P = PlanTables();
    P.nsims = nsims;
    P.tx = tx;
    // Longest waves first (rows per wave) within each class, so that the
    // short ones fill the tail of every launch.
    P.order.resize(nsims);
    for (int i = 0; i < nsims; ++i) P.order[i] = i;
    std::stable_sort(P.order.begin(), P.order.end(), [&](int a, int c) {
        if (tab[a].line_w != tab[c].line_w) return tab[a].line_w < tab[c].line_w;
        const long la = (long)std::min(tab[a].ly, tab[a].ny), lc = (long)std::min(tab[c].ly, tab[c].ny);
        return la > lc;
    });
    P.sims.resize(nsims);
    for (int k = 0; k < nsims; ++k) P.sims[k] = tab[P.order[k]];
    // classes: contiguous table ranges, each with its own prefix (waves or workgroups)
    P.total_waves = 0;
    for (int c = 0, k = 0; c < kClasses; ++c) {
        PlanClass &C = P.cls[c];
        C.w = c;
        C.tx = c == 0 ? tx : line_tx(c, tx);
        C.begin = k;
        C.prefix_at = (int)P.prefix.size();
        P.prefix.push_back(0);
        while (k < nsims && P.sims[k].line_w == c) {
            const SimDev &S = P.sims[k];
            const int units = c == 0 ? S.waves : S.ntx * S.nyc;
            P.prefix.push_back(P.prefix.back() + units);
            P.total_waves += S.waves;
            ++k;
        }
        C.count = k - C.begin;
        C.total = P.prefix.back();
    }
    // workspace carving
    size_t o = 0;
    P.off_sims = o;
    o += align256((size_t)nsims * sizeof(SimDev));
    P.off_prefix = o;
    o += align256((size_t)(nsims + kClasses) * sizeof(int));
    P.off_status = o;
    o += align256((size_t)nsims * sizeof(Ctrl));
    P.off_reset = o;
    o += align256((size_t)nsims * sizeof(ResetParams));
    P.off_speed = o;
    o += align256((size_t)(nsims + 1) * sizeof(int));
    P.off_tables_end = o;
    P.off_far = o;
    o += align256((size_t)nsims * kMaxFarPerSim * kFarSplit * sizeof(FarTile));
    P.off_maxp = o;
    o += align256((size_t)nsims * kSlotBytesPerSim);
    P.bytes = o;
    if (P.bytes > rq.workspace_bytes || P.prefix.size() > (size_t)(nsims + kClasses))
        return set_err(err, TF_ERR_ARG, "plan tables exceed the workspace");
    for (int k = 0; k < nsims; ++k)
        P.sims[k].maxp = reinterpret_cast<uint32_t *>(rq.workspace + P.off_maxp + (size_t)k * kSlotBytesPerSim);
    std::vector<ResetParams> rp(nsims);
    std::vector<int> sp(nsims + 1, 0);
    int64_t speed_blocks = 0;
    for (int k = 0; k < nsims; ++k) {
        const tf_sim_desc &d = rq.sims[P.order[k]];
        rp[k] = ResetParams{d.grid_spacing, d.c_cfl, d.t_final, d.max_steps};
        speed_blocks += ((int64_t)d.nx * d.ny * d.nz + kSpeedCellsPerBlock - 1) / kSpeedCellsPerBlock;
        if (speed_blocks > 2000000000) return set_err(err, TF_ERR_ARG, "batch too large");
        sp[k + 1] = (int)speed_blocks;
    }
    P.speed_blocks = (int)speed_blocks;
    for (int k = 0; k < nsims; ++k) {
        const size_t before = P.far.size();
        far_tiles(P.sims[k], tx, k, P.far, true);
        if (P.far.size() - before > (size_t)kMaxFarPerSim * kFarSplit)
            return set_err(err, TF_ERR_UNSUPPORTED, "sim %d: %zu far-row tiles exceed the plan's %d", P.order[k],
                           P.far.size() - before, kMaxFarPerSim * kFarSplit);
        if (P.far.size() > before) P.far_max_ntz = std::max(P.far_max_ntz, P.sims[k].ntz);
    }
    P.nlaunch = P.far.empty() ? 0 : 1;
    for (int c = 0; c < kClasses; ++c) P.nlaunch += P.cls[c].count > 0 ? 1 : 0;
    // The host image of the workspace range [off_sims, off_tables_end): the
    // status region in it is scratch.
    P.image.assign(P.off_tables_end, 0);
    std::memcpy(P.image.data() + P.off_sims, P.sims.data(), (size_t)nsims * sizeof(SimDev));
    std::memcpy(P.image.data() + P.off_prefix, P.prefix.data(), P.prefix.size() * sizeof(int));
    std::memcpy(P.image.data() + P.off_reset, rp.data(), rp.size() * sizeof(ResetParams));
    std::memcpy(P.image.data() + P.off_speed, sp.data(), sp.size() * sizeof(int));
    return TF_OK;
}

}  // namespace tf
