// Host-side check of the batch-plan builder (tinyfluids_amd/csrc/tf_plan.cpp),
// built with the host compiler under address and undefined-behaviour
// sanitizers by tests/test_plan_host.py. It builds plans over seeded uneven
// batches, the benchmark batch, grids of exactly 2^30 cells and bad inputs,
// and checks every table a plan uploads:
//   - the table order is a permutation and each entry is its simulation;
//   - the tiling classes are contiguous, their prefixes consistent;
//   - decoding every wave / workgroup of every launch with the kernels' own
//     decoders (tf_layout.h: XCD block order, halo and line tile decode,
//     far-row predicate) writes every output cell exactly once, the far-row
//     tiles included;
//   - workspace offsets, the host image and the slot addresses stay inside
//     the workspace;
//   - bad requests return an error instead of a plan.
// No device memory is touched: the device addresses are fake and never
// dereferenced. Prints "plan_check: OK <n> plans" on success, else a message
// and exits non-zero.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "tf_plan.h"

using namespace tf;

namespace {

int g_plans = 0;

[[noreturn]] void die(const std::string &msg) {
    std::fprintf(stderr, "plan_check: FAIL %s\n", msg.c_str());
    std::exit(1);
}
#define CHECK(cond, ...)                                                        \
    do {                                                                        \
        if (!(cond)) {                                                          \
            char b_[512];                                                       \
            std::snprintf(b_, sizeof b_, __VA_ARGS__);                          \
            die(std::string(#cond) + ": " + b_);                                \
        }                                                                       \
    } while (0)

constexpr uintptr_t kWs = 0x7f0000000000ull;  // fake, 256-aligned device address

// Fake, distinct device buffers per simulation.
tf_sim_desc desc(int i, int nx, int ny, int nz, int hb) {
    tf_sim_desc d;
    std::memset(&d, 0, sizeof d);
    d.buf0 = reinterpret_cast<float *>(0x100000000000ull + (uintptr_t)i * 0x100000000ull);
    d.buf1 = reinterpret_cast<float *>(0x200000000000ull + (uintptr_t)i * 0x100000000ull);
    d.ctrl = reinterpret_cast<void *>(0x300000000000ull + (uintptr_t)i * 128);
    d.nx = nx;
    d.ny = ny;
    d.nz = nz;
    d.handle_boundaries = hb;
    d.grid_spacing = 1.0 / nx;
    d.gamma = 1.4;
    d.c_cfl = 0.4;
    d.t_final = 1e30;
    d.max_steps = 10 + i;
    return d;
}

// What the library's registry would choose with line mode on.
bool line_default(int nz) { return nz <= 64 * kMaxLineW && line_lane_use(nz) >= halo_lane_use(nz) - 1e-9; }

// Every output cell of simulation k written exactly once by the plan's
// launches (cells counted per (x, y, z-tile) for the halo tiling).
void check_coverage(const PlanTables &P, int k) {
    const SimDev &S = P.sims[k];
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    if (S.line_w > 0) {
        const PlanClass &C = P.cls[S.line_w];
        const int txl = C.tx;
        const int units = P.prefix[C.prefix_at + (k - C.begin) + 1] - P.prefix[C.prefix_at + (k - C.begin)];
        CHECK(units == S.ntx * S.nyc, "sim %d: %d line blocks for ntx %d x nyc %d", k, units, S.ntx, S.nyc);
        CHECK(S.groups * nz <= 64 * S.line_w, "sim %d: %d lines of %d cells in %d waves", k, S.groups, nz, S.line_w);
        std::vector<unsigned char> cnt((size_t)nx * ny, 0);
        for (int local = 0; local < units; ++local) {
            int xt, yc;
            decode_line_block(local, S.ntx, xt, yc);
            const int y0 = yc * S.ly, rows = std::min(S.ly, ny - y0);
            CHECK(rows >= 1, "sim %d: empty y-chunk %d", k, yc);
            for (int g = 0; g < S.groups; ++g)
                for (int s = 0; s < txl; ++s) {
                    const int x = xt * S.groups * txl + g * txl + s;
                    if (x >= nx) continue;
                    for (int y = y0; y < y0 + rows; ++y) ++cnt[(size_t)x * ny + y];
                }
        }
        for (size_t i = 0; i < cnt.size(); ++i)
            CHECK(cnt[i] == 1, "sim %d (%dx%dx%d line W=%d): line (%zu, %zu) written %d times", k, nx, ny, nz,
                  S.line_w, i / ny, i % ny, (int)cnt[i]);
        return;
    }
    const int tx = P.tx;
    CHECK(S.ntz * kZT >= nz && (S.ntz - 1) * kZT < nz, "sim %d: ntz %d for nz %d", k, S.ntz, nz);
    const int units = P.prefix[P.cls[0].prefix_at + (k - P.cls[0].begin) + 1] -
                      P.prefix[P.cls[0].prefix_at + (k - P.cls[0].begin)];
    CHECK(units == S.waves && S.waves == S.ntx * S.ntz * S.nyc, "sim %d: %d waves", k, units);
    std::vector<unsigned char> cnt((size_t)nx * ny * S.ntz, 0);
    auto mark = [&](int xt, int zt, int y0, int rows) {
        for (int s = 0; s < tx; ++s) {
            const int x = xt * tx + s;
            if (x >= nx) continue;
            for (int y = y0; y < y0 + rows; ++y) ++cnt[((size_t)x * ny + y) * S.ntz + zt];
        }
    };
    for (int w = 0; w < units; ++w) {
        int xt, zt, yc;
        decode_tile(w, S.ntz, S.ntx, xt, zt, yc);
        if (S.far && touches_last_row(S, tx, xt, yc)) continue;  // the far kernel's
        const int y0 = yc * S.ly;
        mark(xt, zt, y0, std::min(S.ly, ny - y0));
    }
    for (const FarTile &f : P.far) {
        if (f.sim != k) continue;
        CHECK(S.far, "sim %d: far tile without a 2^30-cell field", k);
        CHECK(f.xt >= 0 && f.xt < S.ntx && f.rows >= 1 && f.y0 >= 0 && f.y0 + f.rows <= ny, "sim %d: far tile",
              k);
        CHECK(touches_last_row(S, tx, f.xt, f.y0 / S.ly), "sim %d: far tile off the last row", k);
        CHECK(f.y0 / S.ly == (f.y0 + f.rows - 1) / S.ly, "sim %d: far tile spans two y-chunks", k);
        for (int zt = 0; zt < S.ntz; ++zt) mark(f.xt, zt, f.y0, f.rows);
    }
    for (size_t i = 0; i < cnt.size(); ++i)
        CHECK(cnt[i] == 1, "sim %d (%dx%dx%d halo): cell group %zu written %d times", k, nx, ny, nz, i,
              (int)cnt[i]);
}

// Launch-level decode: every logical block / wave of a launch of `total`
// units taken exactly once by the XCD-aware block order.
void check_xcd_order(int nblocks) {
    if (nblocks <= 0) return;
    const int grid = ((nblocks + kXcd - 1) / kXcd) * kXcd;
    std::vector<unsigned char> seen(nblocks, 0);
    for (int b = 0; b < grid; ++b) {
        const int lb = xcd_logical_block(b, nblocks);
        if (lb < 0) continue;
        CHECK(lb < nblocks, "logical block %d of %d", lb, nblocks);
        ++seen[lb];
    }
    for (int i = 0; i < nblocks; ++i) CHECK(seen[i] == 1, "logical block %d seen %d times", i, (int)seen[i]);
}

void check_plan(const PlanRequest &rq, const PlanTables &P) {
    ++g_plans;
    const int n = rq.nsims;
    CHECK(P.nsims == n && (int)P.sims.size() == n && (int)P.order.size() == n, "sizes");
    std::vector<int> seen(n, 0);
    for (int k = 0; k < n; ++k) {
        const int i = P.order[k];
        CHECK(i >= 0 && i < n, "order[%d] = %d", k, i);
        ++seen[i];
        const SimDev &S = P.sims[k];
        const tf_sim_desc &d = rq.sims[i];
        CHECK(S.nx == d.nx && S.ny == d.ny && S.nz == d.nz && S.buf[0] == d.buf0 && S.buf[1] == d.buf1 &&
                  (void *)S.ctrl == d.ctrl && S.open_bc == (d.handle_boundaries ? 1 : 0),
              "table entry %d is not simulation %d", k, i);
        CHECK(S.ly >= 1 && S.ly <= S.ny && S.nyc == (S.ny + S.ly - 1) / S.ly, "sim %d: ly %d nyc %d", i, S.ly, S.nyc);
        const bool want_line = rq.line && rq.line[i];
        CHECK((S.line_w > 0) == want_line, "sim %d: line tiling %d, asked %d", i, S.line_w, (int)want_line);
        const uintptr_t mp = reinterpret_cast<uintptr_t>(S.maxp);
        CHECK(mp == rq.workspace + P.off_maxp + (size_t)k * kSlotBytesPerSim && mp + kSlotBytesPerSim <= rq.workspace + P.bytes,
              "sim %d: signal-speed slots outside the workspace", i);
    }
    for (int i = 0; i < n; ++i) CHECK(seen[i] == 1, "simulation %d appears %d times", i, seen[i]);
    // classes: contiguous, in order, each holding only its own kind
    int k = 0, launches = 0;
    for (int c = 0; c < kClasses; ++c) {
        const PlanClass &C = P.cls[c];
        CHECK(C.w == c && C.begin == k && C.count >= 0, "class %d", c);
        CHECK(C.tx == (c == 0 ? P.tx : line_tx(c, P.tx)), "class %d tile width %d", c, C.tx);
        for (int j = 0; j < C.count; ++j) CHECK(P.sims[k + j].line_w == c, "class %d holds sim of W %d", c, P.sims[k + j].line_w);
        CHECK(C.prefix_at >= 0 && C.prefix_at + C.count < (int)P.prefix.size() && P.prefix[C.prefix_at] == 0,
              "class %d prefix start", c);
        for (int j = 0; j < C.count; ++j)
            CHECK(P.prefix[C.prefix_at + j + 1] > P.prefix[C.prefix_at + j], "class %d prefix not increasing", c);
        CHECK(C.total == P.prefix[C.prefix_at + C.count], "class %d total", c);
        if (C.count > 0) {
            ++launches;
            check_xcd_order(c == 0 ? (C.total + 3) / 4 : C.total);
        }
        k += C.count;
    }
    CHECK(k == n, "classes hold %d of %d", k, n);
    CHECK(P.nlaunch == launches + (P.far.empty() ? 0 : 1), "nlaunch %d", P.nlaunch);
    // workspace carving and the upload image
    CHECK(P.off_sims == 0 && P.off_sims < P.off_prefix && P.off_prefix < P.off_status && P.off_status < P.off_reset &&
              P.off_reset < P.off_speed && P.off_speed < P.off_tables_end && P.off_tables_end == P.off_far &&
              P.off_far < P.off_maxp && P.off_maxp < P.bytes,
          "workspace offsets out of order");
    CHECK(P.bytes <= rq.workspace_bytes && P.bytes == plan_workspace_bytes(n), "plan bytes %zu", P.bytes);
    CHECK(P.off_prefix - P.off_sims >= n * sizeof(SimDev) && P.off_status - P.off_prefix >= P.prefix.size() * sizeof(int),
          "tables overlap");
    CHECK(P.image.size() == P.off_tables_end, "image size");
    CHECK(std::memcmp(P.image.data() + P.off_sims, P.sims.data(), n * sizeof(SimDev)) == 0, "image: sims");
    CHECK(std::memcmp(P.image.data() + P.off_prefix, P.prefix.data(), P.prefix.size() * sizeof(int)) == 0,
          "image: prefix");
    CHECK(P.far.size() * sizeof(FarTile) <= P.off_maxp - P.off_far, "far tiles exceed their region");
    const int *sp = reinterpret_cast<const int *>(P.image.data() + P.off_speed);
    CHECK(sp[0] == 0 && sp[n] == P.speed_blocks, "speed prefix");
    const ResetParams *rp = reinterpret_cast<const ResetParams *>(P.image.data() + P.off_reset);
    for (int j = 0; j < n; ++j) {
        const tf_sim_desc &d = rq.sims[P.order[j]];
        const int64_t cells = (int64_t)d.nx * d.ny * d.nz;
        CHECK(sp[j + 1] - sp[j] == (int)((cells + kSpeedCellsPerBlock - 1) / kSpeedCellsPerBlock), "speed blocks %d", j);
        CHECK(rp[j].max_steps == d.max_steps && rp[j].dx == d.grid_spacing, "reset params %d", j);
    }
    for (int j = 0; j < n; ++j)
        if ((int64_t)P.sims[j].nx * P.sims[j].ny * (P.sims[j].line_w ? 1 : P.sims[j].ntz) <= (int64_t)1 << 25)
            check_coverage(P, j);
}

PlanTables make(const std::vector<tf_sim_desc> &d, const std::vector<unsigned char> &line, int tx, int ty,
                const std::vector<int> *tuned_ly = nullptr, int tuned_tx = 0) {
    PlanRequest rq;
    rq.sims = d.data();
    rq.nsims = (int)d.size();
    rq.workspace = kWs;
    rq.workspace_bytes = plan_workspace_bytes(rq.nsims);
    rq.tile_x = tx;
    rq.tile_y = ty;
    rq.line = line.data();
    rq.tuned_ly = tuned_ly ? tuned_ly->data() : nullptr;
    rq.tuned_tx = tuned_tx;
    PlanTables P;
    std::string err;
    const int rc = build_plan(rq, P, err);
    CHECK(rc == TF_OK, "build_plan failed: %s", err.c_str());
    check_plan(rq, P);
    return P;
}

int expect_error(const PlanRequest &rq, int code, const char *what) {
    PlanTables P;
    std::string err;
    const int rc = build_plan(rq, P, err);
    CHECK(rc == code && !err.empty(), "%s: got %d (%s), want %d", what, rc, err.c_str(), code);
    return rc;
}

void bad_requests() {
    std::vector<tf_sim_desc> d{desc(0, 8, 8, 8, 0), desc(1, 16, 9, 70, 1)};
    std::vector<unsigned char> line{1, 1};
    PlanRequest rq;
    rq.sims = d.data();
    rq.nsims = 2;
    rq.workspace = kWs;
    rq.workspace_bytes = plan_workspace_bytes(2);
    rq.line = line.data();
    PlanRequest r = rq;
    r.nsims = 0;
    expect_error(r, TF_ERR_ARG, "no simulations");
    r = rq;
    r.sims = nullptr;
    expect_error(r, TF_ERR_ARG, "null sims");
    r = rq;
    r.workspace_bytes -= 1;
    expect_error(r, TF_ERR_ARG, "small workspace");
    r = rq;
    r.workspace += 8;
    expect_error(r, TF_ERR_ARG, "misaligned workspace");
    r = rq;
    r.tile_x = 3;
    expect_error(r, TF_ERR_ARG, "tile_x 3");
    r = rq;
    r.tile_y = -1;
    expect_error(r, TF_ERR_ARG, "tile_y -1");
    std::vector<tf_sim_desc> e = d;
    e[1].nz = 0;
    r = rq;
    r.sims = e.data();
    expect_error(r, TF_ERR_ARG, "zero extent");
    e = d;
    e[0].buf1 = nullptr;
    r.sims = e.data();
    expect_error(r, TF_ERR_ARG, "null buffer");
    e = d;
    e[0] = desc(0, 1, 8, 8, 1);
    r.sims = e.data();
    expect_error(r, TF_ERR_ARG, "open boundaries on an axis of 1");
    e = d;
    e[1] = desc(1, 8, 8, 2000, 0);
    r.sims = e.data();
    expect_error(r, TF_ERR_ARG, "whole lines of 2000 cells");
    e = d;
    e[1] = desc(1, 1024, 1024, 1025, 0);
    r.sims = e.data();
    line[1] = 0;
    expect_error(r, TF_ERR_UNSUPPORTED, "more than 2^30 cells");
    // far_list of the single-step path: at most kMaxFarPerSim tiles
    for (int tx : {2, 4, 6, 8}) {
        SimDev S = make_simdev(1024, 1024, 1024, 0, 1.4, tx, 8);
        FarList F;
        std::string err;
        CHECK(far_list(S, tx, F, err) == TF_OK && F.n >= 1 && F.n <= kMaxFarPerSim, "far list tx %d: %d", tx, F.n);
    }
}

}  // namespace

int main() {
    std::mt19937 rng(12345);
    auto uni = [&](int lo, int hi) { return std::uniform_int_distribution<int>(lo, hi)(rng); };
    // the benchmark's uneven batch shapes: 256 cubes of {32,48,64,96,128}
    {
        const int sides[5] = {32, 48, 64, 96, 128};
        std::vector<tf_sim_desc> d;
        std::vector<unsigned char> line;
        for (int i = 0; i < 256; ++i) {
            const int n = sides[uni(0, 4)];
            d.push_back(desc(i, n, n, n, 0));
            line.push_back(line_default(n) ? 1 : 0);
        }
        for (int tx : {0, 2, 4, 6, 8}) make(d, line, tx, 0);
        make(d, line, 4, 3);
        std::vector<unsigned char> none(d.size(), 0);
        make(d, none, 8, 16);
    }
    // seeded uneven batches: any extents, open boundaries, line and halo
    // tilings mixed, tile widths, y-chunks and measured tilings
    for (int rep = 0; rep < 150; ++rep) {
        const int n = uni(1, 24);
        std::vector<tf_sim_desc> d;
        std::vector<unsigned char> line;
        std::vector<int> tuned;
        for (int i = 0; i < n; ++i) {
            const int nx = uni(1, 40), ny = uni(1, 40);
            const int nz = uni(0, 5) == 0 ? uni(1025, 1300) : uni(1, 300);
            const int hb = (nx >= 2 && ny >= 2 && nz >= 2 && uni(0, 2) == 0) ? 1 : 0;
            d.push_back(desc(i, nx, ny, nz, hb));
            line.push_back(nz <= 64 * kMaxLineW && uni(0, 3) != 0 ? 1 : 0);
            tuned.push_back(uni(0, 2) == 0 ? 0 : uni(1, 50));
        }
        const int txs[5] = {0, 2, 4, 6, 8};
        const int tx = txs[uni(0, 4)];
        const int ty = uni(0, 3) == 0 ? uni(1, 45) : 0;
        make(d, line, tx, ty, uni(0, 1) ? &tuned : nullptr, uni(0, 1) ? txs[uni(1, 4)] : 0);
    }
    // fields of exactly 2^30 cells: the far-row tiles of the halo tiling
    for (int tx : {2, 4, 6, 8})
        for (int ty : {0, 1, 5, 64, 1024}) {
            std::vector<tf_sim_desc> d{desc(0, 1024, 1024, 1024, 0), desc(1, 20, 30, 40, 0), desc(2, 8, 128, 1 << 20, 1)};
            std::vector<unsigned char> line{0, 1, 0};
            make(d, line, tx, ty);
        }
    {
        std::vector<tf_sim_desc> d{desc(0, 1024, 1024, 1024, 0), desc(1, 1536, 512, 512, 0)};
        std::vector<unsigned char> line{1, 1};
        make(d, line, 0, 0);
        make(d, line, 2, 128);
    }
    bad_requests();
    std::printf("plan_check: OK %d plans\n", g_plans);
    return 0;
}
