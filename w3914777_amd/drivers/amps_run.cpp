// amps_run.cpp — command-line drivers with the reference applications' surface,
// written against the C ABI (include/amps.h) only.
//
//   amps_run <model> purehigh <run_idx> <csv>     PURE_FP64     (snowflake|graingrowth/purehigh.cu)
//   amps_run <model> end2end_mix <run_idx> <csv>  ADAPTIVE      (snowflake|graingrowth/end2end_mix.cu)
//   amps_run <model> purelow                      PURE_HALF2    (snowflake|graingrowth/purelow.cu)
//   amps_run snowflake puremix                    static map    (snowflake/puremix.cu)
//   amps_run sintering main [<para_idx> <dtype>]  mode from the paras flags (sintering/main.cu)
//   amps_run <model> monitor_independent_withmonitor     ADAPTIVE, time of steps 5..N-1 with the
//                                                        monitor passes (…/monitor_independent_withmonitor.cu)
//   amps_run <model> monitor_independent_withoutmonitor  ADAPTIVE, sweep-kernel time of steps 5..N-1
//                                                        only (…/monitor_independent_withoutmonitor.cu:85-127)
//   amps_run <model> motivation_mix <csv>                N FP64 steps, then a tile map (AMSTENCIL:
//                                                        ranked monitor values grown by BFS to the
//                                                        FP64 share 1-ratio; GRAM1/2: static map) and
//                                                        50 steps on it (…/motivation_mix.cu, tools.h:66-106)
//   (<model> snowflake or graingrowth for the last three)
//
// <model> is snowflake, graingrowth or sintering.  Options:
//   --paras FILE   parameter file in the reference's header syntax; repeatable,
//                  applied in order over the model's defaults (the reference
//                  compiles total.h and paraK.h in; here they are read at run time)
//   --steps N      override `timesteps`
//   --ratio R      low-tile ratio of the GRAM/puremix static maps (total.h `ratio`)
//   --gram 1|2     static map kind: 1 = first ratio*N tiles low, 2 = tile%100 < 100*ratio
//                  low (snowflake/function.h:72-79); default from `#define GRAM1/2`
//
// Outputs (current directory), as the reference writes them:
//   time_tmp.csv  milliseconds of steps 5..N-1 (HIP events, the reference's cudaEvent pair)
//   result CSV    header ",0,1,...,X-1", rows "y,v0,...,v(X-1)", stream default format
//                 (snowflake/tools.h:5-23): phi (snowflake), eta2 (grain growth), c (sintering)
// Errors print the status and message and exit with status 1.
#include "amps.h"

#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <queue>
#include <sstream>
#include <string>
#include <vector>

namespace {

int g_model = 0;

[[noreturn]] void die(const std::string& what, int rc) {
    std::cerr << "amps_run: " << what << ": " << amps_status_string(rc) << " (" << amps_last_error() << ")\n";
    std::exit(1);
}

void check(int rc, const char* what) {
    if (rc != AMPS_OK) die(what, rc);
}

[[noreturn]] void usage(const char* msg) {
    if (msg) std::cerr << "amps_run: " << msg << "\n";
    std::cerr << "usage: amps_run <snowflake|graingrowth|sintering> <purehigh|end2end_mix|purelow|puremix|main|\n"
                 "                monitor_independent_withmonitor|monitor_independent_withoutmonitor|motivation_mix>\n"
                 "                [reference arguments] [--paras FILE]... [--steps N] [--ratio R] [--gram 1|2]\n";
    std::exit(2);
}

// writetocsv of the reference (snowflake/tools.h:5-23): default ostream format.
void write_csv(const std::string& path, const std::vector<double>& v, int X, int Y) {
    std::ofstream f(path);
    if (!f) {
        std::cerr << "amps_run: cannot write " << path << "\n";
        std::exit(1);
    }
    f << ",";
    for (int x = 0; x < X; ++x) f << x << (x == X - 1 ? "\n" : ",");
    for (int y = 0; y < Y; ++y) {
        f << y << ",";
        for (int x = 0; x < X; ++x) f << v[(size_t)y * X + x] << (x == X - 1 ? "\n" : ",");
    }
}

void write_time(double ms) {
    std::ofstream f("time_tmp.csv");
    f << (float)ms;
}

void mkdir_p(const std::string& d) { (void)mkdir(d.c_str(), 0777); }

// The field each application writes: phi, eta2 or c.
int result_field() { return g_model == AMPS_GRAINGROWTH ? 1 : 0; }

std::vector<double> field(amps_ctx* ctx, const amps_config& c, int f) {
    std::vector<double> v((size_t)c.nx * c.ny);
    check(amps_get_field(ctx, 0, f, v.data(), v.size()), "amps_get_field");
    return v;
}

// Static tile map of the GRAM comparators / puremix (snowflake/function.h:72-79,
// puremix.cu:26-32): 1 = low.
std::vector<int32_t> static_map(const amps_config& c, int kind, double ratio) {
    const int tx = c.nx / c.tile, ty = c.ny / c.tile, n = tx * ty;
    std::vector<int32_t> t(n);
    for (int i = 0; i < n; ++i)
        t[i] = kind == 2 ? ((i % 100) < 100.0 * ratio ? 1 : 2) : (i < n * ratio ? 1 : 2);
    return t;
}

// Monitor values of the motivation study on host FP64 fields, one per tile
// (the reference evaluates get_max_diff1/2 on its FP64 state,
// snowflake/tools.h:177-211): Monitor1 = max - min of the 4x4 sample lattice
// {0,(t-1)/3,2(t-1)/3,t-1} starting from min 1, max 0; Monitor2 = max |last -
// now| on {t/8,3t/8,5t/8,7t/8}.  No neighbour max (get_type is not used).
std::vector<double> tile_max_diff(const amps_config& c, const std::vector<double>& now,
                                  const std::vector<double>& last) {
    const int t = c.tile, tx = c.nx / t, ty = c.ny / t;
    std::vector<double> md((size_t)tx * ty);
    for (int j = 0; j < ty; ++j)
        for (int i = 0; i < tx; ++i) {
            double mn = 1.0, mx = 0.0, d = 0.0;
            for (int b = 0; b < 4; ++b)
                for (int a = 0; a < 4; ++a) {
                    if (c.monitor == AMPS_MON_SPATIAL) {
                        const int oy = b == 0 ? 0 : (b == 1 ? (t - 1) / 3 : (b == 2 ? 2 * (t - 1) / 3 : t - 1));
                        const int ox = a == 0 ? 0 : (a == 1 ? (t - 1) / 3 : (a == 2 ? 2 * (t - 1) / 3 : t - 1));
                        const double v = now[(size_t)(j * t + oy) * c.nx + i * t + ox];
                        if (v < mn) mn = v;
                        if (v > mx) mx = v;
                    } else {
                        const int oy = (2 * b + 1) * t / 8, ox = (2 * a + 1) * t / 8;
                        const size_t k = (size_t)(j * t + oy) * c.nx + i * t + ox;
                        const double v = std::fabs(last[k] - now[k]);
                        if (v > d) d = v;
                    }
                }
            md[(size_t)j * tx + i] = c.monitor == AMPS_MON_SPATIAL ? mx - mn : d;
        }
    return md;
}

// Tile map of the motivation study (snowflake/motivation_mix.cu:75-103): the
// tiles ranked by monitor value (descending, ties in index order, as the
// reference's bubble sort leaves them) are made FP64 while their value is
// nonzero, up to high = (int)(N - N*ratio) tiles; the set then grows breadth
// first over the 8 periodic neighbours (tools.h:83-106) until it holds `high`
// tiles.  As in the reference, the first seed and the first grown tile are
// set before the count is checked.
std::vector<int32_t> motivation_map(const amps_config& c, const std::vector<double>& md, double ratio) {
    const int tx = c.nx / c.tile, ty = c.ny / c.tile, n = tx * ty;
    const int high = (int)((double)n - (double)n * ratio);
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return md[a] > md[b]; });
    std::vector<int32_t> type(n, 1);
    std::queue<int> q;
    int sum = 0;
    for (int j = 0; j < n && md[idx[j]] != 0.0; ++j) {
        type[idx[j]] = 2;
        ++sum;
        q.push(idx[j]);
        if (sum >= high) break;
    }
    auto neighbour = [&](int k, int dir) {
        int x = k % tx, y = k / tx;
        if (dir == 1 || dir == 5 || dir == 8) x = (x - 1 + tx) % tx;
        if (dir == 2 || dir == 6 || dir == 7) x = (x + 1) % tx;
        if (dir == 3 || dir == 5 || dir == 6) y = (y - 1 + ty) % ty;
        if (dir == 4 || dir == 7 || dir == 8) y = (y + 1) % ty;
        return y * tx + x;
    };
    while (!q.empty()) {
        const int k = q.front();
        q.pop();
        for (int dir = 1; dir <= 8; ++dir) {
            const int nb = neighbour(k, dir);
            if (type[nb] != 2) {
                ++sum;
                type[nb] = 2;
                q.push(nb);
                if (sum >= high) break;
            }
        }
        if (sum >= high) break;
    }
    return type;
}

// Runs `steps` steps; returns the event time of steps 5..steps-1 (the
// reference starts its events at i == 5).
double run_timed(amps_ctx* ctx, int steps) {
    double ms = 0.0;
    const int warm = steps > 5 ? 5 : steps;
    check(amps_step(ctx, warm), "amps_step");
    check(amps_sync(ctx, nullptr), "amps_sync");
    if (steps > 5) {
        check(amps_step(ctx, steps - 5), "amps_step");
        check(amps_sync(ctx, &ms), "amps_sync");
    }
    return ms;
}

// `#define GRAM1` / `#define GRAM2` in a parameter file (the C ABI maps both
// onto STATIC_MAP; the map kind is the driver's business).
int gram_kind_of(const std::vector<std::string>& files) {
    int kind = 0;
    for (const auto& p : files) {
        std::ifstream f(p);
        std::string line;
        while (std::getline(f, line)) {
            std::istringstream ws(line);
            std::string a, b;
            ws >> a >> b;
            if (a == "#define" && b == "GRAM1") kind = 1;
            if (a == "#define" && b == "GRAM2") kind = 2;
        }
    }
    return kind;
}

}  // namespace

// motivation_mix <csv>: `timesteps` FP64 steps, a tile map, 50 steps on it,
// then the result field (snowflake phi, grain growth eta2).
int motivation(amps_config cfg, int timesteps, const std::vector<std::string>& pos, int gram, double ratio) {
    if (pos.size() != 1) usage("motivation_mix expects <csv>");
    const bool ranked = gram == 0;  // AMSTENCIL: ranked monitor values
    if (ranked && cfg.monitor != AMPS_MON_SPATIAL && cfg.monitor != AMPS_MON_TEMPORAL)
        usage("motivation_mix ranks tiles by Monitor1 or Monitor2");
    if (ranked && cfg.monitor == AMPS_MON_TEMPORAL && timesteps < 10)
        usage("motivation_mix with Monitor2 needs at least 10 steps");
    cfg.mode = AMPS_STATIC_MAP;
    amps_ctx* ctx = nullptr;
    check(amps_create(&cfg, &ctx), "amps_create");
    check(amps_init(ctx), "amps_init");
    amps_info info;
    check(amps_get_info(ctx, &info), "amps_get_info");
    const int tiles = (cfg.nx / cfg.tile) * (cfg.ny / cfg.tile);
    const int groups = info.type_groups > 0 ? info.type_groups : 1;
    auto impose = [&](const std::vector<int32_t>& m) {  // one map for every type group
        std::vector<int32_t> all((size_t)groups * tiles);
        for (int g = 0; g < groups; ++g) std::copy(m.begin(), m.end(), all.begin() + (size_t)g * tiles);
        check(amps_set_tile_types(ctx, 0, all.data(), all.size()), "amps_set_tile_types");
    };
    impose(std::vector<int32_t>(tiles, 2));  // the first phase is the FP64 update
    const int f = result_field();
    std::vector<double> last;
    if (ranked && cfg.monitor == AMPS_MON_TEMPORAL) {
        // Monitor2's history: the field after step timesteps-10 (0-based), i.e.
        // after timesteps-9 steps (motivation_mix.cu:63-68)
        check(amps_step(ctx, timesteps - 9), "amps_step");
        check(amps_sync(ctx, nullptr), "amps_sync");
        last = field(ctx, cfg, f);
        check(amps_step(ctx, 9), "amps_step");
    } else {
        check(amps_step(ctx, timesteps), "amps_step");
    }
    check(amps_sync(ctx, nullptr), "amps_sync");
    impose(ranked ? motivation_map(cfg, tile_max_diff(cfg, field(ctx, cfg, f), last), ratio)
                  : static_map(cfg, gram, ratio));
    check(amps_step(ctx, 50), "amps_step");
    check(amps_sync(ctx, nullptr), "amps_sync");
    write_csv(pos[0], field(ctx, cfg, f), cfg.nx, cfg.ny);
    amps_destroy(ctx);
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 3) usage(nullptr);
    const std::string model = argv[1], drv = argv[2];
    if (model == "snowflake") g_model = AMPS_SNOWFLAKE;
    else if (model == "graingrowth") g_model = AMPS_GRAINGROWTH;
    else if (model == "sintering") g_model = AMPS_SINTERING;
    else usage("unknown model");
    std::vector<std::string> pos, paras;
    int steps_override = -1, gram = 0;
    double ratio = 1.0;
    for (int i = 3; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) usage(("missing value after " + a).c_str());
            return argv[++i];
        };
        if (a == "--paras") paras.push_back(next());
        else if (a == "--steps") steps_override = std::atoi(next().c_str());
        else if (a == "--ratio") ratio = std::atof(next().c_str());
        else if (a == "--gram") gram = std::atoi(next().c_str());
        else if (a.rfind("--", 0) == 0) usage(("unknown option " + a).c_str());
        else pos.push_back(a);
    }
    amps_config cfg;
    check(amps_default_config(g_model, &cfg), "amps_default_config");
    int timesteps = 500;
    for (const auto& p : paras) {
        int ts = -1;
        check(amps_load_paras(p.c_str(), &cfg, &ts), ("amps_load_paras " + p).c_str());
        if (ts >= 0) timesteps = ts;
    }
    if (steps_override >= 0) timesteps = steps_override;
    if (gram == 0) gram = gram_kind_of(paras);

    bool write_result = false;
    std::string out_csv;
    if (drv == "purehigh" || drv == "end2end_mix") {
        if (g_model == AMPS_SINTERING) usage("sintering uses the `main` driver");
        if (pos.size() != 2) usage("expects <run_idx> <csv>");
        cfg.mode = drv == "purehigh" ? AMPS_PURE_FP64 : AMPS_ADAPTIVE;
        write_result = pos[0] == "4";  // the harness records the field on its last run
        out_csv = pos[1];
    } else if (drv == "purelow") {
        if (g_model == AMPS_SINTERING) usage("sintering uses the `main` driver");
        if (!pos.empty()) usage("purelow takes no arguments");
        cfg.mode = AMPS_PURE_HALF2;
        write_result = true;
        out_csv = "half2result.csv";
    } else if (drv == "puremix") {
        if (g_model != AMPS_SNOWFLAKE) usage("puremix is a snowflake driver");
        if (!pos.empty()) usage("puremix takes no arguments");
        cfg.mode = AMPS_STATIC_MAP;
        write_result = true;
        out_csv = gram == 2 ? "gram2.csv" : (gram == 1 ? "gram1.csv" : "amstencil.csv");
    } else if (drv == "monitor_independent_withmonitor" || drv == "monitor_independent_withoutmonitor" ||
               drv == "motivation_mix") {
        if (g_model == AMPS_SINTERING) usage((drv + " is a snowflake / grain-growth driver").c_str());
        if (drv == "motivation_mix") return motivation(cfg, timesteps, pos, gram, ratio);
        if (!pos.empty()) usage((drv + " takes no arguments").c_str());
        cfg.mode = AMPS_ADAPTIVE;
        amps_ctx* ctx = nullptr;
        check(amps_create(&cfg, &ctx), "amps_create");
        check(amps_init(ctx), "amps_init");
        double ms = 0.0;
        if (drv == "monitor_independent_withmonitor") {
            ms = run_timed(ctx, timesteps);
        } else {
            // steps 0..4 untimed; then the sweep kernels alone (monitor passes and
            // conversions run but are not timed)
            const int warm = timesteps > 5 ? 5 : timesteps;
            check(amps_step(ctx, warm), "amps_step");
            check(amps_sync(ctx, nullptr), "amps_sync");
            if (timesteps > 5) {
                check(amps_set_kernel_timing(ctx, 1), "amps_set_kernel_timing");
                check(amps_step(ctx, timesteps - 5), "amps_step");
                check(amps_sync(ctx, nullptr), "amps_sync");
                int64_t launches = 0;
                check(amps_get_kernel_time(ctx, &launches, &ms), "amps_get_kernel_time");
            }
        }
        write_time(ms);
        amps_destroy(ctx);
        return 0;
    } else if (drv == "main") {
        if (g_model != AMPS_SINTERING) usage("main is the sintering driver");
        if (!pos.empty() && pos.size() != 2) usage("main expects no arguments or <para_idx> <dtype>");
        write_result = pos.size() == 2;
        if (cfg.mode == AMPS_STATIC_MAP && gram == 0) gram = 1;
    } else {
        usage("unknown driver");
    }

    amps_ctx* ctx = nullptr;
    check(amps_create(&cfg, &ctx), "amps_create");
    check(amps_init(ctx), "amps_init");
    if (cfg.mode == AMPS_STATIC_MAP) {
        const std::vector<int32_t> map = static_map(cfg, gram == 0 ? 1 : gram, ratio);
        check(amps_set_tile_types(ctx, 0, map.data(), map.size()), "amps_set_tile_types");
    }
    const double ms = run_timed(ctx, timesteps);
    if (drv != "puremix") write_time(ms);

    if (drv == "main" && write_result) {
        // ./result/end2end/<dtype>/<variant>/para<idx>/<steps>[_<thr|ratio>].csv (sintering/main.cu:446-600)
        std::string s = "./result/";
        mkdir_p(s);
        s += "end2end/";
        mkdir_p(s);
        s += pos[1] + "/";
        mkdir_p(s);
        std::string name = std::to_string(timesteps);
        if (cfg.mode == AMPS_PURE_FP64) {
            s += "pure/";
        } else if (cfg.mode == AMPS_STATIC_MAP) {
            s += gram == 2 ? "gram2/" : "gram1/";
            name += "_" + std::to_string(ratio);
        } else {
            s += "amstencil/";
            mkdir_p(s);
            if (cfg.monitor == AMPS_MON_SPATIAL) {
                s += "monitor1/";
                name += "_" + std::to_string(cfg.threshold);
            } else {
                s += cfg.monitor == AMPS_MON_GRADIENT ? "monitorgradient/" : "monitor2/";
                std::ostringstream o;
                o << cfg.threshold;
                name += "_" + o.str();
            }
        }
        mkdir_p(s);
        s += "para" + pos[0] + "/";
        mkdir_p(s);
        out_csv = s + name + ".csv";
    }
    if (write_result) {
        write_csv(out_csv, field(ctx, cfg, result_field()), cfg.nx, cfg.ny);
        if (drv == "purelow") std::cout << out_csv << std::endl;
    }
    amps_destroy(ctx);
    return 0;
}
