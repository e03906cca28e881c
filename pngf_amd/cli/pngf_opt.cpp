// pngf_opt: the reference program's driver (pngf-opt/main.cpp) on the gfx950 library.
//
//   pngf_opt [options]
//     --tiles T          tiles per side (21)
//     --layout shared|disjoint   (shared: the reference Yee grid, reads acmefdfd files)
//     --freqs f1,f2,...  frequencies in GHz (25,27.5,30,32.5,35 as main.cpp:116-117)
//     --gdir DIR         read DIR/Gmat_sub_XX.bin (main.cpp:235); default: synthetic G
//     --driver ref|raster, --objective ref|s11, --method lowrank|plain
//     --symmetry auto|off|project   LOWRANK accept (project: a G symmetric to 1e-5, such as
//                        acmefdfd's widened single-precision output, replaced by (G + G^T) / 2)
//     --seed S           MagicHat seed (-514213645, main.cpp:154)
//     --init empty|random (main.cpp:155, 187-219)
//     --log INDEX        write the optimization history to log_INDEX.txt in the reference
//                        Logger's format (logger.hpp:75-107; main.cpp:470-472), at most 10000
//                        records (main.cpp:185)
//
// Prints the reference's start/finish report: initial and final design as ASCII art,
// per-frequency |S11| (dB) and directivity, time elapsed, flips evaluated.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "pngf_optimizer.hpp"

static void print_map(const std::vector<int>& tm, int T) {
    for (int j = 0; j < T; j++) {
        for (int i = 0; i < T; i++) std::putchar(tm[i + ((T - 1) - j) * T] ? 'x' : '.');
        std::putchar('\n');
    }
}

// The Logger's text file (logger.hpp:75-107): "seed records", the start tile map as x/. in
// index order, one line "i tile decision perf..." per record, the final tile map.
static void save_history(const std::string& fname, int seed, const std::vector<int>& start,
                         const std::vector<int>& end, int64_t n, const int32_t* tile, const int32_t* dec,
                         const double* perf, int nperf) {
    std::ofstream out(fname);
    out << seed << " " << n << std::endl;
    for (int v : start) out << (v ? "x" : ".");
    out << std::endl;
    for (int64_t i = 0; i < n; i++) {
        out << i << " " << tile[i] << " " << dec[i] << " ";
        for (int j = 0; j < nperf; j++) out << perf[i * nperf + j] << " ";
        out << std::endl;
    }
    for (int v : end) out << (v ? "x" : ".");
    out << std::endl;
}

int main(int argc, char** argv) {
    int T = 21, layout = PNGF_LAYOUT_SHARED, driver = PNGF_DRIVER_REF, objective = PNGF_OBJ_REF;
    int method = PNGF_METHOD_LOWRANK, seed = -514213645, symmetry = PNGF_SYM_AUTO;
    bool random_init = false;
    int log_index = -1;
    std::string gdir;
    std::vector<double> freqs = {25.0, 27.5, 30.0, 32.5, 35.0};
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto next = [&]() -> std::string { return i + 1 < argc ? argv[++i] : ""; };
        if (a == "--tiles") T = std::atoi(next().c_str());
        else if (a == "--layout") layout = next() == "disjoint" ? PNGF_LAYOUT_DISJOINT : PNGF_LAYOUT_SHARED;
        else if (a == "--driver") driver = next() == "raster" ? PNGF_DRIVER_RASTER : PNGF_DRIVER_REF;
        else if (a == "--objective") objective = next() == "s11" ? PNGF_OBJ_S11 : PNGF_OBJ_REF;
        else if (a == "--method") method = next() == "plain" ? PNGF_METHOD_PLAIN : PNGF_METHOD_LOWRANK;
        else if (a == "--symmetry") {
            const std::string v = next();
            symmetry = v == "off" ? PNGF_SYM_OFF : v == "project" ? PNGF_SYM_PROJECT : PNGF_SYM_AUTO;
        }
        else if (a == "--seed") seed = std::atoi(next().c_str());
        else if (a == "--init") random_init = next() == "random";
        else if (a == "--gdir") gdir = next();
        else if (a == "--log") log_index = std::atoi(next().c_str());
        else if (a == "--freqs") {
            freqs.clear();
            std::string s = next();
            for (size_t p = 0; p < s.size();) {
                size_t q = s.find(',', p);
                freqs.push_back(std::atof(s.substr(p, q - p).c_str()));
                p = q == std::string::npos ? s.size() : q + 1;
            }
        } else {
            std::fprintf(stderr, "unknown option %s\n", a.c_str());
            return 2;
        }
    }
    const int nf = (int)freqs.size();
    const int ne = objective == PNGF_OBJ_REF ? 3 : 1;
    const int max_flips = driver == PNGF_DRIVER_REF ? 4 : 1;
    const double lambda = 0.01, dx = lambda / 60.0, dy = lambda / 60.0, dz = lambda / 180.0;  // main.cpp:157-160
    try {
        pngf::Optimizer opt(layout, T, T, 3, 3, ne, max_flips, nf, dx, dy, dz, method, objective, driver, 0,
                            symmetry);
        // initial design (main.cpp:187-219): mirrored random quadrant or the dipole
        std::vector<int> tm(T * T, 0);
        if (random_init) {
            std::mt19937 gen(seed);
            const int h = (T - 1) / 2;
            for (int j = 0; j <= h; j++)
                for (int i = 0; i <= h; i++) {
                    int v = gen() & 1;
                    tm[i + j * T] = v;
                    if (i < h) tm[(T - 1 - i) + j * T] = v;
                    if (j < h) {
                        tm[i + (T - 1 - j) * T] = v;
                        if (i < h) tm[(T - 1 - i) + (T - 1 - j) * T] = v;
                    }
                }
        }
        const int mx = (T - 1) / 2, my = (T - 1) / 2;
        tm[my * T + mx - 1] = 1;
        tm[my * T + mx + 1] = 1;
        tm[my * T + mx] = 0;
        tm[(my - 1) * T + mx] = 0;
        tm[(my + 1) * T + mx] = 0;
        auto t0 = std::chrono::high_resolution_clock::now();
        for (int f = 0; f < nf; f++) {
            opt.set_frequency(f, freqs[f] * 1e9);
            if (gdir.empty()) opt.synth_g(f);
            else {
                char name[64];
                std::snprintf(name, sizeof(name), "/Gmat_sub_%02d.bin", f + 1);
                opt.load_g(f, (gdir + name).c_str());
            }
        }
        opt.set_tile_map(tm);
        opt.set_source_port(mx, my, 1, 2);
        opt.init_state();
        double init_s = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
        std::printf("Time elapsed: %g\n", init_s);
        print_map(tm, T);
        std::vector<int32_t> acc_t(100000);
        std::vector<int64_t> acc_e(100000);
        std::vector<double> acc_F(100000);
        pngf_dbs_params p = {};
        p.seed = seed;
        p.max_accepts = (int32_t)acc_t.size();
        p.acc_tiles = acc_t.data();
        p.acc_eval = acc_e.data();
        p.acc_F = acc_F.data();
        p.rng_discard = random_init ? ((T - 1) / 2 + 1) * ((T - 1) / 2 + 1) : 0;
        const int nperf = 2 + 3 * nf;
        std::vector<int32_t> h_tile, h_dec;
        std::vector<double> h_perf;
        if (log_index >= 0) {
            const int cap = 10000;
            h_tile.resize(cap);
            h_dec.resize(cap);
            h_perf.resize((size_t)cap * nperf);
            p.hist_tile = h_tile.data();
            p.hist_dec = h_dec.data();
            p.hist_perf = h_perf.data();
            p.hist_cap = cap;
        }
        pngf_dbs_result r = opt.dbs_run(p);
        if (log_index >= 0)
            save_history("log_" + std::to_string(log_index) + ".txt", seed, tm, opt.tile_map(), r.hist_count,
                         h_tile.data(), h_dec.data(), h_perf.data(), nperf);
        std::printf("starting F: %g\n", r.F_start);
        for (int64_t i = 0; i < r.accepts; i++)
            std::printf("%lld: tile %d F %.10g\n", (long long)acc_e[i], acc_t[i], acc_F[i]);
        std::printf("Time elapsed: %g\n", r.seconds);
        std::printf("Converged after %lld flips!\n", (long long)r.flip_count);
        std::printf("Final design:\n\n");
        print_map(opt.tile_map(), T);
    } catch (const pngf::Error& e) {
        std::fprintf(stderr, "pngf error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
