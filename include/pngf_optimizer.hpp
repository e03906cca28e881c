// pngf_optimizer.hpp -- header-only C++ view of the C ABI (pngf_gpu.h) with the
// method names of the reference optimizer (pngf-opt/pngfopt.hpp:18-679), so that a
// main.cpp-shaped driver reads the same.  One object holds every frequency; the
// per-frequency public members of the reference (newEobj, Eobj, Vinc, LN1inc,
// LN2inc, tile_map) become accessors indexed by frequency.
//
// Errors: every call checks the status and throws pngf::Error carrying
// pngf_last_error(); the C ABI itself never throws.
#pragma once
#include <complex>
#include <stdexcept>
#include <string>
#include <vector>

#include "pngf_gpu.h"

namespace pngf {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

class Optimizer {
public:
    typedef std::complex<double> cplx;

    // reference: PNGF_Optimizer::init(file, num_tiles_x, num_tiles_y, cells_x, cells_y,
    //            num_eobj, max_flips, dx, dy, dz) -- here for all frequencies at once.
    Optimizer(int layout, int tiles_x, int tiles_y, int cells_x, int cells_y, int num_eobj, int max_flips,
              int num_freq, double dx, double dy, double dz, int method = PNGF_METHOD_LOWRANK,
              int objective = PNGF_OBJ_REF, int driver = PNGF_DRIVER_REF, int device = 0,
              int symmetry = PNGF_SYM_AUTO) {
        pngf_config c = {};
        c.layout = layout; c.tiles_x = tiles_x; c.tiles_y = tiles_y; c.cells_x = cells_x; c.cells_y = cells_y;
        c.num_eobj = num_eobj; c.max_flips = max_flips; c.num_freq = num_freq; c.method = method;
        c.objective = objective; c.driver = driver; c.device = device; c.dx = dx; c.dy = dy; c.dz = dz;
        c.symmetry = symmetry;
        c.z0 = 50.0;
        nf_ = num_freq; ne_ = num_eobj; tiles_ = tiles_x * tiles_y; max_flips_ = max_flips;
        int rc = pngf_create(&c, &h_);
        if (rc != PNGF_OK) {
            std::string m = h_ ? pngf_last_error(h_) : "invalid configuration";
            if (h_) pngf_destroy(h_);
            h_ = nullptr;
            throw Error(rc, m);
        }
    }
    ~Optimizer() { if (h_) pngf_destroy(h_); }
    Optimizer(const Optimizer&) = delete;
    Optimizer& operator=(const Optimizer&) = delete;

    void set_frequency(int f, double hz) { ck(pngf_set_frequency(h_, f, hz)); }
    // reference init's fread of Gmat_sub_XX.bin (pngfopt.hpp:121-184)
    void load_g(int f, const char* path) { ck(pngf_load_g(h_, f, path)); }
    void synth_g(int f) { ck(pngf_synth_g(h_, f)); }
    void set_tile_map(const std::vector<int>& tm) { ck(pngf_set_tile_map(h_, tm.data())); }
    void set_source_port(int tx, int ty, int yoff, int ysz) { ck(pngf_set_source_port(h_, tx, ty, yoff, ysz)); }
    // compute_C_matrix + LinAlg::invertMatrix + compute_PEinc + compute_R_matrix(1)
    void init_state() { ck(pngf_init_state(h_)); }

    // reference: compute_updated_Eobj (pngfopt.hpp:379-539); returns newEobj(0) of frequency 0
    // and keeps all newEobj values for newEobj(f, r).
    cplx compute_updated_Eobj(const int* tiles, int n) {
        last_.assign(nf_ * ne_, cplx(0.0));
        last_tiles_.assign(tiles, tiles + n);
        ck(pngf_eval(h_, tiles, n, reinterpret_cast<pngf_c128*>(last_.data()), &last_F_));
        return last_[0];
    }
    // reference: update_C_inv (+ compute_PEinc + compute_R_matrix(0)); commits the flip
    // scored by the last compute_updated_Eobj call, like the reference.
    void update_C_inv() { ck(pngf_accept(h_, last_tiles_.data(), (int)last_tiles_.size())); }

    cplx newEobj(int f, int r) const { return last_[f * ne_ + r]; }
    double last_objective() const { return last_F_; }
    cplx Eobj(int f, int r) {
        std::vector<cplx> e(nf_ * ne_);
        ck(pngf_get_state(h_, reinterpret_cast<pngf_c128*>(e.data()), nullptr, nullptr, nullptr, nullptr));
        return e[f * ne_ + r];
    }
    cplx Vinc(int f) {
        std::vector<cplx> v(nf_);
        ck(pngf_get_state(h_, nullptr, reinterpret_cast<pngf_c128*>(v.data()), nullptr, nullptr, nullptr));
        return v[f];
    }
    cplx LNinc(int f, int which) {
        std::vector<cplx> l(2 * nf_);
        ck(pngf_get_state(h_, nullptr, nullptr, reinterpret_cast<pngf_c128*>(l.data()), nullptr, nullptr));
        return l[2 * f + which];
    }
    std::vector<int> tile_map() {
        std::vector<int> t(tiles_);
        ck(pngf_get_tile_map(h_, t.data()));
        return t;
    }
    // The whole DBS loop of main.cpp:338-459 on the device.
    pngf_dbs_result dbs_run(pngf_dbs_params& p) {
        pngf_dbs_result r;
        ck(pngf_dbs_run(h_, &p, &r));
        return r;
    }
    pngf_handle* handle() { return h_; }

private:
    void ck(int rc) {
        if (rc != PNGF_OK) throw Error(rc, pngf_last_error(h_));
    }
    pngf_handle* h_ = nullptr;
    int nf_ = 0, ne_ = 0, tiles_ = 0, max_flips_ = 0;
    std::vector<cplx> last_;
    std::vector<int> last_tiles_;
    double last_F_ = 0.0;
};

}  // namespace pngf
