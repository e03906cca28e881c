// Host-side description of the design region: component indexing, metal mask,
// port components, component positions (for the synthetic G) and the candidate
// flip groups of each DBS driver.
#pragma once
#include <vector>
#include <algorithm>
#include <cmath>
#include <cstdint>

namespace pngf {

struct Layout {
    int kind = 1, Tx = 0, Ty = 0, cx = 3, cy = 3, N = 0;
    int yxx = 0, yxy = 0, yxs = 0, yyx = 0, yyy = 0, yys = 0;  // SHARED Yee block sizes
    int per_tile = 0, n_design = 0;
    int ptx = 0, pty = 0, pyoff = 1, pysz = 2;
    std::vector<int> port;

    // SHARED: pngfopt.hpp:86-94.  DISJOINT: per_tile components per tile, tile
    // after tile in raster order, then cx * pysz port components.
    void setup(int k, int tx, int ty, int ccx, int ccy) {
        kind = k; Tx = tx; Ty = ty; cx = ccx; cy = ccy;
        yxx = Tx * cx; yxy = Ty * cy + 1; yyx = Tx * cx + 1; yyy = Ty * cy;
        yxs = yxx * yxy; yys = yyx * yyy;
        per_tile = cx * (cy + 1) + (cx + 1) * cy;
        if (kind == 0) { n_design = yxs + yys; N = n_design; }
        else { n_design = per_tile * Tx * Ty; N = n_design + cx * 2; }
        set_port(Tx / 2, Ty / 2, 1, 2);
    }
    bool port_ok(int yoff, int ysz) const { return yoff >= 0 && ysz > 0 && yoff + ysz <= cy; }
    // pngfopt.hpp:584-590 for SHARED; the separate port components for DISJOINT.
    void set_port(int tx, int ty, int yoff, int ysz) {
        ptx = tx; pty = ty; pyoff = yoff; pysz = ysz;
        port.clear();
        if (kind == 0) {
            for (int jj = 0; jj < ysz; jj++)
                for (int ii = 0; ii < cx; ii++)
                    port.push_back((tx * cx + ii) + (ty * cy + jj + yoff) * yxx);
        } else {
            // DISJOINT keeps exactly cx * 2 port slots in N
            for (int p = 0; p < cx * 2 && p < cx * ysz; p++) port.push_back(n_design + p);
        }
    }
    // Components of one tile, ascending (pngfopt.hpp:640-661 for SHARED).
    void tile_components(int tile, std::vector<int>& out) const {
        int ty = tile / Tx, tx = tile - ty * Tx;
        if (kind == 0) {
            int bx = tx * cx + ty * cy * yxx, by = yxs + tx * cx + ty * cy * yyx;
            for (int j = 0; j < cy + 1; j++) for (int i = 0; i < cx; i++) out.push_back(bx + i + j * yxx);
            for (int j = 0; j < cy; j++) for (int i = 0; i < cx + 1; i++) out.push_back(by + i + j * yyx);
        } else {
            for (int m = 0; m < per_tile; m++) out.push_back(tile * per_tile + m);
        }
    }
    // Tiles containing component i (up to 2 for SHARED edges; -1 = none).
    void owners(int i, int& o0, int& o1) const {
        o0 = o1 = -1;
        if (kind == 1) { if (i < n_design) o0 = i / per_tile; return; }
        if (i < yxs) {  // Ex at (ix, line iy)
            int ix = i % yxx, iy = i / yxx, tx = ix / cx;
            int ty = iy / cy;
            if (ty < Ty) o0 = tx + ty * Tx;
            if (iy % cy == 0 && iy > 0) { int t2 = tx + (ty - 1) * Tx; if (o0 < 0) o0 = t2; else o1 = t2; }
        } else {        // Ey at (line ix, iy)
            int j = i - yxs, ix = j % yyx, iy = j / yyx, ty = iy / cy;
            int tx = ix / cx;
            if (tx < Tx) o0 = tx + ty * Tx;
            if (ix % cx == 0 && ix > 0) { int t2 = (tx - 1) + ty * Tx; if (o0 < 0) o0 = t2; else o1 = t2; }
        }
    }
    // Metal mask from a tile map (pngfopt.hpp:624-664).
    void mask(const int32_t* tm, std::vector<int8_t>& P) const {
        P.assign(N, 0);
        std::vector<int> c;
        for (int t = 0; t < Tx * Ty; t++) {
            if (!tm[t]) continue;
            c.clear(); tile_components(t, c);
            for (int i : c) P[i] = 1;
        }
    }
    // Position (units of cells) and orientation (0 = x, 1 = y) of component i.
    void position(int i, double& x, double& y, int& o) const {
        if (kind == 0) {
            if (i < yxs) { x = (i % yxx) + 0.5; y = i / yxx; o = 0; }
            else { int j = i - yxs; x = j % yyx; y = j / yyx + 0.5; o = 1; }
            return;
        }
        if (i >= n_design) {  // port: interior x-edges of the port tile
            int p = i - n_design, ii = p % cx, jj = p / cx;
            x = ptx * cx + ii + 0.5; y = pty * cy + pyoff + jj; o = 0; return;
        }
        int tile = i / per_tile, m = i - tile * per_tile, ty = tile / Tx, tx = tile - ty * Tx;
        double lx, ly;
        if (m < cx * (cy + 1)) { lx = (m % cx) + 0.5; ly = m / cx; o = 0; }
        else { int mm = m - cx * (cy + 1); lx = mm % (cx + 1); ly = mm / (cx + 1) + 0.5; o = 1; }
        // pull each tile's edges 20% toward the tile centre so neighbouring tiles'
        // components do not coincide
        lx = 0.5 * cx + 0.8 * (lx - 0.5 * cx); ly = 0.5 * cy + 0.8 * (ly - 0.5 * cy);
        x = tx * cx + lx; y = ty * cy + ly;
    }
};

// A candidate flip: up to 4 tiles, and the union of their components (ascending).
struct Groups {
    int n = 0, gmax = 0, tmax = 4;
    std::vector<int32_t> tiles;   // [n][4], -1 padded
    std::vector<int32_t> ntiles;  // [n]
    std::vector<int32_t> comps;   // [n][gmax], -1 padded
    std::vector<int32_t> ncomp;   // [n]
    std::vector<int32_t> prim;    // [n] primary tile (what the reference logs)

    void add(const Layout& L, const int* t, int nt, int primary) {
        std::vector<int> c;
        for (int i = 0; i < nt; i++) L.tile_components(t[i], c);
        std::sort(c.begin(), c.end());
        c.erase(std::unique(c.begin(), c.end()), c.end());
        for (int i = 0; i < 4; i++) tiles.push_back(i < nt ? t[i] : -1);
        ntiles.push_back(nt);
        ncomp.push_back((int)c.size());
        prim.push_back(primary);
        for (int i = 0; i < gmax; i++) comps.push_back(i < (int)c.size() ? c[i] : -1);
        n++;
    }
};

// Flip set of the reference driver for a quadrant index r (main.cpp:354-391).
// Returns the number of tiles (0 for the draws the reference skips).
inline int ref_flip_set(int r, int Tx, int Ty, int* t) {
    int qx = (Tx - 1) / 2 + 1;
    int ty = r / qx, tx = r - ty * qx;
    int midx = (Tx - 1) / 2, midy = (Ty - 1) / 2;
    if (tx == midx && ty == midy) return 0;
    if (tx == midx - 1 && ty == midy) return 0;
    if (tx == midx) { t[0] = tx + ty * Tx; t[1] = tx + ((Ty - 1) - ty) * Tx; return 2; }
    if (ty == midy) { t[0] = tx + ty * Tx; t[1] = ((Tx - 1) - tx) + ty * Tx; return 2; }
    t[0] = tx + ty * Tx; t[1] = ((Tx - 1) - tx) + ty * Tx;
    t[2] = tx + ((Ty - 1) - ty) * Tx; t[3] = ((Tx - 1) - tx) + ((Ty - 1) - ty) * Tx;
    return 4;
}

}  // namespace pngf
