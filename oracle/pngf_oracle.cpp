// CPU oracle for the pngf-opt direct-binary-search (DBS) Woodbury path.
//
// TEST INFRASTRUCTURE ONLY.  This file is what the tests compare the HIP library
// against, and the CPU baseline that bench.py times beside it.  It is
// never linked into, loaded by, or called from the product library.
//
// It restates the reference optimizer (pngf-opt/pngfopt.hpp, pngf-opt/linalg.hpp,
// pngf-opt/main.cpp, pngf-opt/magichat.hpp) in plain C++17 + OpenMP with its own
// complex LU (no Eigen, no BLAS: neither is present here).  Storage is column-major
// like the reference's Eigen matrices, and the loop structure of the two hot
// methods is kept (stride-N row gather of G, k^2 N capacitance dot products,
// naive k N^2 V*Cinv loop, rank-k update), so that its timing is a fair CPU
// baseline of the reference algorithm.
//
// Pinning (see tests/test_oracle.py):
//   * compute_updated_Eobj  -> brute-force dense re-solve Gobj (C+dC)^-1 PEinc'
//                              and numpy golden vectors (tests/golden/)
//   * update_C_inv          -> Woodbury-updated Cinv equals inv(C') (numpy)
//   * layout / P mask       -> closed-form counts (924, 8064, k in {12..24})
//   * MagicHat              -> mt19937 10000th output KAT, draw-without-replacement
//                              invariants
//   * objective             -> hand-computed Zin/S11/D known answers
//   * zgetrf/zgetrs/zgetri  -> scipy.linalg.lapack on the same matrices
//
// Built twice (oracle/Makefile): in complex128, and with PNGF_ORACLE_C64 in complex64 -- the
// reference's USE_SINGLE_PRECISION build (main.cpp:58-72: `precision` = float throughout, the
// c128 file cast to c64 at load, pngfopt.hpp:128-156; c-prefixed LAPACK/BLAS,
// pngfopt.hpp:285-347, 488-510, linalg.hpp:25-78).  The C ABI below stays in double either way:
// arrays crossing it are converted at the boundary.
#include <complex>
#include <vector>
#include <random>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <cstdint>
#include <chrono>
#include <algorithm>
#include <omp.h>

#ifdef PNGF_ORACLE_C64
typedef float real;
#else
typedef double real;
#endif
typedef std::complex<real> cd;
typedef std::complex<double> c128;

namespace {

// ---------------------------------------------------------------- layout ----
// kind 0 = shared-edge Yee layout of the reference (pngfopt.hpp:86-94, 624-664):
//          Ex block (cx*T) x (cy*T+1), then Ey block (cx*T+1) x (cy*T), x fastest;
//          the port is a set of interior Ex components of the port tile
//          (pngfopt.hpp:584-590).
// kind 1 = disjoint layout: each tile owns cx*(cy+1) + (cx+1)*cy components
//          (24 for 3x3 cells) stored contiguously tile after tile in raster order,
//          then cx*ysz separate x-directed port components that are never metal.
struct Layout {
    int kind, T, cx, cy, N;
    int yxx, yxy, yxs, yyx, yyy, yys;  // shared-layout Yee block sizes
    int per_tile;
    int ptx, pty, pyoff, pysz;
    std::vector<int> port;

    void setup(int k, int t, int ccx, int ccy) {
        kind = k; T = t; cx = ccx; cy = ccy;
        yxx = T * cx; yxy = T * cy + 1; yyx = T * cx + 1; yyy = T * cy;
        yxs = yxx * yxy; yys = yyx * yyy;
        per_tile = cx * (cy + 1) + (cx + 1) * cy;
        ptx = pty = T / 2; pyoff = 1; pysz = 2;
        if (kind == 0) N = yxs + yys;
        else N = per_tile * T * T + cx * pysz;
    }
    // reference: pngfopt.hpp:624-664 (compute_P_matrix)
    void mask(const int* tile_map, int* P) const {
        if (kind == 0) {
            for (int i = 0; i < N; i++) P[i] = 0;
            for (int idx = 0; idx < T * T; idx++) {
                if (!tile_map[idx]) continue;
                int ty = idx / T, tx = idx - ty * T;
                int bx = tx * cx + ty * cy * yxx;
                int by = yxs + tx * cx + ty * cy * yyx;
                for (int j = 0; j < cy + 1; j++)
                    for (int i = 0; i < cx; i++) P[bx + i + j * yxx] = 1;
                for (int j = 0; j < cy; j++)
                    for (int i = 0; i < cx + 1; i++) P[by + i + j * yyx] = 1;
            }
        } else {
            for (int c = 0; c < T * T; c++)
                for (int m = 0; m < per_tile; m++) P[c * per_tile + m] = tile_map[c] ? 1 : 0;
            for (int i = per_tile * T * T; i < N; i++) P[i] = 0;
        }
    }
    // reference: pngfopt.hpp:584-590 (port component indices)
    void set_port(int tx, int ty, int yoff, int ysz) {
        ptx = tx; pty = ty; pyoff = yoff; pysz = ysz;
        port.clear();
        if (kind == 0) {
            for (int jj = 0; jj < ysz; jj++)
                for (int ii = 0; ii < cx; ii++)
                    port.push_back((tx * cx + ii) + (ty * cy + jj + yoff) * yxx);
        } else {
            int base = per_tile * T * T;
            for (int p = 0; p < cx * ysz; p++) port.push_back(base + p);
        }
    }
};

// ------------------------------------------------------ dense complex LU ----
// LAPACK zgetrf semantics (unblocked right-looking form): pivot = first row of
// max |Re|+|Im| (izamax), full-row swaps, unit-lower L stored below the diagonal.
// Column-major a(i,j) = a[i + j*lda].  Returns info (0 ok, j+1 if U(j,j)==0).
int lu_factor(int n, cd* a, int lda, int* ipiv) {
    int info = 0;
    for (int j = 0; j < n; j++) {
        int p = j; real best = -1;
        for (int i = j; i < n; i++) {
            real v = std::fabs(a[i + (size_t)j * lda].real()) + std::fabs(a[i + (size_t)j * lda].imag());
            if (v > best) { best = v; p = i; }
        }
        ipiv[j] = p;
        if (a[p + (size_t)j * lda] == cd(0, 0)) { if (!info) info = j + 1; continue; }
        if (p != j)
            for (int c = 0; c < n; c++) std::swap(a[j + (size_t)c * lda], a[p + (size_t)c * lda]);
        cd r = real(1) / a[j + (size_t)j * lda];
        for (int i = j + 1; i < n; i++) a[i + (size_t)j * lda] *= r;
        for (int c = j + 1; c < n; c++) {
            cd f = a[j + (size_t)c * lda];
            if (f == cd(0, 0)) continue;
            cd* col = a + (size_t)c * lda;
            const cd* lj = a + (size_t)j * lda;
            for (int i = j + 1; i < n; i++) col[i] -= lj[i] * f;
        }
    }
    return info;
}

// zgetrs('N') with one right-hand side.
void lu_solve(int n, const cd* a, int lda, const int* ipiv, cd* b) {
    for (int j = 0; j < n; j++) if (ipiv[j] != j) std::swap(b[j], b[ipiv[j]]);
    for (int j = 0; j < n; j++) {
        cd bj = b[j];
        for (int i = j + 1; i < n; i++) b[i] -= a[i + (size_t)j * lda] * bj;
    }
    for (int j = n - 1; j >= 0; j--) {
        b[j] /= a[j + (size_t)j * lda];
        cd bj = b[j];
        for (int i = 0; i < j; i++) b[i] -= a[i + (size_t)j * lda] * bj;
    }
}

// zgetri: inverse from the LU factors, in place.  inv(A) = inv(U) inv(L) P.
void lu_inverse(int n, cd* a, int lda, const int* ipiv) {
    // inv(U) in place (upper triangle)
    for (int j = 0; j < n; j++) {
        a[j + (size_t)j * lda] = real(1) / a[j + (size_t)j * lda];
        cd ajj = -a[j + (size_t)j * lda];
        // column j above the diagonal: x = inv(U(0:j,0:j)) * U(0:j, j)
        std::vector<cd> x(j);
        for (int i = 0; i < j; i++) {
            cd s = 0;
            for (int l = i; l < j; l++) s += a[i + (size_t)l * lda] * a[l + (size_t)j * lda];
            x[i] = s;
        }
        for (int i = 0; i < j; i++) a[i + (size_t)j * lda] = x[i] * ajj;
    }
    // solve X * L = inv(U) for X, column by column from the right
    std::vector<cd> work(n);
    for (int j = n - 2; j >= 0; j--) {
        for (int i = j + 1; i < n; i++) { work[i] = a[i + (size_t)j * lda]; a[i + (size_t)j * lda] = 0; }
        for (int i = 0; i < n; i++) {
            cd s = 0;
            for (int l = j + 1; l < n; l++) s += a[i + (size_t)l * lda] * work[l];
            a[i + (size_t)j * lda] -= s;
        }
    }
    // undo the row interchanges as column interchanges, in reverse order
    for (int j = n - 2; j >= 0; j--) {
        int p = ipiv[j];
        if (p != j)
            for (int i = 0; i < n; i++) std::swap(a[i + (size_t)j * lda], a[i + (size_t)p * lda]);
    }
}

// In-place N x N inverse for the one-time initial C^-1 (reference:
// linalg.hpp:34-85, zgetrf + zgetri).  Blocked by columns of the identity so
// that OpenMP can split the N right-hand sides; same pivot rule as lu_factor.
int invert_dense(int n, cd* a) {
    std::vector<int> ipiv(n);
    // right-looking LU with the trailing update parallel over columns
    int info = 0;
    for (int j = 0; j < n; j++) {
        int p = j; real best = -1;
        for (int i = j; i < n; i++) {
            real v = std::fabs(a[i + (size_t)j * n].real()) + std::fabs(a[i + (size_t)j * n].imag());
            if (v > best) { best = v; p = i; }
        }
        ipiv[j] = p;
        if (a[p + (size_t)j * n] == cd(0, 0)) { info = j + 1; break; }
        if (p != j) {
            #pragma omp parallel for
            for (int c = 0; c < n; c++) std::swap(a[j + (size_t)c * n], a[p + (size_t)c * n]);
        }
        cd r = real(1) / a[j + (size_t)j * n];
        for (int i = j + 1; i < n; i++) a[i + (size_t)j * n] *= r;
        #pragma omp parallel for schedule(static)
        for (int c = j + 1; c < n; c++) {
            cd f = a[j + (size_t)c * n];
            cd* col = a + (size_t)c * n;
            const cd* lj = a + (size_t)j * n;
            for (int i = j + 1; i < n; i++) col[i] -= lj[i] * f;
        }
    }
    if (info) return info;
    // inverse by solving A X = I column by column (parallel over columns)
    std::vector<cd> inv((size_t)n * n);
    #pragma omp parallel for schedule(dynamic, 8)
    for (int c = 0; c < n; c++) {
        cd* b = inv.data() + (size_t)c * n;
        for (int i = 0; i < n; i++) b[i] = 0;
        b[c] = 1;
        lu_solve(n, a, n, ipiv.data(), b);
    }
    std::memcpy(a, inv.data(), sizeof(cd) * (size_t)n * n);
    return 0;
}

// ------------------------------------------------------------ optimizer ----
// One object per frequency, as in pngf-opt/main.cpp:166,224.
struct Opt {
    Layout L;
    int N, ne, max_k;
    std::vector<cd> G, C, Gobj, R, Einc, PEinc, S, Snew;   // column-major
    std::vector<cd> V, CinvU, VCinv, VinvCU, RU, VSnew;
    std::vector<int> tile_map, yee_grid, diff_P, P_idx, ipiv;
    cd Vinc = 0, LN1inc = 0, LN2inc = 0;
    std::vector<cd> Eobj, newEobj;
    int nk = 0;                   // num_yee_flipped
    int last_flipped[16];
    int num_flipped_tiles = 0;
    int last_lu_info = 0;

    // reference: pngfopt.hpp:66-191 (init).  `file` is the acmefdfd raw buffer:
    // M x M column-major complex128 with M = N + ne - 1; G = file[0:N,0:N],
    // Gobj(1+r, c) = file[N+r, c] (cast to complex64 in the c64 build, pngfopt.hpp:128-156).
    void init(int kind, int T, int cx, int cy, int tne, int max_flips, const c128* file) {
        L.setup(kind, T, cx, cy);
        N = L.N; ne = tne;
        max_k = L.per_tile * max_flips;
        size_t NN = (size_t)N * N;
        G.assign(NN, 0); C.assign(NN, 0);
        Gobj.assign((size_t)ne * N, 0); R.assign((size_t)ne * N, 0);
        Einc.assign(N, 0); PEinc.assign(N, 0); S.assign(N, 0); Snew.assign(N, 0);
        V.assign((size_t)max_k * N, 0); CinvU.assign((size_t)max_k * N, 0);
        VCinv.assign((size_t)max_k * N, 0); VinvCU.assign((size_t)max_k * max_k, 0);
        RU.assign((size_t)ne * N, 0); VSnew.assign(N, 0);
        tile_map.assign(T * T, 0); yee_grid.assign(N, 0); diff_P.assign(N, 0);
        P_idx.assign(N, 0); ipiv.assign(N, 0);
        Eobj.assign(ne, 0); newEobj.assign(ne, 0);
        size_t M = (size_t)N + ne - 1;
        #pragma omp parallel for
        for (long long ii = 0; ii < (long long)NN; ii++) {
            size_t col = ii / N, row = ii - col * N;
            G[ii] = (cd)file[row + M * col];
        }
        for (int r = 0; r < ne - 1; r++)
            for (int c = 0; c < N; c++) Gobj[(1 + r) + (size_t)ne * c] = (cd)file[(N + r) + M * c];
    }
    // reference: pngfopt.hpp:669-674
    void set_tile_map(const int* tm) {
        for (int i = 0; i < L.T * L.T; i++) tile_map[i] = tm[i];
        L.mask(tile_map.data(), yee_grid.data());
    }
    // reference: pngfopt.hpp:559-618
    int set_source_port(int tx, int ty, int yoff, int ysz) {
        if (yoff + ysz > L.cy) return -1;
        L.set_port(tx, ty, yoff, ysz);
        for (int c = 0; c < N; c++) Gobj[0 + (size_t)ne * c] = 0;
        std::fill(Einc.begin(), Einc.end(), cd(0));
        LN1inc = LN2inc = 0;
        for (int p : L.port) {
            for (int c = 0; c < N; c++) Gobj[0 + (size_t)ne * c] += G[p + (size_t)N * c];
            for (int r = 0; r < N; r++) Einc[r] += G[r + (size_t)N * p];
            if (ne > 1) { LN1inc += Gobj[1 + (size_t)ne * p]; LN2inc += Gobj[2 + (size_t)ne * p]; }
        }
        Vinc = 0;
        for (int p : L.port) Vinc += Einc[p];
        return 0;
    }
    // reference: pngfopt.hpp:196-233 (column-major branch); C = I - P + P G
    void compute_C() {
        #pragma omp parallel for
        for (int col = 0; col < N; col++) {
            size_t base = (size_t)col * N;
            for (int row = 0; row < N; row++)
                C[base + row] = yee_grid[row] == 1 ? G[base + row] : (row == col ? cd(1) : cd(0));
        }
    }
    int invert() { return invert_dense(N, C.data()); }
    // reference: pngfopt.hpp:542-552
    void compute_PEinc() {
        for (int i = 0; i < N; i++) PEinc[i] = yee_grid[i] ? -Einc[i] : cd(0);
    }
    // reference: pngfopt.hpp:237-250;  R = Gobj*Cinv, S = Cinv*PEinc
    void compute_R(int compute_eobj) {
        #pragma omp parallel for
        for (int c = 0; c < N; c++) {
            for (int r = 0; r < ne; r++) {
                cd s = 0;
                const cd* cc = C.data() + (size_t)c * N;
                for (int i = 0; i < N; i++) s += Gobj[r + (size_t)ne * i] * cc[i];
                R[r + (size_t)ne * c] = s;
            }
        }
        std::vector<cd> s(N, cd(0));
        for (int c = 0; c < N; c++) {
            cd pe = PEinc[c];
            if (pe == cd(0)) continue;
            const cd* cc = C.data() + (size_t)c * N;
            for (int i = 0; i < N; i++) s[i] += cc[i] * pe;
        }
        S = s;
        if (compute_eobj) {
            for (int r = 0; r < ne; r++) {
                cd e = 0;
                for (int c = 0; c < N; c++) e += R[r + (size_t)ne * c] * PEinc[c];
                Eobj[r] = e;
            }
            newEobj = Eobj;
        } else {
            Eobj = newEobj;
        }
    }

    // reference: pngfopt.hpp:379-539 (compute_updated_Eobj), formulation and loop
    // structure kept: dP scan, stride-N gather of V = (G-I)[t,:], k^2 N dot
    // products for the capacitance, Snew, V*Snew, LU + solve, Gobj*Snew - RU*y.
    cd compute_updated_Eobj(const int* tiles, int ntiles) {
        num_flipped_tiles = ntiles;
        for (int i = 0; i < ntiles; i++) { last_flipped[i] = tiles[i]; tile_map[tiles[i]] ^= 1; }
        L.mask(tile_map.data(), diff_P.data());
        for (int i = 0; i < ntiles; i++) tile_map[tiles[i]] ^= 1;
        for (int i = 0; i < N; i++) diff_P[i] -= yee_grid[i];
        nk = 0;
        for (int i = 0; i < N; i++) {
            if (diff_P[i] != 0) {
                P_idx[nk] = i;
                real sg = diff_P[i] > 0 ? 1 : -1;
                for (int r = 0; r < ne; r++) RU[(size_t)nk * ne + r] = sg * R[r + (size_t)ne * i];
                nk++;
            }
        }
        const int k = nk;
        #pragma omp parallel for
        for (int jj = 0; jj < k; jj++) {
            size_t base = (size_t)jj * N;
            int t = P_idx[jj];
            for (int ii = 0; ii < N; ii++) V[base + ii] = G[t + (size_t)ii * N];
            V[base + t] -= real(1);
        }
        #pragma omp parallel for
        for (int ii = 0; ii < k * k; ii++) {
            int col = ii / k, row = ii - col * k;
            const cd* v = V.data() + (size_t)row * N;
            const cd* cc = C.data() + (size_t)P_idx[col] * N;
            cd s = 0;
            for (int i = 0; i < N; i++) s += v[i] * cc[i];
            if (diff_P[P_idx[col]] < 0) s = -s;
            if (row == col) s += real(1);
            VinvCU[ii] = s;
        }
        #pragma omp parallel for
        for (int ii = 0; ii < N; ii++) {
            cd s = 0;
            for (int jj = 0; jj < k; jj++)
                s += C[ii + (size_t)P_idx[jj] * N] * Einc[P_idx[jj]] * (real)diff_P[P_idx[jj]];
            Snew[ii] = S[ii] - s;
        }
        #pragma omp parallel for
        for (int ii = 0; ii < k; ii++) {
            const cd* v = V.data() + (size_t)ii * N;
            cd s = 0;
            for (int jj = 0; jj < N; jj++) s += v[jj] * Snew[jj];
            VSnew[ii] = s;
        }
        last_lu_info = lu_factor(k, VinvCU.data(), k, ipiv.data());
        lu_solve(k, VinvCU.data(), k, ipiv.data(), VSnew.data());
        for (int r = 0; r < ne; r++) {
            cd s = 0;
            for (int i = 0; i < N; i++) s += Gobj[r + (size_t)ne * i] * Snew[i];
            newEobj[r] = s;
        }
        for (int r = 0; r < ne; r++) {
            cd s = 0;
            for (int i = 0; i < k; i++) s += RU[(size_t)i * ne + r] * VSnew[i];
            newEobj[r] -= s;
        }
        return newEobj[0];
    }

    // reference: pngfopt.hpp:256-367 (update_C_inv); valid only right after the
    // compute_updated_Eobj call that scored the flip being committed.
    void update_C_inv() {
        const int k = nk;
        #pragma omp parallel for
        for (int ii = 0; ii < k; ii++) {
            size_t base = (size_t)ii * N, cb = (size_t)P_idx[ii] * N;
            real sg = diff_P[P_idx[ii]] > 0 ? 1 : -1;
            for (int jj = 0; jj < N; jj++) CinvU[base + jj] = sg * C[cb + jj];
        }
        lu_inverse(k, VinvCU.data(), k, ipiv.data());
        #pragma omp parallel for
        for (long long ii = 0; ii < (long long)k * N; ii++) {
            int col = (int)(ii / k), row = (int)(ii - (long long)col * k);
            const cd* v = V.data() + (size_t)row * N;
            const cd* cc = C.data() + (size_t)col * N;
            cd s = 0;
            for (int i = 0; i < N; i++) s += v[i] * cc[i];
            VCinv[ii] = s;
        }
        // V <- cap^-1 * VCinv   (k x N, column-major with ld k)
        #pragma omp parallel for
        for (int col = 0; col < N; col++) {
            cd tmp[128];
            for (int r = 0; r < k; r++) {
                cd s = 0;
                for (int l = 0; l < k; l++) s += VinvCU[r + (size_t)l * k] * VCinv[l + (size_t)col * k];
                tmp[r] = s;
            }
            for (int r = 0; r < k; r++) V[r + (size_t)col * k] = tmp[r];
        }
        // C <- C - CinvU * V  (N x N rank-k update)
        #pragma omp parallel for schedule(static)
        for (int col = 0; col < N; col++) {
            cd* cc = C.data() + (size_t)col * N;
            for (int l = 0; l < k; l++) {
                cd f = V[l + (size_t)col * k];
                const cd* u = CinvU.data() + (size_t)l * N;
                for (int i = 0; i < N; i++) cc[i] -= u[i] * f;
            }
        }
        for (int i = 0; i < num_flipped_tiles; i++) tile_map[last_flipped[i]] ^= 1;
        for (int i = 0; i < N; i++) yee_grid[i] += diff_P[i];
    }
};

// ------------------------------------------------------------ objective ----
const double kC0 = 2.997924579998211E8;
const double kEta0 = 3.767303134118051e+02;

// The objective in the arithmetic type T of the build (the reference computes it in `precision`,
// main.cpp:265-300, 401-423); T = double is the complex128 build's.
template <class T>
struct Objective {
    typedef std::complex<T> C;
    T dx, dy, dz;
    // reference: main.cpp:277-289 / 407-413
    C zin(C e0, C vinc) const { return -(e0 + vinc) * dx * T(0.5) / (T(2.0) * dy * dz); }
    static C s11(C zin) { return (zin - T(50.0)) / (zin + T(50.0)); }
    T dir(const C* e, C vinc, C ln1, C ln2, T f) const {
        T k0 = T(2.0 * M_PI) * f / T(kC0);
        C a = e[1] + ln1, b = e[2] + ln2;
        T d = (k0 * k0 / T(32 * M_PI * M_PI * kEta0)) * std::real(std::conj(a) * a + std::conj(b) * b);
        return T(4.0 * M_PI) * d / std::real(-(e[0] + vinc) * dx * T(0.5) * (T(2.0) * dy * dz) * T(0.5));
    }
    // F of main.cpp:407-420 (mode 0) or mean_f |S11_f|^2 (mode 1); S11 and D per frequency if asked
    T operator()(int mode, int nf, const C* E, int ne, const C* Vi, const C* LN, const double* freqs, double* s11_out,
                 double* d_out) const {
        T acc = 0;
        for (int f = 0; f < nf; f++) {
            C s = s11(zin(E[f * ne], Vi[f]));
            T m = std::abs(s);
            if (s11_out) { s11_out[2 * f] = s.real(); s11_out[2 * f + 1] = s.imag(); }
            if (mode == 0) {
                T D = dir(E + f * ne, Vi[f], LN[2 * f], LN[2 * f + 1], T(freqs[f]));
                if (d_out) d_out[f] = D;
                acc += T(200.0) * m * m + (T(15.0) - D) * (T(15.0) - D);
            } else {
                acc += m * m;
            }
        }
        return mode == 0 ? acc : acc / T(nf);
    }
};

// double arrays across the C ABI <-> the build's complex type
std::vector<cd> from_abi(const double* p, size_t n) {
    std::vector<cd> v(n);
    for (size_t i = 0; i < n; i++) v[i] = cd((real)p[2 * i], (real)p[2 * i + 1]);
    return v;
}
void to_abi(const cd* v, size_t n, double* p) {
    for (size_t i = 0; i < n; i++) { p[2 * i] = v[i].real(); p[2 * i + 1] = v[i].imag(); }
}

// ---------------------------------------------------------------- MagicHat --
// reference: magichat.hpp:53-107 (seed != 0 branch only: seed 0 reads
// std::random_device and is not reproducible).
struct Hat {
    std::mt19937 gen;
    std::uniform_int_distribution<> dis;
    int total, left;
    std::vector<int> idx;
    void init(int seed, int n) {
        gen.seed(seed);
        dis = std::uniform_int_distribution<>(0, n - 1);
        total = left = n;
        idx.resize(n);
        for (int i = 0; i < n; i++) idx[i] = i;
    }
    int rand_binary() { return gen() & 1; }
    int draw() {
        if (left <= 0) return -1;
        dis.param(std::uniform_int_distribution<int>::param_type(0, left - 1));
        int i = dis(gen);
        int v = idx[i];
        idx[i] = idx[left - 1];
        left--;
        return v;
    }
    void reset() { for (int i = 0; i < total; i++) idx[i] = i; left = total; }
};

}  // namespace

// ======================================================================= C ABI
extern "C" {

// ---- layout
int or_layout_size(int kind, int T, int cx, int cy) { Layout L; L.setup(kind, T, cx, cy); return L.N; }
void or_layout_mask(int kind, int T, int cx, int cy, const int* tile_map, int* P) {
    Layout L; L.setup(kind, T, cx, cy); L.mask(tile_map, P);
}
int or_layout_port(int kind, int T, int cx, int cy, int tx, int ty, int yoff, int ysz, int* out) {
    Layout L; L.setup(kind, T, cx, cy); L.set_port(tx, ty, yoff, ysz);
    for (size_t i = 0; i < L.port.size(); i++) out[i] = L.port[i];
    return (int)L.port.size();
}

// ---- dense LAPACK-equivalent helpers (for pinning against scipy)
int or_zgetrf(int n, double* a, int* ipiv) {
    std::vector<cd> m = from_abi(a, (size_t)n * n);
    int info = lu_factor(n, m.data(), n, ipiv);
    to_abi(m.data(), m.size(), a);
    return info;
}
void or_zgetrs(int n, const double* a, const int* ipiv, double* b) {
    std::vector<cd> m = from_abi(a, (size_t)n * n), x = from_abi(b, n);
    lu_solve(n, m.data(), n, ipiv, x.data());
    to_abi(x.data(), n, b);
}
void or_zgetri(int n, double* a, const int* ipiv) {
    std::vector<cd> m = from_abi(a, (size_t)n * n);
    lu_inverse(n, m.data(), n, ipiv);
    to_abi(m.data(), m.size(), a);
}
int or_invert(int n, double* a) {
    std::vector<cd> m = from_abi(a, (size_t)n * n);
    int info = invert_dense(n, m.data());
    to_abi(m.data(), m.size(), a);
    return info;
}
// 1 in the complex64 build
int or_is_c64() { return sizeof(real) == sizeof(float) ? 1 : 0; }

// ---- optimizer object (mirrors PNGF_Optimizer's methods)
void* or_opt_create(int kind, int T, int cx, int cy, int ne, int max_flips, const double* file) {
    Opt* o = new Opt();
    o->init(kind, T, cx, cy, ne, max_flips, (const c128*)file);
    return o;
}
void or_opt_destroy(void* h) { delete (Opt*)h; }
int or_opt_N(void* h) { return ((Opt*)h)->N; }
void or_opt_set_tile_map(void* h, const int* tm) { ((Opt*)h)->set_tile_map(tm); }
int or_opt_set_source_port(void* h, int tx, int ty, int yoff, int ysz) { return ((Opt*)h)->set_source_port(tx, ty, yoff, ysz); }
void or_opt_compute_C(void* h) { ((Opt*)h)->compute_C(); }
int or_opt_invert(void* h) { return ((Opt*)h)->invert(); }
void or_opt_set_cinv(void* h, const double* cinv) {  // column-major N x N
    Opt* o = (Opt*)h; o->C = from_abi(cinv, (size_t)o->N * o->N);
}
void or_opt_get_cinv(void* h, double* out) {
    Opt* o = (Opt*)h; to_abi(o->C.data(), (size_t)o->N * o->N, out);
}
void or_opt_compute_PEinc(void* h) { ((Opt*)h)->compute_PEinc(); }
void or_opt_compute_R(void* h, int flag) { ((Opt*)h)->compute_R(flag); }
void or_opt_eval(void* h, const int* tiles, int n, double* new_eobj) {
    Opt* o = (Opt*)h; o->compute_updated_Eobj(tiles, n);
    to_abi(o->newEobj.data(), o->ne, new_eobj);
}
int or_opt_last_k(void* h) { return ((Opt*)h)->nk; }
int or_opt_last_info(void* h) { return ((Opt*)h)->last_lu_info; }
void or_opt_update(void* h) { ((Opt*)h)->update_C_inv(); }
// out: Eobj[ne], Vinc, LN1inc, LN2inc (complex)
void or_opt_scalars(void* h, double* out) {
    Opt* o = (Opt*)h;
    std::vector<cd> c(o->Eobj);
    c.push_back(o->Vinc); c.push_back(o->LN1inc); c.push_back(o->LN2inc);
    to_abi(c.data(), c.size(), out);
}
void or_opt_get_vec(void* h, int which, double* out) {  // 0 S, 1 Einc, 2 PEinc, 3 R (ne x N col-major)
    Opt* o = (Opt*)h;
    const std::vector<cd>* v = which == 0 ? &o->S : which == 1 ? &o->Einc : which == 2 ? &o->PEinc : &o->R;
    to_abi(v->data(), v->size(), out);
}
void or_opt_get_tile_map(void* h, int* out) { Opt* o = (Opt*)h; for (int i = 0; i < o->L.T * o->L.T; i++) out[i] = o->tile_map[i]; }

// ---- objective known answers (reference: main.cpp:277-289, 407-413)
void or_objective(int mode, int nf, const double* eobj, int ne, const double* vinc, const double* ln,
                  const double* freqs, double dx, double dy, double dz, double* F, double* s11_out, double* d_out) {
    // always in double: the objective's known answers
    Objective<double> obj{dx, dy, dz};
    *F = obj(mode, nf, (const c128*)eobj, ne, (const c128*)vinc, (const c128*)ln, freqs, s11_out, d_out);
}

// ---- MagicHat
void* or_hat_create(int seed, int n) { Hat* h = new Hat(); h->init(seed, n); return h; }
void or_hat_destroy(void* h) { delete (Hat*)h; }
int or_hat_draw(void* h) { return ((Hat*)h)->draw(); }
void or_hat_reset(void* h) { ((Hat*)h)->reset(); }
int or_hat_rand_binary(void* h) { return ((Hat*)h)->rand_binary(); }
unsigned or_mt19937_nth(unsigned seed, int n) {  // n-th output (1-based) of mt19937(seed)
    std::mt19937 g(seed); unsigned v = 0; for (int i = 0; i < n; i++) v = g(); return v;
}

// ---- DBS driver ------------------------------------------------------------
// mode 1 (S11 objective, raster order): candidates = every tile in raster order
//   (tiles whose flip would touch a port component are skipped), single-tile
//   flips, cyclic passes, stop after ncand consecutive non-improving evaluations
//   or after max_evals evaluations (0 = unlimited).  Objective mean_f |S11_f|^2.
// mode 0 (reference driver, main.cpp:338-459): MagicHat order over the lower-left
//   quadrant with the reference's skip rules, mirror-symmetric 2/4-tile flips,
//   objective sum_f 200|S11|^2 + (15-D)^2, stop when the hat is empty.
// All frequencies share the tile map.  Outputs (caller-sized):
//   acc_tiles[]  primary tile index of every accepted flip, in order
//   acc_eval[]   flip_count at which it was accepted
//   acc_F[]      objective after each accept
//   F_trace[]    objective of every evaluation (may be null)
//   hist_*       the Logger's history (logger.hpp:64-73; main.cpp:304-328, 401-426), may be
//                null: record 0 is the start state (tile -1, decision -1), then one record per
//                evaluation: primary tile, F < F_best, perf = F, F_best (before the
//                evaluation), then per frequency Re S11, Im S11, D (0 in mode 1); at most
//                hist_cap records, hist_count[0] receives the number written
//   stats[0..5]  flip_count, accepts, F_start, F_final, loop seconds, eval seconds
// max_seconds > 0 stops the loop before the first evaluation that would start
// after that much wall time (a bounded sample for timing).
// Returns 0, or -1 if a buffer would overflow.
int or_dbs_run(void** opts, int nf, int mode, const double* freqs, double dx, double dy, double dz,
               int seed, int max_evals, int max_acc,
               int* acc_tiles, int* acc_eval, double* acc_F, double* F_trace, int trace_cap,
               int* final_map, double* stats, double max_seconds,
               int* hist_tile, int* hist_dec, double* hist_perf, int hist_cap, int* hist_count) {
    Opt** O = (Opt**)opts;
    const Objective<real> obj{(real)dx, (real)dy, (real)dz};
    const int T = O[0]->L.T, ne = O[0]->ne;
    std::vector<cd> E((size_t)nf * ne), Vi(nf), LN(2 * nf);
    auto objective = [&](bool use_new) {
        for (int f = 0; f < nf; f++) {
            for (int r = 0; r < ne; r++) E[(size_t)f * ne + r] = use_new ? O[f]->newEobj[r] : O[f]->Eobj[r];
            Vi[f] = O[f]->Vinc; LN[2 * f] = O[f]->LN1inc; LN[2 * f + 1] = O[f]->LN2inc;
        }
        return (double)obj(mode, nf, E.data(), ne, Vi.data(), LN.data(), freqs, nullptr, nullptr);
    };
    const int nperf = 2 + 3 * nf;
    std::vector<double> s11(2 * nf), dir(nf, 0.0);
    int nhist = 0;
    // one history record: objective and per-frequency S11 / D of the current (use_new = false)
    // or the candidate's (use_new = true) Eobj
    auto hist = [&](int tile, int dec, bool use_new, double F_best_now) {
        if (!hist_perf || nhist >= hist_cap) return;
        for (int f = 0; f < nf; f++) {
            for (int r = 0; r < ne; r++) E[(size_t)f * ne + r] = use_new ? O[f]->newEobj[r] : O[f]->Eobj[r];
            Vi[f] = O[f]->Vinc; LN[2 * f] = O[f]->LN1inc; LN[2 * f + 1] = O[f]->LN2inc;
        }
        const double F = obj(mode, nf, E.data(), ne, Vi.data(), LN.data(), freqs, s11.data(), dir.data());
        double* pr = hist_perf + (size_t)nhist * nperf;
        pr[0] = use_new ? F : F_best_now;
        pr[1] = F_best_now;
        for (int f = 0; f < nf; f++) {
            pr[2 + 3 * f] = s11[2 * f];
            pr[2 + 3 * f + 1] = s11[2 * f + 1];
            pr[2 + 3 * f + 2] = mode == 0 ? dir[f] : 0.0;
        }
        hist_tile[nhist] = tile;
        hist_dec[nhist] = dec;
        nhist++;
    };
    double F_best = objective(false);
    stats[2] = F_best;
    hist(-1, -1, false, F_best);
    int flip_count = 0, nacc = 0;
    double eval_s = 0.0;
    auto t0 = std::chrono::high_resolution_clock::now();

    // candidate list for mode 1
    std::vector<int> cands;
    if (mode == 1) {
        std::vector<int> tm(T * T, 0), P(O[0]->N);
        for (int c = 0; c < T * T; c++) {
            std::fill(tm.begin(), tm.end(), 0); tm[c] = 1;
            O[0]->L.mask(tm.data(), P.data());
            bool touches = false;
            for (int p : O[0]->L.port) if (P[p]) touches = true;
            if (!touches) cands.push_back(c);
        }
    }
    const int mid = (T - 1) / 2;
    Hat hat;
    if (mode == 0) hat.init(seed, ((T - 1) / 2 + 1) * ((T - 1) / 2 + 1) - 1);
    int last_tried = -1, nonimp = 0, ci = 0;
    int rc = 0;
    while (true) {
        int tiles[4], nt = 0, prim = -1;
        if (mode == 1) {
            if (nonimp >= (int)cands.size() || (max_evals > 0 && flip_count >= max_evals)) break;
            prim = cands[ci]; tiles[0] = prim; nt = 1;
            ci = (ci + 1) % (int)cands.size();
        } else {
            bool done = false;
            while (true) {
                int r = hat.draw();
                if (r < 0) { done = true; break; }
                int ty = r / ((T - 1) / 2 + 1), tx = r - ty * ((T - 1) / 2 + 1);
                if (r == last_tried) continue;
                if (tx == mid && ty == mid) continue;
                if (tx == mid - 1 && ty == mid) continue;
                if (tx == mid) { nt = 2; tiles[0] = tx + ty * T; tiles[1] = tx + ((T - 1) - ty) * T; }
                else if (ty == mid) { nt = 2; tiles[0] = tx + ty * T; tiles[1] = ((T - 1) - tx) + ty * T; }
                else {
                    nt = 4; tiles[0] = tx + ty * T; tiles[1] = ((T - 1) - tx) + ty * T;
                    tiles[2] = tx + ((T - 1) - ty) * T; tiles[3] = ((T - 1) - tx) + ((T - 1) - ty) * T;
                }
                last_tried = r; prim = tx + ty * T;
                break;
            }
            if (done || (max_evals > 0 && flip_count >= max_evals)) break;
        }
        auto e0 = std::chrono::high_resolution_clock::now();
        if (max_seconds > 0 && std::chrono::duration<double>(e0 - t0).count() > max_seconds) break;
        for (int f = 0; f < nf; f++) O[f]->compute_updated_Eobj(tiles, nt);
        double F = objective(true);
        eval_s += std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - e0).count();
        if (F_trace && flip_count < trace_cap) F_trace[flip_count] = F;
        hist(prim, F < F_best ? 1 : 0, true, F_best);
        if (F < F_best) {
            if (nacc >= max_acc) { rc = -1; break; }
            acc_tiles[nacc] = prim; acc_eval[nacc] = flip_count; acc_F[nacc] = F;
            nacc++;
            F_best = F;
            for (int f = 0; f < nf; f++) { O[f]->update_C_inv(); O[f]->compute_PEinc(); O[f]->compute_R(0); }
            if (mode == 0) hat.reset();
            nonimp = 0;
        } else {
            nonimp++;
        }
        flip_count++;
    }
    double loop_s = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
    for (int i = 0; i < T * T; i++) final_map[i] = O[0]->tile_map[i];
    stats[0] = flip_count; stats[1] = nacc; stats[3] = F_best; stats[4] = loop_s; stats[5] = eval_s;
    if (hist_count) hist_count[0] = nhist;
    return rc;
}

int or_num_threads() { return omp_get_max_threads(); }

}  // extern "C"
