// Launch wrappers for the gfx950 kernels in kernels.hip (host-callable).
#pragma once
#include "common.h"

namespace pngf {

// Per-frequency pointer tables live in device memory (arrays of nf pointers).  Z is the complex
// type of the state: zd (complex128, the default) or zf (complex64 precision mode, LOWRANK).
template <class Z>
struct DevTabsT {
    Z** W0;      // LOWRANK: N x ldN row-major
    Z** Aall;    // LOWRANK: N x Kmax row-major
    Z** Mall;    // LOWRANK: Kmax x ldN row-major (null when the factors are kept as Lam)
    Z** Grow;    // PLAIN: G row-major
    Z** Cinv;    // PLAIN: C^-1 column-major
    // LOWRANK, symmetric W, groups of <= kLdsGmax components: M is not stored.  Row K' of M
    // comes from the accept that appended A column K' (columns fstart[K'] .. + flen[K'] - 1):
    //   M[K'][n] = sum_l Lam[K'][l] A[n][fstart[K'] + l],  Lam[K'] = s_c cap^-1[c][:]
    // (c = K' - fstart[K']), which is M_new = diag(s) cap^-1 W[t,:] with W[t,:] = A_new^T.
    Z** Lam;           // [Kmax][gmax] per frequency
    int32_t* fstart;   // [Kmax]
    int32_t* flen;     // [Kmax]
};
typedef DevTabsT<zd> DevTabs;

// Contiguous per-frequency small arrays.  Eobj (ne scalars per frequency, accumulated over every
// accept) and the objective stay complex128 in both precision modes.
template <class Z>
struct DevVecsT {
    Z* q;        // [nf][ldN]    LOWRANK: W PEinc
    Z* S;        // [nf][ldN]    PLAIN:   C^-1 PEinc
    Z* R;        // [nf][ne][ldN]
    Z* Gobj;     // [nf][ne][ldN]
    Z* Einc;     // [nf][ldN]
    zd* Eobj;    // [nf][ne]
    Z* Dblk;     // [nf][ngroups][gmax][gmax]  LOWRANK blocks of W over group components
    FreqConst* fc;  // [nf]
};
typedef DevVecsT<zd> DevVecs;

// The same device memory viewed with another complex type (a handle in complex64 mode allocates
// its state at 8 bytes per complex and keeps the pointers in its zd-typed tables).
template <class Z>
inline DevTabsT<Z> tabs_as(const DevTabs& t) {
    return DevTabsT<Z>{(Z**)t.W0, (Z**)t.Aall, (Z**)t.Mall, (Z**)t.Grow, (Z**)t.Cinv, (Z**)t.Lam, t.fstart, t.flen};
}
template <class Z>
inline DevVecsT<Z> vecs_as(const DevVecs& v) {
    return DevVecsT<Z>{(Z*)v.q, (Z*)v.S, (Z*)v.R, (Z*)v.Gobj, (Z*)v.Einc, v.Eobj, (Z*)v.Dblk, v.fc};
}

struct Dims {
    int N, ldN, nf, ne, ngroups, gmax, Kmax;
    // LOWRANK with symmetric W: W0 and every cap^-1-derived factor are kept exactly symmetric
    // and the diagonal blocks are read as (D + D^T) / 2
    int sym;
    // factors kept as A and Lam (no M; symmetric W, groups of <= kLdsGmax components)
    int lam;
};

// gmax above which the k x (k+1) capacitance matrix moves from LDS to global scratch
constexpr int kLdsGmax = 32;

// Device-resident accept record for the device-driven DBS loop: when a kernel gets a non-null
// `dyn`, it reads k = dyn[0], K = dyn[1] and the group's tile count dyn[3] from device memory
// and does nothing unless dyn[2] (an accept is pending) is set; batch-sized kernels likewise
// read the batch size from `nb_dev`.  Null pointers mean the by-value arguments apply.
enum { kDynK = 0, kDynKK = 1, kDynActive = 2, kDynNt = 3 };

// State of the device-driven raster DBS loop (first improvement), in device memory.
struct LoopState {
    int64_t flip_count, spec, batches, max_evals;
    double F_best, F_sel;
    int32_t pos, nonimp, bs, nb, sel, done, nacc, ncl, bmin, bmax, acc_cap, status, g;
    int32_t pad;
    int32_t dyn[4];
};
enum { kLoopOk = 0, kLoopFull = 1, kLoopSingular = 2 };

// Loop control of the device-driven loop (loop_control_kernel, after the scoring kernel):
// every candidate's objective, the first improving one, the accept record and the next batch.
struct LoopCtl {
    LoopState* ls;
    ObjParams op;
    const FreqConst* fc;
    double* F;              // [ncl] objective per candidate of the batch
    const int32_t* cl;      // candidate groups in visit order
    int32_t *t, *s, *lp, *gtiles;  // accept record: change set and tiles of the selected group
    int32_t* acc_group;
    int64_t* acc_eval;
    double* acc_F;
};

// First batch of the device-driven loop (later batches come from the loop control kernel).
void launch_loop_batch(hipStream_t st, LoopState* ls, const int32_t* cl, int32_t* batch);
void launch_loop_control(hipStream_t st, const Dims& d, GroupTab gt, const int32_t* tile, const int8_t* P,
                         const int32_t* owners, int32_t* batch, const zd* eobj, const int32_t* info,
                         const LoopCtl& c, int32_t* iterlog = nullptr);

// ---- eval
// batches of at least this many (candidate, frequency) pairs are scored one pair per wave
constexpr int kWaveEvalMin = 2048;
template <class Z>
void launch_lr_eval(hipStream_t st, const Dims& d, const DevVecsT<Z>& v, GroupTab gt, const int32_t* tile,
                    const int8_t* P, const int32_t* owners, const int32_t* batch, int nb, zd* out_eobj,
                    int32_t* out_info, Z* gscratch, const int32_t* nb_dev = nullptr);
void launch_plain_eval(hipStream_t st, const Dims& d, const DevTabs& t, const DevVecs& v, GroupTab gt,
                       const int32_t* tile, const int8_t* P, const int32_t* owners, const int32_t* batch,
                       int nb, int nsplit, zd* partial, zd* out_eobj, int32_t* out_info,
                       zd* store /* null or [nf][acc_stride] */, int acc_stride, zd* gscratch);
void launch_objective(hipStream_t st, const ObjParams& op, const FreqConst* fc, const zd* eobj, int nb,
                      double* F_out, const double* F_best, int32_t* sel, double* F_sel,
                      const int32_t* nb_dev = nullptr);

// ---- LOWRANK accept (all frequencies in grid.y)
// acc buffer per frequency (zd): capinv [0,k^2), scratch [k^2,2k^2), y [2k^2,+k),
// w = dPE - s o y [2k^2+k,+k), newEobj [2k^2+2k,+4), Rt = R[:,t] [2k^2+2k+4,+ne*k)
template <class Z>
void launch_lr_accept_prep(hipStream_t st, const Dims& d, const DevVecsT<Z>& v, int g, const int32_t* t,
                           const int32_t* s, const int32_t* lp, int k, Z* acc, int acc_stride,
                           int32_t* info, Z* gscratch, const int32_t* dyn = nullptr,
                           const int32_t* g_dev = nullptr);
template <class Z>
void launch_lr_gather(hipStream_t st, const Dims& d, const DevTabsT<Z>& tb, const int32_t* t, int k, int K,
                      Z* Mc, Z* Ar, int small_stride, const int32_t* dyn = nullptr);
// complex128: v_mfma_f64_16x16x4f64; complex64: v_mfma_f32_16x16x4f32
void launch_lr_colblk(hipStream_t st, const Dims& d, const DevTabs& tb, const int32_t* t, int k, int K,
                      const zd* Mc, int small_stride, int variant, const int32_t* dyn = nullptr);
void launch_lr_colblk(hipStream_t st, const Dims& d, const DevTabsT<zf>& tb, const int32_t* t, int k, int K,
                      const zf* Mc, int small_stride, int variant, const int32_t* dyn = nullptr);
// Rows kk0 .. kk0 + nrows - 1 of M (as ld-strided rows of out, all N columns) from Lam and A
// (the factors kept without M); one frequency f.
template <class Z>
void launch_lr_build_m(hipStream_t st, const Dims& d, const DevTabsT<Z>& tb, int f, int kk0, int nrows, Z* out, int ld);
void launch_lr_rowblk(hipStream_t st, const Dims& d, const DevTabs& tb, const int32_t* t, int k, int K,
                      const zd* Ar, int small_stride, zd* Bt, size_t bt_stride, int variant,
                      const int32_t* dyn = nullptr);
void launch_lr_rowblk(hipStream_t st, const Dims& d, const DevTabsT<zf>& tb, const int32_t* t, int k, int K,
                      const zf* Ar, int small_stride, zf* Bt, size_t bt_stride, int variant,
                      const int32_t* dyn = nullptr);
// number of tile-shape variants of the skinny products for flips of k components (c64: FP32 kernel)
int skinny_variants(int k, bool c64 = false);
template <class Z>
void launch_apply_capinv(hipStream_t st, const Dims& d, const Z* acc, int acc_stride, const Z* Bt,
                         size_t bt_stride, int ldb, int k, const int32_t* s, Z* const* out_tab,
                         Z* out_single, size_t out_stride, int out_row0, int ld_out,
                         const Z* const* b_tab = nullptr, int b_col0 = 0);
template <class Z>
void launch_lr_small(hipStream_t st, const Dims& d, const DevTabsT<Z>& tb, const DevVecsT<Z>& v, const Z* acc,
                     int acc_stride, int k, int K);
// fused apply_capinv + lr_small + lr_blocks for gmax <= 32
// left != null: partition layout (groups [0, ngroups_part) have disjoint component lists; `left`
// holds the nleft components outside every group); null: every group's block plus 32-component
// chunks of all N
// The device-driven loop's accepted flip, applied to the component mask and the tile map by the
// finish launch (no workgroup of that kernel reads them): P[t[c]] += s[c] for c < k and
// tile[gtiles[i]] ^= 1 for i < dyn[kDynNt].
struct FlipUpdate {
    int8_t* P = nullptr;
    int32_t* tile = nullptr;
    const int32_t* t = nullptr;
    const int32_t* gtiles = nullptr;
};
template <class Z>
void launch_lr_finish(hipStream_t st, const Dims& d, const DevTabsT<Z>& tb, const DevVecsT<Z>& v, GroupTab gt,
                      const Z* acc, int acc_stride, const Z* Bt, size_t bt_stride, const int32_t* s, int k,
                      int K, const int32_t* dyn = nullptr, const int32_t* left = nullptr, int nleft = 0,
                      int ngroups_part = 0, FlipUpdate fu = FlipUpdate());
template <class Z>
void launch_lr_blocks(hipStream_t st, const Dims& d, const DevTabsT<Z>& tb, const DevVecsT<Z>& v, GroupTab gt,
                      int k, int K);
void launch_mask_update(hipStream_t st, int8_t* P, int32_t* tile, const int32_t* t, const int32_t* s, int k,
                        const int32_t* gtiles, int ntiles);

// ---- PLAIN accept
void launch_build_V(hipStream_t st, const Dims& d, const DevTabs& tb, const int32_t* t, int k, zd* V,
                    size_t v_stride);
void launch_snapshot_cinvu(hipStream_t st, const Dims& d, const DevTabs& tb, const int32_t* t,
                           const int32_t* s, int k, zd* CU, size_t cu_stride);
// VC (k x N, ld k) = V C^-1 on FP64 MFMA, V (k x N, ld k); every frequency (grid.y).  `variant`
// picks the tile shape (the column strip's shapes; k <= 24: 0 .. 5).
void launch_plain_vcinv(hipStream_t st, const Dims& d, const DevTabs& tb, int k, const zd* V, size_t v_stride,
                        zd* VC, size_t vc_stride, int variant = 1);
// C^-1 -= CU M on FP64 MFMA, CU (N x k, ld ldN), M (k x N, row c at c * ldN); every frequency
void launch_plain_rankk(hipStream_t st, const Dims& d, const DevTabs& tb, const zd* CU, size_t cu_stride,
                        const zd* M, size_t m_stride, int k);
void launch_plain_small(hipStream_t st, const Dims& d, const DevVecs& v, const zd* acc, int acc_stride,
                        const zd* CU, size_t cu_stride, const zd* M, size_t m_stride, const int32_t* t,
                        const int32_t* s, int k, zd* rut);

// ---- learned ranker inputs / labels (LOWRANK state)
void launch_features(hipStream_t st, const Dims& d, const DevVecs& v, GroupTab gt, const int32_t* tile,
                     const int8_t* P, const int32_t* owners, const int32_t* batch, int nb, int Tx, int Ty,
                     const ObjParams& op, float* out);
void launch_labels(hipStream_t st, const Dims& d, const DevVecs& v, const ObjParams& op, const zd* eobj_out, int nb,
                   float* out);

// ---- init
void launch_synth_g(hipStream_t st, int N, int ldN, const double* pos /*[N][3]*/, double kr, zd scale,
                    double soft, int row_major, zd* out);
template <class Z>
void launch_build_C(hipStream_t st, int N, int ldN, const int8_t* P, const Z* Gcm, Z* C);
template <class Z>
void launch_add_diag(hipStream_t st, int N, int ldN, Z* A, double v);
// out (complex64, ld ldN) = the complex128 matrix in rounded to nearest (the reference's cast of
// the file to c64, pngfopt.hpp:128-156)
void launch_cast_c64(hipStream_t st, int N, int ldN, const zd* in, zf* out);
void launch_transpose(hipStream_t st, int N, int ldN, const zd* in, zd* out);
void launch_asym(hipStream_t st, int N, int ldN, const zd* G, unsigned long long* out);
// G(i, j) += amp (u1 + i u2), uniform u from a hash of (seed, i, j): a non-reciprocal perturbation
void launch_perturb(hipStream_t st, int N, int ldN, zd* G, double amp, unsigned long long seed);
// *flag |= 1 when the column-major N x N matrix holds a NaN or an infinity (flag zeroed by the caller)
void launch_nonfinite(hipStream_t st, int N, int ldN, const zd* G, int32_t* flag);
template <class Z>
void launch_symmetrize(hipStream_t st, int N, int ldN, Z* W);
template <class Z>
void launch_port_sums(hipStream_t st, int N, int ldN, const Z* Gcm, const int32_t* port, int np, Z* gobj0,
                      Z* einc);
template <class Z>
void launch_pe(hipStream_t st, int N, const int8_t* P, const Z* einc, Z* pe);
// out[r] = sum_n A[r][n] x[n] in the arithmetic of Z, written as complex128
template <class Z>
void launch_dot_rows(hipStream_t st, int N, int ldN, int rows, const Z* A, const Z* x, zd* out);
template <class Z>
void launch_lr_block_current(hipStream_t st, const Dims& d, const DevTabsT<Z>& tb, GroupTab gt, int g, int K,
                             Z* Dblk);
template <class Z>
void launch_extract_blocks(hipStream_t st, const Dims& d, const Z* W0, GroupTab gt, Z* Dblk_f);

}  // namespace pngf
