// gfx950 kernels of the DBS Woodbury path.
//
// Notation (pngf-opt/pngfopt.hpp): N design components, P the 0/1 metal mask,
// C = I + P(G - I), S = C^-1 PEinc, R = Gobj C^-1, Eobj = R PEinc.  A flip changes
// the k components t (ascending) with signs s = +-1.
//   PLAIN   (reference formulation, pngfopt.hpp:379-539): cap = I + (G-I)[t,:] C^-1[:,t] diag(s)
//   LOWRANK (W = (G-I) C^-1, q = W PEinc): cap = I + W[t,t] diag(s), rhs = q_t + W[t,t] dPE
// Both give the same newEobj in exact arithmetic (the accept paths are in runtime.hip).
#include "kernels.h"

#include <type_traits>

namespace pngf {

// Bit pattern of a non-negative real as an unsigned key that orders like the value.
__device__ inline unsigned long long rbits(double x) { return (unsigned long long)__double_as_longlong(x); }
__device__ inline unsigned long long rbits(float x) { return (unsigned long long)__float_as_uint(x); }

// Wave-wide maximum and minimum of the ROCm device library (the HIP headers declare the 64-bit
// one only under HIP_ENABLE_EXTRA_WARP_SYNC_TYPES); every lane gets the result.
extern "C" __device__ __attribute__((const)) unsigned long long __ockl_wfred_max_u64(unsigned long long);
extern "C" __device__ __attribute__((const)) unsigned int __ockl_wfred_min_u32(unsigned int);

// x of lane `src` in every lane (src uniform)
__device__ inline zd readlane_z(zd x, int src) {
    const unsigned long long rx = (unsigned long long)__double_as_longlong(x.x);
    const unsigned long long ry = (unsigned long long)__double_as_longlong(x.y);
    const unsigned long long ox = (unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(unsigned)rx, src) |
                                  ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(unsigned)(rx >> 32), src) << 32);
    const unsigned long long oy = (unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(unsigned)ry, src) |
                                  ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(unsigned)(ry >> 32), src) << 32);
    return zmk(__longlong_as_double((long long)ox), __longlong_as_double((long long)oy));
}
__device__ inline zf readlane_z(zf x, int src) {
    return zmkf(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(x.x), src)),
                __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x.y), src)));
}

// ------------------------------------------------------------------ helpers --

// Components of flip group g whose metal state changes, ascending, with sign
// s = new - old and the position in the group's component list.  Serial (one
// thread); gmax <= 96.  reference: pngfopt.hpp:384-418 (dP scan / P_indices).
__device__ int change_set(const GroupTab& gt, int g, const int32_t* tile, const int8_t* P,
                          const int32_t* owners, int* t, int* sg, int* lp) {
    const int nt = gt.ntiles[g];
    const int32_t* gtl = gt.tiles + 4 * g;
    const int nc = gt.ncomp[g];
    const int32_t* cl = gt.comps + (size_t)g * gt.gmax;
    int k = 0;
    for (int a = 0; a < nc; a++) {
        const int i = cl[a];
        int nv = 0;
        for (int h = 0; h < 2; h++) {
            const int o = owners[2 * i + h];
            if (o < 0) continue;
            int v = tile[o];
            for (int q = 0; q < nt; q++) v ^= (gtl[q] == o) ? 1 : 0;
            nv |= v;
        }
        const int ov = P[i];
        if (nv != ov) { t[k] = i; sg[k] = nv - ov; lp[k] = a; k++; }
    }
    return k;
}

// LU factorisation with partial pivoting of the k x k leading block of the
// row-major LDS matrix a (leading dimension lda), applying the row operations to
// columns k..ncol-1 as well.  Pivot = first row of max |Re|+|Im| (LAPACK
// izamax / zgetf2); rows are swapped across all ncol columns and the multipliers
// l = a_ij * (1/a_jj) are stored below the diagonal.  piv[j] = pivot row.
// All threads of the block call it (blockDim.x a multiple of 64); k <= 128.  Returns
// LAPACK-style info (0 or j+1).
//
// Look-ahead: in the trailing update of column j, wave 0 updates column j + 1 first (lane i its
// rows j+1+i and j+65+i) and searches it for the next pivot from those registers, while the other
// waves update the columns after it; so a column costs two barriers and no search phase of its
// own.  Every entry goes through the same operations as in a column-at-a-time elimination.
template <class Z>
__device__ __forceinline__ int lds_lu(Z* a, int lda, int k, int ncol, int* piv) {
    typedef typename zreal<Z>::type R;
    const int tid = threadIdx.x, nth = blockDim.x, lane = tid & 63, nw = nth >> 6;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    __shared__ int s_info;
    __shared__ R s_pv[6];  // pivot, old diagonal entry, reciprocal of the pivot
    // the pivot of column j from this wave's values of it (rows j+lane and j+lane+64; wave 0
    // only): the first row of max |Re|+|Im| as two wave reductions (|.| >= 0 orders like its bit
    // pattern; NaN and empty lanes: key 0; a column of NaN only takes row j, as izamax does)
    auto pick = [&](int j, Z y0, Z y1) __attribute__((always_inline)) {
        R best = -1;
        int bi = 0x7fffffff;
        if (j + lane < k) {
            const R v = zabs1(y0);
            if (v > best) { best = v; bi = j + lane; }
        }
        if (j + lane + 64 < k) {
            const R v = zabs1(y1);
            if (v > best) { best = v; bi = j + lane + 64; }
        }
        const unsigned long long key = best >= 0 ? rbits(best) + 1ull : 0ull;
        const unsigned long long mk = __ockl_wfred_max_u64(key);
        const unsigned mp = __ockl_wfred_min_u32(key == mk ? (unsigned)bi : 0xffffffffu);
        const int p = mp < (unsigned)k ? (int)mp : j;
        // the pivot entry and the diagonal one, from the lanes holding them
        const Z pv = p - j < 64 ? readlane_z(y0, p - j) : readlane_z(y1, p - j - 64);
        const Z ajj = readlane_z(y0, 0);
        if (lane == 0) {
            piv[j] = p;
            const Z r = (pv.x == 0 && pv.y == 0) ? pv : zrecip(pv);
            s_pv[0] = pv.x; s_pv[1] = pv.y; s_pv[2] = ajj.x; s_pv[3] = ajj.y; s_pv[4] = r.x; s_pv[5] = r.y;
        }
    };
    // the search of column j read from LDS (the first column, and the one after a zero pivot)
    auto search = [&](int j) __attribute__((always_inline)) {
        if (wv == 0) {
            const Z y0 = j + lane < k ? a[(j + lane) * lda + j] : zc<Z>(0, 0);
            const Z y1 = j + lane + 64 < k ? a[(j + lane + 64) * lda + j] : zc<Z>(0, 0);
            pick(j, y0, y1);
        }
    };
    if (tid == 0) s_info = 0;
    search(0);
    __syncthreads();
    for (int j = 0; j < k; j++) {
        // the row interchange and the multipliers in one phase: column j is left to the
        // multiplier loop (row p's entry there is the old diagonal one, read before the barrier)
        const int p = piv[j];
        const Z pv = zc<Z>(s_pv[0], s_pv[1]), ajj = zc<Z>(s_pv[2], s_pv[3]), r = zc<Z>(s_pv[4], s_pv[5]);
        const bool zero = pv.x == 0 && pv.y == 0;
        if (p != j)
            for (int c = tid; c < ncol; c += nth)
                if (c != j || zero) { Z x = a[j * lda + c]; a[j * lda + c] = a[p * lda + c]; a[p * lda + c] = x; }
        if (zero) {
            if (tid == 0 && s_info == 0) s_info = j + 1;
            __syncthreads();
            if (j + 1 < k) search(j + 1);
            __syncthreads();
            continue;
        }
        for (int i = j + 1 + tid; i < k; i += nth) a[i * lda + j] = zmul(i == p ? ajj : a[i * lda + j], r);
        if (tid == 0) a[j * lda + j] = pv;
        __syncthreads();
        // wave 0: column j + 1 (rows j + 1 .. k - 1), then the next pivot from it
        if (wv == 0) {
            const int c = j + 1;
            const Z u = a[j * lda + c];
            const int i0 = j + 1 + lane, i1 = i0 + 64;
            Z y0 = zc<Z>(0, 0), y1 = zc<Z>(0, 0);
            if (i0 < k) {
                y0 = zfms(a[i0 * lda + j], u, a[i0 * lda + c]);
                a[i0 * lda + c] = y0;
            }
            if (i1 < k) {
                y1 = zfms(a[i1 * lda + j], u, a[i1 * lda + c]);
                a[i1 * lda + c] = y1;
            }
            if (j + 1 < k) pick(j + 1, y0, y1);
        }
        // the columns after it: lanes take columns, the other waves take rows (every wave reads
        // the same multiplier, a broadcast), four rows' loads issued before their stores; with one
        // wave, wave 0 does them as well
        if (wv > 0 || nw == 1) {
            const int rw = nw > 1 ? wv - 1 : 0, nr = nw > 1 ? nw - 1 : 1;
            for (int c = j + 2 + lane; c < ncol; c += 64) {
                const Z u = a[j * lda + c];
                int i = j + 1 + rw;
                for (; i + 3 * nr < k; i += 4 * nr) {
                    Z l[4], x[4];
#pragma unroll
                    for (int m = 0; m < 4; m++) {
                        l[m] = a[(i + m * nr) * lda + j];
                        x[m] = a[(i + m * nr) * lda + c];
                    }
#pragma unroll
                    for (int m = 0; m < 4; m++) a[(i + m * nr) * lda + c] = zfms(l[m], u, x[m]);
                }
                for (; i < k; i += nr) a[i * lda + c] = zfms(a[i * lda + j], u, a[i * lda + c]);
            }
        }
        __syncthreads();
    }
    __syncthreads();
    return s_info;
}

// Back substitution U x = c for the column `col` of the factored LDS matrix.
template <class Z>
__device__ void lds_backsolve(Z* a, int lda, int k, int col) {
    const int tid = threadIdx.x, nth = blockDim.x;
    for (int j = k - 1; j >= 0; j--) {
        if (tid == 0) a[j * lda + col] = zdiv(a[j * lda + col], a[j * lda + j]);
        __syncthreads();
        const Z xj = a[j * lda + col];
        for (int i = tid; i < j; i += nth) a[i * lda + col] = zfms(a[i * lda + j], xj, a[i * lda + col]);
        __syncthreads();
    }
}

// Inverse from the LU factors of lds_lu: column c of cap^-1 solves cap x = e_c
// (row interchanges, unit-lower forward and upper back substitution).  One thread per
// column, each writing only its own column of X (row-major, leading dimension ldx),
// so the columns need no synchronisation between them.
template <class Z>
__device__ void lu_inverse_cols(const Z* a, int lda, int k, const int* piv, Z* X, int ldx) {
    for (int c = threadIdx.x; c < k; c += blockDim.x) {
        for (int i = 0; i < k; i++) X[i * ldx + c] = zc<Z>(i == c ? 1 : 0, 0);
        for (int j = 0; j < k; j++) {
            const int p = piv[j];
            if (p != j) { Z x = X[j * ldx + c]; X[j * ldx + c] = X[p * ldx + c]; X[p * ldx + c] = x; }
        }
        for (int j = 0; j < k; j++) {
            const Z xj = X[j * ldx + c];
            for (int i = j + 1; i < k; i++) X[i * ldx + c] = zfms(a[i * lda + j], xj, X[i * ldx + c]);
        }
        for (int j = k - 1; j >= 0; j--) {
            const Z xj = zdiv(X[j * ldx + c], a[j * lda + j]);
            X[j * ldx + c] = xj;
            for (int i = 0; i < j; i++) X[i * ldx + c] = zfms(a[i * lda + j], xj, X[i * ldx + c]);
        }
    }
    __syncthreads();
}

// Entry (r, c) of the maintained diagonal block D (row-major, leading dimension gmax).  For
// symmetric W the block is read as (D + D^T) / 2, which removes the rounding asymmetry that the
// eager block updates accumulate (a symmetric-representation Woodbury update fed an asymmetric
// block amplifies that asymmetry at every near-singular capacitance matrix).
template <class Z>
__device__ inline Z dblk(const Z* D, int gmax, int r, int c, int sym) {
    typedef typename zreal<Z>::type R;
    const Z w = D[r * gmax + c];
    if (!sym) return w;
    const Z u = D[c * gmax + r];
    return zc<Z>(R(0.5) * (w.x + u.x), R(0.5) * (w.y + u.y));
}

// change_set computed by the 64 lanes of one wave (all of them call it): lane a tests
// component a of the group, and a ballot keeps the changed ones in ascending list order.
__device__ int change_set_wave(const GroupTab& gt, int g, const int32_t* tile, const int8_t* P,
                               const int32_t* owners, int* t, int* sg, int* lp) {
    const int lane = threadIdx.x & 63;
    const int nt = gt.ntiles[g];
    const int32_t* gtl = gt.tiles + 4 * g;
    const int nc = gt.ncomp[g];
    const int32_t* cl = gt.comps + (size_t)g * gt.gmax;
    int k = 0;
    for (int base = 0; base < nc; base += 64) {
        const int a = base + lane;
        bool ch = false;
        int sgn = 0, i = -1;
        if (a < nc) {
            i = cl[a];
            int nv = 0;
            for (int h = 0; h < 2; h++) {
                const int o = owners[2 * i + h];
                if (o < 0) continue;
                int v = tile[o];
                for (int q = 0; q < nt; q++) v ^= (gtl[q] == o) ? 1 : 0;
                nv |= v;
            }
            const int ov = P[i];
            ch = nv != ov;
            sgn = nv - ov;
        }
        const unsigned long long m = __ballot(ch);
        if (ch) {
            const int pos = k + __popcll(m & ((1ull << lane) - 1ull));
            t[pos] = i;
            sg[pos] = sgn;
            lp[pos] = a;
        }
        k += __popcll(m);
    }
    return k;
}

// Gauss-Jordan elimination with partial pivoting of the k x k leading block (k <= 32) of the
// LDS matrix a (leading dimension lda), applying the row operations to columns k..ncol-1.
// Rows stay in place: perm[i] is the storage row of logical row i, the order that zgetf2's
// row interchanges produce (pivot = first remaining logical row of max |Re|+|Im|).  Step j
// removes column j from every other row and writes only columns > j (column j of the other
// rows is not read again).  On return storage row perm[i] holds d_i in column i and d_i times
// the solution in columns k..ncol-1.  Returns LAPACK-style info (0, or j+1 for the first zero
// pivot, whose elimination step is skipped).  All threads of the block call it, after a barrier
// that follows the writes of a; blockDim.x is a multiple of 64.
template <class Z>
__device__ int lds_gj(Z* a, int lda, int k, int ncol, int* perm) {
    const int tid = threadIdx.x, lane = tid & 63;
    const int ri = tid & 31, cg = tid >> 5, ncg = blockDim.x >> 5;
    // every wave keeps its own copy of the row permutation (lane i: position i) and repeats the
    // pivot search, so one barrier per column suffices: the search for column j only reads
    // column j, which the elimination of column j does not write
    int pm = lane;
    int info = 0;
    for (int j = 0; j < k; j++) {
        // pivot search as one wave reduction: |Re|+|Im| >= 0, so its bit pattern orders like
        // the value; + 1 puts every candidate above the lanes outside [j, k) (key 0), and the
        // first lane holding the maximum key is the first logical row of the maximum
        const unsigned long long key =
            lane >= j && lane < k ? rbits(zabs1(a[pm * lda + j])) + 1ull : 0ull;
        const unsigned long long mk = __ockl_wfred_max_u64(key);
        const int bi = __ffsll((unsigned long long)__ballot(key == mk)) - 1;
        const int pj = __builtin_amdgcn_readlane(pm, j), pb = __builtin_amdgcn_readlane(pm, bi);
        if (lane == j) pm = pb;
        else if (lane == bi) pm = pj;
        if (mk == 1ull && info == 0) info = j + 1;  // the maximum is 0
        const int pr = pb;
        const Z pv = a[pr * lda + j];
        if ((pv.x != 0 || pv.y != 0) && ri < k && ri != pr) {
            // all reads of this thread's columns first, then the updates (rows pr and ri differ,
            // and only this thread writes row ri's columns j+1+cg, j+1+cg+ncg, ...)
            constexpr int U = 9;
            Z pvv[U], rvv[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int c = j + 1 + cg + u * ncg;
                if (c < ncol) {
                    pvv[u] = a[pr * lda + c];
                    rvv[u] = a[ri * lda + c];
                }
            }
            const Z l = zmul(a[ri * lda + j], zrecip(pv));
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int c = j + 1 + cg + u * ncg;
                if (c < ncol) a[ri * lda + c] = zfms(l, pvv[u], rvv[u]);
            }
            for (int c = j + 1 + cg + U * ncg; c < ncol; c += ncg) a[ri * lda + c] = zfms(l, a[pr * lda + c], a[ri * lda + c]);
        }
        __syncthreads();
    }
    if (tid < k) perm[tid] = pm;
    __syncthreads();
    return info;
}

// The k x (k+1) capacitance | rhs matrix and the k x k inverse live in LDS when
// gmax <= kLdsGmax (33 KB); larger flip sets (2- and 4-tile flips, k <= 96) use a
// per-block slot of a global scratch buffer instead, except in the LOWRANK scoring and accept
// kernels, whose k x (k+1) system alone (149 KB at k = 96) stays in LDS where the device allows
// it (the accept then builds the inverse in registers and writes it to the global slot).
template <class Z>
static size_t lds_matrix_bytes(int gmax, const Z* gscratch) {
    return gscratch ? 0 : (size_t)gmax * (2 * gmax + 1) * sizeof(Z);
}

// --------------------------------------------------------- LOWRANK eval ------
// One workgroup per (candidate b, frequency f): cap = I + D[t,t] diag(s),
// rhs = q_t + D[t,t] dPE, y = cap^-1 rhs, newEobj = Eobj + R[:,t] (dPE - s y),
// where dPE = -s Einc_t and D the maintained block of W over the group's components.
template <class Z>
__device__ void lr_eval_one(const Dims& d, const DevVecsT<Z>& v, const GroupTab& gt, const int32_t* __restrict__ tile,
                            const int8_t* __restrict__ P, const int32_t* __restrict__ owners,
                            const int32_t* batch, zd* __restrict__ out_eobj, int32_t* __restrict__ out_info,
                            Z* gscratch) {
    typedef typename zreal<Z>::type R;
    extern __shared__ __align__(16) unsigned char smraw[];
    Z* smz = (Z*)smraw;
    const int b = blockIdx.y, f = blockIdx.x, tid = threadIdx.x, nth = blockDim.x;
    const int gmax = d.gmax;
    Z* dpe = smz;                                 // [gmax]
    Z* wdg = dpe + gmax;                          // [gmax] diagonal of the block
    Z* qt = wdg + gmax;                           // [gmax] q[t]
    Z* rt = qt + gmax;                            // [ne][gmax] R[:, t]
    int* t = (int*)(rt + 3 * gmax);
    int* sg = t + gmax;
    int* lp = sg + gmax;
    int* piv = lp + gmax;
    __shared__ int s_k;
    const int g = batch[b];
    if (!gscratch) {
        if (tid < 64) {
            const int kk = change_set_wave(gt, g, tile, P, owners, t, sg, lp);
            if (tid == 0) s_k = kk;
        }
    } else if (tid == 0) {
        s_k = change_set(gt, g, tile, P, owners, t, sg, lp);
    }
    __syncthreads();
    const int k = s_k;
    // [k][k+1] capacitance | rhs: LDS when it fits, else this block's global slot.  The rest of
    // the candidate is written once for a matrix pointer and expanded for each case, so that
    // the LDS case compiles to LDS instructions (a pointer that may be either is a flat one).
    auto rest = [&](Z* a) __attribute__((always_inline)) {
    const int lda = k + 1;
    const Z* D = v.Dblk + ((size_t)f * d.ngroups + g) * gmax * gmax;
    const Z* q = v.q + (size_t)f * d.ldN;
    const Z* einc = v.Einc + (size_t)f * d.ldN;
    // every global read of the candidate in one phase: dPE, q_t, R[:, t] and the block (the
    // block gather: independent loads, several in flight per thread; the diagonal of W is kept
    // apart (wdg) so that the rhs below reads W exactly from LDS)
    for (int c = tid; c < k; c += nth) {
        Z e = einc[t[c]];
        dpe[c] = sg[c] > 0 ? zc<Z>(-e.x, -e.y) : e;
        qt[c] = q[t[c]];
    }
    for (int e = tid; e < d.ne * k; e += nth) {
        const int r = e / k, c = e - r * k;
        rt[r * gmax + c] = v.R[((size_t)f * d.ne + r) * d.ldN + t[c]];
    }
#pragma unroll 4
    for (int e = tid; e < k * k; e += nth) {
        const int r = e / k, c = e - r * k;
        Z w = dblk(D, gmax, lp[r], lp[c], d.sym);
        Z val = sg[c] > 0 ? w : zc<Z>(-w.x, -w.y);
        if (r == c) {
            wdg[r] = w;
            val.x += R(1);
        }
        a[r * lda + c] = val;
    }
    __syncthreads();
    // rhs_r = q_t_r + sum_c W[t_r][t_c] dPE_c, with W[t_r][t_c] = s_c a[r][c] off the diagonal
    for (int r = tid; r < k; r += nth) {
        Z acc = qt[r];
        for (int c = 0; c < k; c++) {
            Z w = wdg[r];
            if (c != r) {
                const Z v = a[r * lda + c];
                w = sg[c] > 0 ? v : zc<Z>(-v.x, -v.y);
            }
            acc = zfma(w, dpe[c], acc);
        }
        a[r * lda + k] = acc;
    }
    __syncthreads();
    int info;
    if (!gscratch && gmax <= kLdsGmax) {
        // Gauss-Jordan in LDS; y_c then goes to column k of storage row c (after a barrier, so
        // that no thread still reads the eliminated matrix)
        info = lds_gj(a, lda, k, k + 1, piv);
        Z y = zc<Z>(0, 0);
        if (tid < k) y = zdiv(a[piv[tid] * lda + k], a[piv[tid] * lda + tid]);
        __syncthreads();
        if (tid < k) a[tid * lda + k] = y;
        __syncthreads();
    } else {
        info = lds_lu(a, lda, k, k + 1, piv);
        lds_backsolve(a, lda, k, k);
    }
    if (tid < d.ne) {
        const int r = tid;
        if constexpr (std::is_same<Z, zd>::value) {
            zd acc = v.Eobj[f * d.ne + r];
            for (int c = 0; c < k; c++) {
                zd y = a[c * lda + k];
                zd w = sg[c] > 0 ? zsub(dpe[c], y) : zadd(dpe[c], y);
                acc = zfma(rt[r * gmax + c], w, acc);
            }
            out_eobj[((size_t)b * d.nf + f) * d.ne + r] = acc;
        } else {
            // complex64: the correction in single precision, added to the complex128 Eobj
            Z acc = zc<Z>(0, 0);
            for (int c = 0; c < k; c++) {
                Z y = a[c * lda + k];
                Z w = sg[c] > 0 ? zsub(dpe[c], y) : zadd(dpe[c], y);
                acc = zfma(rt[r * gmax + c], w, acc);
            }
            out_eobj[((size_t)b * d.nf + f) * d.ne + r] = zadd(v.Eobj[f * d.ne + r], zwide(acc));
        }
    }
    if (tid == 0 && out_info) out_info[b * d.nf + f] = info;
    };
    if (!gscratch)
        rest((Z*)(piv + gmax));
    else
        rest(gscratch + ((size_t)b * gridDim.x + f) * gmax * (2 * gmax + 1));
}

__device__ void loop_after_select_wave(const Dims& d, LoopState* ls, const GroupTab& gt, const int32_t* tile,
                                       const int8_t* P, const int32_t* owners, const int32_t* batch,
                                       const int32_t* info, const LoopCtl& c);
__device__ void loop_next_batch(LoopState* ls, int* B_out, int* pos_out, int* ncl_out);

// One workgroup per (candidate b, frequency f); workgroups past the batch size (read from
// nb_dev) return at once.
template <class Z>
__global__ void lr_eval_kernel(Dims d, DevVecsT<Z> v, GroupTab gt, const int32_t* __restrict__ tile,
                               const int8_t* __restrict__ P, const int32_t* __restrict__ owners,
                               const int32_t* __restrict__ batch, zd* __restrict__ out_eobj,
                               int32_t* __restrict__ out_info, Z* gscratch, const int32_t* nb_dev) {
    if (nb_dev && (int)blockIdx.y >= *nb_dev) return;
    lr_eval_one(d, v, gt, tile, P, owners, batch, out_eobj, out_info, gscratch);
}

// Loop control of the device-driven loop, one workgroup of 256 after the scoring kernel: every
// candidate's objective, the first improving one in batch order, the accept record and the
// next batch.  (A kernel of its own rather than the scoring kernel's last workgroup: on a
// multi-XCD device the device-scope fence that pattern needs in every workgroup writes back
// the XCD's L2 and cost more than the launch it saves.)
__global__ void __launch_bounds__(256) loop_control_kernel(Dims d, GroupTab gt, const int32_t* __restrict__ tile,
                                                           const int8_t* __restrict__ P,
                                                           const int32_t* __restrict__ owners, int32_t* batch,
                                                           const zd* __restrict__ eobj,
                                                           const int32_t* __restrict__ info, LoopCtl c,
                                                           int32_t* iterlog) {
    __shared__ int s_red[256];
    __shared__ double s_F;
    __shared__ int s_B, s_pos, s_ncl;
    const int tid = threadIdx.x, nth = blockDim.x;
    LoopState* ls = c.ls;
    const int nb = ls->nb;
    const double fb = ls->F_best;
    int best = 0x7fffffff;
    for (int b = tid; b < nb; b += nth) {
        const double F = objective_of(c.op, c.fc, eobj + (size_t)b * d.nf * d.ne);
        c.F[b] = F;
        if (F < fb && best == 0x7fffffff) best = b;
    }
    s_red[tid] = best;
    __syncthreads();
    for (int off = nth >> 1; off > 0; off >>= 1) {
        if (tid < off) s_red[tid] = min(s_red[tid], s_red[tid + off]);
        __syncthreads();
    }
    const int bi = s_red[0];
    if (bi != 0x7fffffff && bi % nth == tid) s_F = c.F[bi];
    __syncthreads();
    if (tid == 0) {
        ls->sel = bi == 0x7fffffff ? -1 : bi;
        ls->F_sel = bi == 0x7fffffff ? fb : s_F;
    }
    __syncthreads();
    if (tid < 64) loop_after_select_wave(d, ls, gt, tile, P, owners, batch, info, c);
    __syncthreads();
    if (tid == 0) loop_next_batch(ls, &s_B, &s_pos, &s_ncl);
    __syncthreads();
    for (int i = tid; i < s_B; i += nth) batch[i] = c.cl[(s_pos + i) % s_ncl];
    // this iteration's accept as the timing of its column strip needs it: pending, k, K
    if (iterlog && tid == 0) {
        iterlog[0] = ls->dyn[kDynActive];
        iterlog[1] = ls->dyn[kDynK];
        iterlog[2] = ls->dyn[kDynKK];
    }
}

void launch_loop_control(hipStream_t st, const Dims& d, GroupTab gt, const int32_t* tile, const int8_t* P,
                         const int32_t* owners, int32_t* batch, const zd* eobj, const int32_t* info,
                         const LoopCtl& c, int32_t* iterlog) {
    hipLaunchKernelGGL(loop_control_kernel, dim3(1), dim3(256), 0, st, d, gt, tile, P, owners, batch, eobj, info, c,
                       iterlog);
}


// LOWRANK scoring of one (candidate b, frequency f) per wave, for flips of at most G <= 32
// components: lane r holds row r of [cap | rhs] in registers (cap = I + D[t,t] diag(s), D the
// maintained block) and the Gauss-Jordan elimination runs with neither LDS traffic nor
// barriers: the pivot search is a wave reduction and the pivot row is read across lanes.  The
// rows stay in their lanes; pos is the lane's logical row, the order zgetf2's interchanges
// produce.  The operations and their order are those of lr_eval_one with lds_gj, so the scores
// are bitwise the same.
template <class Z, int G>
__global__ void __launch_bounds__(64) lr_eval_wave_kernel(Dims d, DevVecsT<Z> v, GroupTab gt,
                                                          const int32_t* __restrict__ tile,
                                                          const int8_t* __restrict__ P,
                                                          const int32_t* __restrict__ owners,
                                                          const int32_t* __restrict__ batch, zd* __restrict__ out_eobj,
                                                          int32_t* __restrict__ out_info, const int32_t* nb_dev) {
    typedef typename zreal<Z>::type R;
    const int b = blockIdx.y, f = blockIdx.x, lane = threadIdx.x;
    if (nb_dev && b >= *nb_dev) return;
    __shared__ Z dpe[G], qt[G], rt[3 * G], ys[G];
    __shared__ int t[G], sg[G], lp[G];
    const int gmax = d.gmax;
    const int g = batch[b];
    const int k = change_set_wave(gt, g, tile, P, owners, t, sg, lp);
    __syncthreads();
    const Z* D = v.Dblk + ((size_t)f * d.ngroups + g) * gmax * gmax;
    if (lane < k) {
        const Z e = v.Einc[(size_t)f * d.ldN + t[lane]];
        dpe[lane] = sg[lane] > 0 ? zc<Z>(-e.x, -e.y) : e;
        qt[lane] = v.q[(size_t)f * d.ldN + t[lane]];
    }
    for (int e = lane; e < d.ne * k; e += 64) {
        const int r = e / k, c = e - r * k;
        rt[r * G + c] = v.R[((size_t)f * d.ne + r) * d.ldN + t[c]];
    }
    const bool row = lane < k;
    const int lr = row ? lp[lane] : 0;
    Z a[G];
    Z wd = zc<Z>(0, 0);  // W[t_r][t_r]
#pragma unroll
    for (int c = 0; c < G; c++) {
        a[c] = zc<Z>(0, 0);
        if (row && c < k) {
            const Z w = dblk(D, gmax, lr, lp[c], d.sym);
            Z val = sg[c] > 0 ? w : zc<Z>(-w.x, -w.y);
            if (c == lane) {
                wd = w;
                val.x += R(1);
            }
            a[c] = val;
        }
    }
    __syncthreads();
    // rhs_r = q_t_r + sum_c W[t_r][t_c] dPE_c, with W[t_r][t_c] = s_c cap[r][c] off the diagonal
    Z bv = zc<Z>(0, 0);
    if (row) {
        bv = qt[lane];
#pragma unroll
        for (int c = 0; c < G; c++) {
            if (c < k) {
                Z w = wd;
                if (c != lane) w = sg[c] > 0 ? a[c] : zc<Z>(-a[c].x, -a[c].y);
                bv = zfma(w, dpe[c], bv);
            }
        }
    }
    int pos = lane, info = 0;
#pragma unroll
    for (int j = 0; j < G; j++) {
        if (j < k) {
            // pivot: the first logical row >= j of maximal |Re|+|Im| in column j (the bit
            // pattern of a non-negative double orders like its value; + 1 keeps every row above
            // the lanes that take no part)
            const unsigned long long key =
                row && pos >= j ? rbits(zabs1(a[j])) + 1ull : 0ull;
            const unsigned long long mk = __ockl_wfred_max_u64(key);
            const unsigned int mp = __ockl_wfred_min_u32(key == mk ? (unsigned)pos : 0xffffffffu);
            const int pr = __ffsll((unsigned long long)__ballot(row && pos == (int)mp)) - 1;
            if (pos == j) pos = (int)mp;
            else if (pos == (int)mp) pos = j;
            if (mk == 1ull && info == 0) info = j + 1;  // the maximum is 0
            const Z pv = readlane_z(a[j], pr);
            if (pv.x != 0 || pv.y != 0) {
                const bool upd = row && lane != pr;
                const Z l = zmul(a[j], zrecip(pv));
#pragma unroll
                for (int c = j + 1; c < G; c++) {
                    if (c < k) {
                        const Z pc = readlane_z(a[c], pr);
                        if (upd) a[c] = zfms(l, pc, a[c]);
                    }
                }
                const Z pb = readlane_z(bv, pr);
                if (upd) bv = zfms(l, pb, bv);
            }
        }
    }
    // this lane's storage row is logical row pos: d_pos in column pos, d_pos y_pos in the rhs
    if (row) {
        Z dd = a[0];
#pragma unroll
        for (int c = 1; c < G; c++)
            if (c == pos) dd = a[c];
        ys[pos] = zdiv(bv, dd);
    }
    __syncthreads();
    if (lane < d.ne) {
        if constexpr (std::is_same<Z, zd>::value) {
            zd acc = v.Eobj[f * d.ne + lane];
            for (int c = 0; c < k; c++) {
                const zd y = ys[c];
                const zd w = sg[c] > 0 ? zsub(dpe[c], y) : zadd(dpe[c], y);
                acc = zfma(rt[lane * G + c], w, acc);
            }
            out_eobj[((size_t)b * d.nf + f) * d.ne + lane] = acc;
        } else {
            Z acc = zc<Z>(0, 0);
            for (int c = 0; c < k; c++) {
                const Z y = ys[c];
                const Z w = sg[c] > 0 ? zsub(dpe[c], y) : zadd(dpe[c], y);
                acc = zfma(rt[lane * G + c], w, acc);
            }
            out_eobj[((size_t)b * d.nf + f) * d.ne + lane] = zadd(v.Eobj[f * d.ne + lane], zwide(acc));
        }
    }
    if (lane == 0 && out_info) out_info[b * d.nf + f] = info;
}

// Dynamic LDS a workgroup may take on the current device, less what kernels declare statically
// besides it (a few scalars), from the device attribute; 0 if the query fails.
static size_t lds_per_workgroup() {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        return 0;
    return v > 256 ? (size_t)v - 256 : 0;
}

template <class Z>
void launch_lr_eval(hipStream_t st, const Dims& d, const DevVecsT<Z>& v, GroupTab gt, const int32_t* tile,
                    const int8_t* P, const int32_t* owners, const int32_t* batch, int nb, zd* out_eobj,
                    int32_t* out_info, Z* gscratch, const int32_t* nb_dev) {
    // frequency fastest in dispatch order: with a device-resident batch size the workgroups of
    // the batch's candidates are dispatched first, the ones past it (which return at once) after
    dim3 grid(d.nf, nb);
    // A wave per candidate takes longer per candidate than a workgroup of two (its column
    // updates run on one wave) but issues far less work in all (no LDS round trips or barriers):
    // it is the faster choice once the batch gives every SIMD a few candidates, as a full sweep
    // does.  The device-driven loop's batches (32 candidates after an accept) stay on the
    // workgroup kernel.  Both give bitwise the same scores.
    const bool wave = !gscratch && d.gmax <= 32 && !nb_dev && (int64_t)nb * d.nf >= kWaveEvalMin;
    if (wave && d.gmax <= 24) {
        hipLaunchKernelGGL((lr_eval_wave_kernel<Z, 24>), grid, dim3(64), 0, st, d, v, gt, tile, P, owners, batch,
                           out_eobj, out_info, nb_dev);
        return;
    }
    if (wave) {
        hipLaunchKernelGGL((lr_eval_wave_kernel<Z, 32>), grid, dim3(64), 0, st, d, v, gt, tile, P, owners, batch,
                           out_eobj, out_info, nb_dev);
        return;
    }
    int nth = d.gmax <= 32 ? 128 : 256;
    const size_t small = (size_t)4 * d.gmax * sizeof(int) + (size_t)6 * d.gmax * sizeof(Z);
    size_t sm = small + lds_matrix_bytes(d.gmax, gscratch);
    if (gscratch) {
        // larger flips (k up to 96: the reference driver's 2- and 4-tile flips): the k x (k+1)
        // system still goes to LDS, for the LU with partial pivoting, when the device's per-
        // workgroup LDS holds it (96 x 97 complex128 entries: 149 KB), instead of the global slot
        // whose every elimination step waits on the L2
        const size_t need = small + (size_t)d.gmax * (d.gmax + 1) * sizeof(Z);
        if (need <= lds_per_workgroup()) {
            (void)hipFuncSetAttribute((const void*)lr_eval_kernel<Z>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)need);
            gscratch = nullptr;
            sm = need;
            // one workgroup per CU (the LDS it takes): 1024 threads, four waves per SIMD to hide
            // the latencies of the elimination steps (each thread's sums run as with 256)
            nth = 1024;
        }
    }
    hipLaunchKernelGGL(lr_eval_kernel<Z>, grid, dim3(nth), sm, st, d, v, gt, tile, P, owners, batch, out_eobj,
                       out_info, gscratch, nb_dev);
}

// ------------------------------------------------------------ PLAIN eval -----
// Split-N streaming pass (reference pngfopt.hpp:420-480, fused): for each n in the
// split, stage V = (G-I)[t, n-chunk] (rows of the row-major G copy) and
// C^-1[n-chunk, t] (columns of the column-major C^-1) in LDS, form
// Snew = S - C^-1[:,t] (s o Einc_t) on the fly and accumulate the augmented
// product  [V ; Gobj] x [C^-1[:,t] , Snew]  whose blocks are V C^-1[:,t] (the
// capacitance), V Snew and Gobj Snew.  Each thread owns a 4x4 block of the product
// and a strided subset of the chunk; partials are reduced in LDS in a fixed order.
constexpr int PL_TILE = 48;     // rows / cols of the product per workgroup tile

__global__ void __launch_bounds__(256) plain_partial_kernel(Dims d, DevTabs tb, DevVecs v, GroupTab gt,
        const int32_t* __restrict__ tile, const int8_t* __restrict__ P, const int32_t* __restrict__ owners,
        const int32_t* __restrict__ batch, int nsplit, int PL_NC, zd* __restrict__ partial) {
    extern __shared__ char smem[];
    const int item = blockIdx.x, split = blockIdx.y, ttile = blockIdx.z;
    const int b = item / d.nf, f = item - b * d.nf;
    const int tid = threadIdx.x;
    const int gmax = d.gmax, ne = d.ne;
    int* t = (int*)smem;
    int* sg = t + 96;
    int* lp = sg + 96;
    __shared__ int s_k;
    zd* es = (zd*)(lp + 96);              // s o Einc_t   [96]
    zd* Cc = es + 96;                     // [k+1][PL_NC]   C^-1 columns + Snew
    const int g = batch[b];
    if (tid == 0) s_k = change_set(gt, g, tile, P, owners, t, sg, lp);
    __syncthreads();
    const int k = s_k;
    const int nrow = k + ne, ncolp = k + 1;
    const int ntr = (nrow + PL_TILE - 1) / PL_TILE, ntc = (ncolp + PL_TILE - 1) / PL_TILE;
    if (ttile >= ntr * ntc) return;
    const int r0 = (ttile / ntc) * PL_TILE, c0 = (ttile % ntc) * PL_TILE;
    const int rows = min(PL_TILE, nrow - r0), cols = min(PL_TILE, ncolp - c0);
    const int nrb = (rows + 3) >> 2, ncb = (cols + 3) >> 2, nblk = nrb * ncb;
    const int nsub = 256 / nblk;
    zd* Vt = Cc + (size_t)(k + 1) * PL_NC;     // [rows_pad][PL_NC]
    const int rows_pad = nrb * 4, cols_pad = ncb * 4;
    const zd* einc = v.Einc + (size_t)f * d.ldN;
    for (int c = tid; c < k; c += 256) { zd e = einc[t[c]]; es[c] = sg[c] > 0 ? e : zmk(-e.x, -e.y); }

    const zd* Gr = tb.Grow[f];
    const zd* Ci = tb.Cinv[f];
    const zd* S = v.S + (size_t)f * d.ldN;
    const zd* Go = v.Gobj + (size_t)f * ne * d.ldN;

    const int per = (d.N + nsplit - 1) / nsplit;
    const int n_begin = split * per, n_end = min(d.N, n_begin + per);

    const int blk = tid / nsub, sub = tid - blk * nsub;
    const bool active = blk < nblk;
    const int rb = active ? blk / ncb : 0, cb = active ? blk - rb * ncb : 0;
    zd acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = zmk(0.0, 0.0);

    __syncthreads();
    for (int nb0 = n_begin; nb0 < n_end; nb0 += PL_NC) {
        const int nlen = min(PL_NC, n_end - nb0);
        // stage C^-1 columns (all k: Snew needs them) and the tile's V / Gobj rows
        for (int e = tid; e < k * PL_NC; e += 256) {
            const int j = e / PL_NC, nn = e - j * PL_NC;
            Cc[j * PL_NC + nn] = nn < nlen ? Ci[(size_t)t[j] * d.ldN + nb0 + nn] : zmk(0.0, 0.0);
        }
        for (int e = tid; e < rows_pad * PL_NC; e += 256) {
            const int i = e / PL_NC, nn = e - i * PL_NC;
            const int gi = r0 + i, n = nb0 + nn;
            zd val = zmk(0.0, 0.0);
            if (nn < nlen && i < rows) {
                if (gi < k) { val = Gr[(size_t)t[gi] * d.ldN + n]; if (n == t[gi]) val.x -= 1.0; }
                else val = Go[(size_t)(gi - k) * d.ldN + n];
            }
            Vt[i * PL_NC + nn] = val;
        }
        __syncthreads();
        for (int nn = tid; nn < PL_NC; nn += 256) {
            zd s = nn < nlen ? S[nb0 + nn] : zmk(0.0, 0.0);
            for (int j = 0; j < k; j++) s = zfms(Cc[j * PL_NC + nn], es[j], s);
            Cc[k * PL_NC + nn] = nn < nlen ? s : zmk(0.0, 0.0);
        }
        __syncthreads();
        if (active) {
            for (int nn = sub; nn < nlen; nn += nsub) {
                zd vr[4], cc[4];
#pragma unroll
                for (int i = 0; i < 4; i++) vr[i] = Vt[(rb * 4 + i) * PL_NC + nn];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int gc = c0 + cb * 4 + j;
                    cc[j] = gc < ncolp ? Cc[gc * PL_NC + nn] : zmk(0.0, 0.0);
                }
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 4; j++) acc[i][j] = zfma(vr[i], cc[j], acc[i][j]);
            }
        }
        __syncthreads();
    }
    // fixed-order reduction of the nsub partial blocks through LDS
    zd* red = Cc;  // reuse: rows_pad x cols_pad entries
    for (int e = tid; e < rows_pad * cols_pad; e += 256) red[e] = zmk(0.0, 0.0);
    __syncthreads();
    for (int sround = 0; sround < nsub; sround++) {
        if (active && sub == sround) {
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    zd* p = &red[(rb * 4 + i) * cols_pad + cb * 4 + j];
                    *p = zadd(*p, acc[i][j]);
                }
        }
        __syncthreads();
    }
    // write this split's partial: full (gmax+ne) x (gmax+1) layout per (item, split)
    const int pstride = (gmax + ne) * (gmax + 1);
    zd* out = partial + ((size_t)item * nsplit + split) * pstride;
    for (int e = tid; e < rows * cols; e += 256) {
        const int i = e / cols, j = e - i * cols;
        out[(r0 + i) * (gmax + 1) + c0 + j] = red[i * cols_pad + j];
    }
}

static int plain_nc(int gmax) { return gmax <= 32 ? 32 : 16; }

static size_t plain_partial_smem(int gmax, int ne) {
    const int PL_NC = plain_nc(gmax);
    const int rp = ((gmax + ne + 3) / 4) * 4, cp = ((gmax + 1 + 3) / 4) * 4;
    const int rows_pad = rp < PL_TILE ? rp : PL_TILE, cols_pad = cp < PL_TILE ? cp : PL_TILE;
    size_t chunk = (size_t)(gmax + 1) * PL_NC + (size_t)rows_pad * PL_NC;
    size_t red = (size_t)rows_pad * cols_pad;
    return (size_t)3 * 96 * sizeof(int) + 96 * sizeof(zd) + (chunk > red ? chunk : red) * sizeof(zd);
}

// Reduce the split partials in a fixed order, then cap = I + (V C^-1[:,t]) diag(s),
// y = cap^-1 (V Snew) (LU, pngfopt.hpp:488-511) and
// newEobj = Gobj Snew - R[:,t] diag(s) y (pngfopt.hpp:513-537).  With `store`,
// also keeps the factors for the accept in the acc layout (capinv, y, newEobj).
__global__ void plain_finish_kernel(Dims d, DevVecs v, GroupTab gt, const int32_t* __restrict__ tile,
                                    const int8_t* __restrict__ P, const int32_t* __restrict__ owners,
                                    const int32_t* __restrict__ batch, int nsplit, const zd* __restrict__ partial,
                                    zd* __restrict__ out_eobj, int32_t* __restrict__ out_info, zd* store,
                                    int acc_stride, zd* gscratch) {
    extern __shared__ char smem[];
    const int item = blockIdx.x, b = item / d.nf, f = item - b * d.nf;
    const int tid = threadIdx.x, nth = blockDim.x;
    const int gmax = d.gmax, ne = d.ne;
    int* t = (int*)smem;
    int* sg = t + 96;
    int* lp = sg + 96;
    int* piv = lp + 96;
    __shared__ int s_k;
    zd* gs = (zd*)(piv + 96);       // [ne]
    zd* tmp = gs + 4;               // [96]
    zd* a = gscratch ? gscratch + (size_t)item * gmax * (2 * gmax + 1) : tmp + 96;  // [k][k+1], then [k][k]
    const int g = batch[b];
    if (tid == 0) s_k = change_set(gt, g, tile, P, owners, t, sg, lp);
    __syncthreads();
    const int k = s_k, lda = k + 1;
    const int pstride = (gmax + ne) * (gmax + 1);
    const zd* part = partial + (size_t)item * nsplit * pstride;
    for (int e = tid; e < (k + ne) * (k + 1); e += nth) {
        const int i = e / (k + 1), j = e - i * (k + 1);
        zd s = zmk(0.0, 0.0);
        for (int sp = 0; sp < nsplit; sp++) s = zadd(s, part[(size_t)sp * pstride + i * (gmax + 1) + j]);
        if (i < k) {
            if (j < k) {
                if (sg[j] < 0) s = zmk(-s.x, -s.y);
                if (i == j) s.x += 1.0;
            }
            a[i * lda + j] = s;
        } else if (j == k) {
            gs[i - k] = s;
        }
    }
    __syncthreads();
    int info = lds_lu(a, lda, k, k + 1, piv);
    lds_backsolve(a, lda, k, k);
    if (tid < ne) {
        const int r = tid;
        const zd* R = v.R + ((size_t)f * ne + r) * d.ldN;
        zd acc = gs[r];
        for (int c = 0; c < k; c++) {
            zd ru = R[t[c]];
            if (sg[c] < 0) ru = zmk(-ru.x, -ru.y);
            acc = zfms(ru, a[c * lda + k], acc);
        }
        out_eobj[((size_t)b * d.nf + f) * ne + r] = acc;
        if (store) store[(size_t)f * acc_stride + 2 * k * k + 2 * k + r] = acc;  // newEobj slot
    }
    if (tid == 0 && out_info) out_info[item] = info;
    if (store) {
        zd* st = store + (size_t)f * acc_stride;
        for (int c = tid; c < k; c += nth) st[2 * k * k + c] = a[c * lda + k];  // y
        zd* X = a + k * lda;
        lu_inverse_cols(a, lda, k, piv, X, k);
        for (int e = tid; e < k * k; e += nth) st[e] = X[e];
    }
}

// split-N pass of the PLAIN score on FP64 MFMA (defined with the other MFMA kernels below); false
// when the group size does not fit its tiles
bool launch_plain_partial_mfma(hipStream_t st, const Dims& d, const DevTabs& t, const DevVecs& v, GroupTab gt,
                               const int32_t* tile, const int8_t* P, const int32_t* owners, const int32_t* batch,
                               int nitems, int nsplit, zd* partial);

void launch_plain_eval(hipStream_t st, const Dims& d, const DevTabs& t, const DevVecs& v, GroupTab gt,
                       const int32_t* tile, const int8_t* P, const int32_t* owners, const int32_t* batch,
                       int nb, int nsplit, zd* partial, zd* out_eobj, int32_t* out_info, zd* store,
                       int acc_stride, zd* gscratch) {
    const int nitems = nb * d.nf;
    // groups of up to 31 components (k + ne <= 32): the FP64-MFMA pass; larger ones: VALU tiles
    if (!launch_plain_partial_mfma(st, d, t, v, gt, tile, P, owners, batch, nitems, nsplit, partial)) {
        const int ntiles_max = ((d.gmax + d.ne + PL_TILE - 1) / PL_TILE) * ((d.gmax + 1 + PL_TILE - 1) / PL_TILE);
        hipLaunchKernelGGL(plain_partial_kernel, dim3(nitems, nsplit, ntiles_max), dim3(256),
                           plain_partial_smem(d.gmax, d.ne), st, d, t, v, gt, tile, P, owners, batch, nsplit,
                           plain_nc(d.gmax), partial);
    }
    size_t sm = 4 * 96 * sizeof(int) + 100 * sizeof(zd) + lds_matrix_bytes(d.gmax, gscratch);
    hipLaunchKernelGGL(plain_finish_kernel, dim3(nitems), dim3(d.gmax <= 32 ? 64 : 256), sm, st, d, v, gt, tile,
                       P, owners, batch, nsplit, partial, out_eobj, out_info, store, acc_stride, gscratch);
}

// ------------------------------------------------------------- objective -----
// One thread per candidate: objective_of (common.h); then the first candidate in
// batch order with F < F_best.
__global__ void objective_kernel(ObjParams op, const FreqConst* __restrict__ fc, const zd* __restrict__ eobj,
                                 int nb, double* __restrict__ F_out, const int32_t* nb_dev) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (nb_dev) nb = *nb_dev;
    if (b >= nb) return;
    F_out[b] = objective_of(op, fc, eobj + (size_t)b * op.nf * op.ne);
}

__global__ void select_kernel(const double* __restrict__ F, int nb, const double* __restrict__ F_best,
                              int32_t* sel, double* F_sel, const int32_t* nb_dev) {
    __shared__ int red[256];
    const int tid = threadIdx.x;
    if (nb_dev) nb = *nb_dev;
    const double fb = *F_best;
    int best = 0x7fffffff;
    for (int b = tid; b < nb; b += 256)
        if (F[b] < fb) { best = b; break; }
    red[tid] = best;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off) red[tid] = min(red[tid], red[tid + off]);
        __syncthreads();
    }
    if (tid == 0) {
        const int bi = red[0];
        *sel = bi == 0x7fffffff ? -1 : bi;
        *F_sel = bi == 0x7fffffff ? fb : F[bi];
    }
}

void launch_objective(hipStream_t st, const ObjParams& op, const FreqConst* fc, const zd* eobj, int nb,
                      double* F_out, const double* F_best, int32_t* sel, double* F_sel, const int32_t* nb_dev) {
    hipLaunchKernelGGL(objective_kernel, dim3((nb + 127) / 128), dim3(128), 0, st, op, fc, eobj, nb, F_out, nb_dev);
    if (sel) hipLaunchKernelGGL(select_kernel, dim3(1), dim3(256), 0, st, F_out, nb, F_best, sel, F_sel, nb_dev);
}

// ------------------------------------------------------- LOWRANK accept -----
// acc layout per frequency (acc_stride zd):
//   [0, k*k)            capinv (row-major)
//   [k*k, 2k*k)         scratch LU
//   [2k*k, 2k*k+k)      y
//   [2k*k+k, 2k*k+2k)   w = dPE - s o y
//   [2k*k+2k, +ne)      newEobj
//   [2k*k+2k+4, +ne*k)  Rt = R[:, t] (snapshot)
template <class Z>
__global__ __launch_bounds__(256) void lr_prep_kernel(Dims d, DevVecsT<Z> v, int g, const int32_t* __restrict__ tt,
                               const int32_t* __restrict__ ss, const int32_t* __restrict__ lpp, int k, Z* acc,
                               int acc_stride, int32_t* info_out, Z* gscratch, const int32_t* dyn,
                               const int32_t* g_dev, int lds_big) {
    typedef typename zreal<Z>::type R;
    extern __shared__ char smem[];
    if (dyn) {
        if (!dyn[kDynActive]) return;
        k = dyn[kDynK];
        g = *g_dev;
    }
    const int f = blockIdx.x, tid = threadIdx.x, nth = blockDim.x;
    const int gmax = d.gmax, ne = d.ne;
    int* piv = (int*)smem;
    Z* dpe = (Z*)(piv + 100);
    Z* tmp = dpe + 96;
    // the matrix: LDS for k <= 32, else this frequency's global slot, or, with lds_big (k <= 96
    // where the workgroup's LDS holds k x (k+1) entries), LDS for the LU, whose factors then go to
    // the global slot (fac) for lr_prep_inv_kernel and lr_prep_lam_kernel (the body is expanded
    // for each, so that the LDS cases compile to LDS instructions)
    auto rest = [&](Z* a, Z* fac) __attribute__((always_inline)) {
    // k <= 32 (LDS): Gauss-Jordan on [cap | rhs | I]; larger k: LU, back substitution, inverse
    const bool gj = gscratch == nullptr;
    const int lda = gj ? 2 * k + 1 : k + 1;
    const Z* D = v.Dblk + ((size_t)f * d.ngroups + g) * gmax * gmax;
    const Z* q = v.q + (size_t)f * d.ldN;
    const Z* einc = v.Einc + (size_t)f * d.ldN;
    Z* st = acc + (size_t)f * acc_stride;
    for (int c = tid; c < k; c += nth) { Z e = einc[tt[c]]; dpe[c] = ss[c] > 0 ? zc<Z>(-e.x, -e.y) : e; }
    __syncthreads();
    Z* wdg = tmp;  // [k] W[t_r][t_r]
#pragma unroll 4
    for (int e = tid; e < k * k; e += nth) {
        const int r = e / k, c = e - r * k;
        Z w = dblk(D, gmax, lpp[r], lpp[c], d.sym);
        Z val = ss[c] > 0 ? w : zc<Z>(-w.x, -w.y);
        if (r == c) {
            wdg[r] = w;
            val.x += R(1);
        }
        a[r * lda + c] = val;
    }
    __syncthreads();
    // rhs from the LDS copy of the block (W[t_r][t_c] = s_c a[r][c] off the diagonal)
    for (int r = tid; r < k; r += nth) {
        Z s = q[tt[r]];
        for (int c = 0; c < k; c++) {
            Z w = wdg[r];
            if (c != r) {
                const Z v = a[r * lda + c];
                w = ss[c] > 0 ? v : zc<Z>(-v.x, -v.y);
            }
            s = zfma(w, dpe[c], s);
        }
        a[r * lda + k] = s;
    }
    if (gj)
        for (int e = tid; e < k * k; e += nth) {
            const int r = e / k, c = e - r * k;
            a[r * lda + k + 1 + c] = zc<Z>(r == c ? 1 : 0, 0);
        }
    __syncthreads();
    int info;
    if (gj) {
        info = lds_gj(a, lda, k, 2 * k + 1, piv);
    } else {
        info = lds_lu(a, lda, k, k + 1, piv);
        lds_backsolve(a, lda, k, k);
    }
    for (int c = tid; c < k; c += nth) {
        Z y = gj ? zdiv(a[piv[c] * lda + k], a[piv[c] * lda + c]) : a[c * lda + k];
        st[2 * k * k + c] = y;
        st[2 * k * k + k + c] = ss[c] > 0 ? zsub(dpe[c], y) : zadd(dpe[c], y);
    }
    for (int e = tid; e < ne * k; e += nth) {
        const int r = e / k, c = e - r * k;
        st[2 * k * k + 2 * k + 4 + e] = v.R[((size_t)f * ne + r) * d.ldN + tt[c]];
    }
    __syncthreads();
    if (tid < ne) {
        const int r = tid;
        if constexpr (std::is_same<Z, zd>::value) {
            zd s = v.Eobj[f * ne + r];
            for (int c = 0; c < k; c++) s = zfma(st[2 * k * k + 2 * k + 4 + r * k + c], st[2 * k * k + k + c], s);
            st[2 * k * k + 2 * k + r] = s;
            v.Eobj[f * ne + r] = s;
        } else {
            Z s = zc<Z>(0, 0);
            for (int c = 0; c < k; c++) s = zfma(st[2 * k * k + 2 * k + 4 + r * k + c], st[2 * k * k + k + c], s);
            const zd e = zadd(v.Eobj[f * ne + r], zwide(s));
            st[2 * k * k + 2 * k + r] = zc<Z>((R)e.x, (R)e.y);
            v.Eobj[f * ne + r] = e;
        }
    }
    if (fac) {
        // the factors (k x k, leading dimension k) and the pivots (as ints in the scratch-LU
        // place of the acc layout) for the inverse, which more workgroups share
        for (int e = tid; e < k * k; e += nth) {
            const int r = e / k, c = e - r * k;
            fac[e] = a[r * lda + c];
        }
        int* pv = (int*)(st + k * k);
        for (int c = tid; c < k; c += nth) pv[c] = piv[c];
        if (tid == 0 && info_out) info_out[f] = info;
        return;
    }
    // cap^-1: X[r][c] = a[perm r][k+1+c] / d_r after Gauss-Jordan, else from the LU factors
    Z* X = a + k * lda;
    if (!gj) lu_inverse_cols(a, lda, k, piv, X, k);
    auto xinv = [&](int r, int c) {
        return gj ? zdiv(a[piv[r] * lda + k + 1 + c], a[piv[r] * lda + r]) : X[r * k + c];
    };
    // Lambda = diag(s) cap^-1 = (diag(s) + W[t,t])^-1 is symmetric with W; for symmetric W it is
    // stored exactly symmetric: X'[r][c] = (X[r][c] + s_r s_c X[c][r]) / 2, so diag(s) X' = sym(Lambda)
    for (int e = tid; e < k * k; e += nth) {
        const int r = e / k, c = e - r * k;
        Z x = xinv(r, c);
        if (d.sym) {
            Z u = xinv(c, r);
            if (ss[r] != ss[c]) u = zc<Z>(-u.x, -u.y);
            x = zc<Z>(R(0.5) * (x.x + u.x), R(0.5) * (x.y + u.y));
        }
        st[e] = x;
    }
    if (tid == 0 && info_out) info_out[f] = info;
    };
    if (!gscratch)
        rest(tmp + 96, nullptr);
    else if (lds_big)
        rest(tmp + 96, gscratch + (size_t)f * gmax * (2 * gmax + 1));
    else
        rest(gscratch + (size_t)f * gmax * (2 * gmax + 1), nullptr);
}

// cap^-1 of the large flips from the factors lr_prep_kernel left in the global slot: one
// workgroup per (frequency, 16 columns), the factors staged in LDS.  Thread (tr, tc) keeps rows
// tr + 16 m (m < 6) of column 16 blockIdx.y + tc in registers; every entry goes through the
// operations of lu_inverse_cols in the same order (a step also gives the rows it does not touch
// x - 0 x_j, which leaves x as it is for a finite x_j).  At each step the owners of row j publish
// it through LDS (two alternating rows of 16, one barrier per step).  X goes to the slot after
// the factors.
template <class Z>
__global__ __launch_bounds__(256) void lr_prep_inv_kernel(Dims d, int k, const Z* __restrict__ acc, int acc_stride,
                                                          Z* gscratch, const int32_t* dyn) {
    constexpr int T = 6, G = 16;
    if (dyn) {
        if (!dyn[kDynActive]) return;
        k = dyn[kDynK];
    }
    const int f = blockIdx.x, c0 = G * blockIdx.y;
    if (c0 >= k) return;
    extern __shared__ __align__(16) unsigned char smraw[];
    Z* L = (Z*)smraw;                 // [k][k] factors
    Z* rowbuf = L + (size_t)k * k;    // [2][16]
    int* piv = (int*)(rowbuf + 2 * G);
    const int tid = threadIdx.x, tc = tid % G, tr = tid / G;
    const Z* fac = gscratch + (size_t)f * d.gmax * (2 * d.gmax + 1);
    const int* pg = (const int*)(acc + (size_t)f * acc_stride + (size_t)k * k);
    for (int e = tid; e < k * k; e += 256) L[e] = fac[e];
    for (int c = tid; c < k; c += 256) piv[c] = pg[c];
    __syncthreads();
    const int c = c0 + tc;
    // the 1 of e_c after the row interchanges piv[0..k-1]
    int pos = c;
    for (int j = 0; j < k; j++) {
        const int p = piv[j];
        if (p != j) {
            if (pos == j) pos = p;
            else if (pos == p) pos = j;
        }
    }
    Z x[T];
#pragma unroll
    for (int m = 0; m < T; m++) x[m] = zc<Z>(tr + G * m == pos ? 1 : 0, 0);
    int buf = 0;
    // unit lower triangle, j increasing: rows i > j
    for (int j = 0; j < k; j++) {
        const int mj = j / G;
        if (tr == j - G * mj)
#pragma unroll
            for (int m = 0; m < T; m++)
                if (m == mj) rowbuf[buf * G + tc] = x[m];
        __syncthreads();
        const Z xj = rowbuf[buf * G + tc];
        Z l[T];
#pragma unroll
        for (int m = 0; m < T; m++) l[m] = L[min(tr + G * m, k - 1) * k + j];
#pragma unroll
        for (int m = 0; m < T; m++)
            if (G * m + G - 1 > j) {
                const Z lm = tr + G * m > j ? l[m] : zc<Z>(0, 0);
                x[m] = zfms(lm, xj, x[m]);
            }
        buf ^= 1;
    }
    // upper triangle, j decreasing: x_j /= u_jj, then rows i < j
    for (int j = k - 1; j >= 0; j--) {
        const int mj = j / G;
        if (tr == j - G * mj) {
            const Z dj = L[j * k + j];
#pragma unroll
            for (int m = 0; m < T; m++)
                if (m == mj) {
                    x[m] = zdiv(x[m], dj);
                    rowbuf[buf * G + tc] = x[m];
                }
        }
        __syncthreads();
        const Z xj = rowbuf[buf * G + tc];
        Z l[T];
#pragma unroll
        for (int m = 0; m < T; m++) l[m] = L[min(tr + G * m, k - 1) * k + j];
#pragma unroll
        for (int m = 0; m < T; m++)
            if (G * m < j) {
                const Z lm = tr + G * m < j ? l[m] : zc<Z>(0, 0);
                x[m] = zfms(lm, xj, x[m]);
            }
        buf ^= 1;
    }
    Z* X = gscratch + (size_t)f * d.gmax * (2 * d.gmax + 1) + (size_t)k * k;
    if (c < k)
#pragma unroll
        for (int m = 0; m < T; m++)
            if (tr + G * m < k) X[(tr + G * m) * k + c] = x[m];
}

// Lambda of the large flips from lr_prep_inv_kernel's cap^-1, as lr_prep_kernel forms it for the
// smaller ones: diag(s) cap^-1, stored exactly symmetric for symmetric W.
template <class Z>
__global__ __launch_bounds__(256) void lr_prep_lam_kernel(Dims d, int k, const int32_t* __restrict__ ss, Z* acc,
                                                          int acc_stride, const Z* gscratch, const int32_t* dyn) {
    typedef typename zreal<Z>::type R;
    if (dyn) {
        if (!dyn[kDynActive]) return;
        k = dyn[kDynK];
    }
    const int f = blockIdx.x;
    const Z* X = gscratch + (size_t)f * d.gmax * (2 * d.gmax + 1) + (size_t)k * k;
    Z* st = acc + (size_t)f * acc_stride;
    for (int e = threadIdx.x; e < k * k; e += blockDim.x) {
        const int r = e / k, c = e - r * k;
        Z x = X[r * k + c];
        if (d.sym) {
            Z u = X[c * k + r];
            if (ss[r] != ss[c]) u = zc<Z>(-u.x, -u.y);
            x = zc<Z>(R(0.5) * (x.x + u.x), R(0.5) * (x.y + u.y));
        }
        st[e] = x;
    }
}

template <class Z>
void launch_lr_accept_prep(hipStream_t st, const Dims& d, const DevVecsT<Z>& v, int g, const int32_t* t,
                           const int32_t* s, const int32_t* lp, int k, Z* acc, int acc_stride,
                           int32_t* info, Z* gscratch, const int32_t* dyn, const int32_t* g_dev) {
    size_t sm = 100 * sizeof(int) + 192 * sizeof(Z) + lds_matrix_bytes(d.gmax, gscratch);
    int lds_big = 0;
    size_t sm_inv = 0;
    if (gscratch && d.gmax <= 96) {
        // the k x (k+1) LU in LDS, then the inverse on (frequency, 16-column) workgroups with the
        // k x k factors in LDS, when the device's per-workgroup LDS takes them
        const size_t need = sm + (size_t)d.gmax * (d.gmax + 1) * sizeof(Z);
        sm_inv = (size_t)d.gmax * d.gmax * sizeof(Z) + 32 * sizeof(Z) + 96 * sizeof(int);
        if (need <= lds_per_workgroup() && sm_inv <= lds_per_workgroup()) {
            (void)hipFuncSetAttribute((const void*)lr_prep_kernel<Z>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)need);
            (void)hipFuncSetAttribute((const void*)lr_prep_inv_kernel<Z>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm_inv);
            sm = need;
            lds_big = 1;
        }
    }
    hipLaunchKernelGGL(lr_prep_kernel<Z>, dim3(d.nf), dim3(256), sm, st, d, v, g, t, s, lp, k, acc,
                       acc_stride, info, gscratch, dyn, g_dev, lds_big);
    if (lds_big) {
        hipLaunchKernelGGL(lr_prep_inv_kernel<Z>, dim3(d.nf, (d.gmax + 15) / 16), dim3(256), sm_inv, st, d, k,
                           (const Z*)acc, acc_stride, gscratch, dyn);
        hipLaunchKernelGGL(lr_prep_lam_kernel<Z>, dim3(d.nf), dim3(256), 0, st, d, k, s, acc, acc_stride,
                           (const Z*)gscratch, dyn);
    }
}

// M[kk][n]: stored, or from Lam and A (see DevTabs); the sum runs in the order the accept's
// epilogue used when it stored M rows, with the sign folded into Lam.
template <class Z>
__device__ inline Z m_entry(const DevTabsT<Z>& tb, const Dims& d, int f, int kk, int n) {
    if (!d.lam) return tb.Mall[f][(size_t)kk * d.ldN + n];
    const Z* L = tb.Lam[f] + (size_t)kk * d.gmax;
    const Z* A = tb.Aall[f] + (size_t)n * d.Kmax + tb.fstart[kk];
    const int len = tb.flen[kk];
    Z s = zc<Z>(0, 0);
    for (int l = 0; l < len; l++) s = zfma(L[l], A[l], s);
    return s;
}

// Mc[K'][c] = M_all[K'][t_c] (K x k) and, unless Ar is null, Ar[c][K'] = A_all[t_c][K'] (k x K).
template <class Z>
__global__ void lr_gather_kernel(Dims d, DevTabsT<Z> tb, const int32_t* __restrict__ t, int k, int K, Z* Mc,
                                 Z* Ar, int small_stride, const int32_t* dyn) {
    const int f = blockIdx.y;
    if (dyn) {
        if (!dyn[kDynActive]) return;
        k = dyn[kDynK];
        K = dyn[kDynKK];
    }
    const Z* A = tb.Aall[f];
    Z* mc = Mc + (size_t)f * small_stride;
    Z* ar = Ar ? Ar + (size_t)f * small_stride : nullptr;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < K * k; e += gridDim.x * blockDim.x) {
        const int kk = e / k, c = e - kk * k;
        mc[e] = m_entry(tb, d, f, kk, t[c]);
        if (!Ar) continue;
        const int c2 = e / K, k2 = e - c2 * K;
        ar[e] = A[(size_t)t[c2] * d.Kmax + k2];
    }
}

// Mc[K'][c] = M[K'][t_c] with M rebuilt from Lam and A (factors kept without M): a workgroup
// takes 32 factor rows of one frequency; the A columns those rows use lie within gmax - 1 of
// the rows' range, so they are staged in LDS for every t_c in the same phase as the Lam rows.
// Same sums as m_entry.
constexpr int kGatherRows = 32;
template <class Z>
__global__ void __launch_bounds__(256) lr_gather_lam_kernel(Dims d, DevTabsT<Z> tb, const int32_t* __restrict__ t,
                                                           int k, int K, Z* Mc, int small_stride, const int32_t* dyn) {
    if (dyn) {
        if (!dyn[kDynActive]) return;
        k = dyn[kDynK];
        K = dyn[kDynKK];
    }
    extern __shared__ __align__(16) unsigned char smlraw[];
    Z* sml = (Z*)smlraw;
    const int f = blockIdx.y, tid = threadIdx.x, gmax = d.gmax;
    const int span = kGatherRows + 2 * (gmax - 1);  // A columns a workgroup can need
    Z* Ls = sml;                                     // [kGatherRows][gmax] Lam rows
    Z* As = sml + kGatherRows * gmax;                // [k][span] A[t_c][lo ..]
    __shared__ int s_fs[kGatherRows], s_len[kGatherRows];
    const Z* A = tb.Aall[f];
    const Z* Lam = tb.Lam[f];
    Z* mc = Mc + (size_t)f * small_stride;
    for (int r0 = blockIdx.x * kGatherRows; r0 < K; r0 += gridDim.x * kGatherRows) {
        const int nr = min(kGatherRows, K - r0);
        const int lo = max(0, r0 - (gmax - 1)), hi = min(K, r0 + nr + gmax - 1);
        __syncthreads();
        for (int e = tid; e < nr * gmax; e += blockDim.x) Ls[e] = Lam[(size_t)r0 * gmax + e];
        if (tid < nr) {
            s_fs[tid] = tb.fstart[r0 + tid];
            s_len[tid] = tb.flen[r0 + tid];
        }
        const int w = hi - lo;
        for (int e = tid; e < k * w; e += blockDim.x) {
            const int c = e / w, l = e - c * w;
            As[c * span + l] = A[(size_t)t[c] * d.Kmax + lo + l];
        }
        __syncthreads();
        for (int e = tid; e < nr * k; e += blockDim.x) {
            const int r = e / k, c = e - r * k;
            const int off = s_fs[r] - lo, len = s_len[r];
            Z acc = zc<Z>(0, 0);
            for (int l = 0; l < len; l++) acc = zfma(Ls[r * gmax + l], As[c * span + off + l], acc);
            mc[(size_t)(r0 + r) * k + c] = acc;
        }
    }
}

template <class Z>
void launch_lr_gather(hipStream_t st, const Dims& d, const DevTabsT<Z>& tb, const int32_t* t, int k, int K,
                      Z* Mc, Z* Ar, int small_stride, const int32_t* dyn) {
    if (!dyn && K == 0) return;
    if (d.lam) {
        const int rows = dyn ? d.Kmax : K;
        const int blocks = std::min(1024, (rows + kGatherRows - 1) / kGatherRows);
        const size_t sm = ((size_t)kGatherRows * d.gmax + (size_t)d.gmax * (kGatherRows + 2 * (d.gmax - 1))) * sizeof(Z);
        hipLaunchKernelGGL(lr_gather_lam_kernel<Z>, dim3(blocks, d.nf), dim3(256), sm, st, d, tb, t, k, K, Mc,
                           small_stride, dyn);
        return;
    }
    // grid-stride loop: with a device-resident K the grid is the largest one
    int blocks = dyn ? 1024 : min(1024, (K * k + 255) / 256);
    hipLaunchKernelGGL(lr_gather_kernel<Z>, dim3(blocks, d.nf), dim3(256), 0, st, d, tb, t, k, K, Mc, Ar,
                       small_stride, dyn);
}

template <class Z>
__global__ void lr_build_m_kernel(Dims d, DevTabsT<Z> tb, int f, int kk0, Z* out, int ld) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (n < d.N) out[(size_t)r * ld + n] = m_entry(tb, d, f, kk0 + r, n);
}

template <class Z>
void launch_lr_build_m(hipStream_t st, const Dims& d, const DevTabsT<Z>& tb, int f, int kk0, int nrows, Z* out, int ld) {
    hipLaunchKernelGGL(lr_build_m_kernel<Z>, dim3((d.N + 255) / 256, nrows), dim3(256), 0, st, d, tb, f, kk0, out, ld);
}

// Skinny products over the deferred factors (k <= 96 columns), on FP64 MFMA
// (v_mfma_f64_16x16x4f64):
//   colblk: A_all[n][K+c] = W0[n][t_c] - sum_K' A_all[n][K'] Mc[K'][c]
//   rowblk: Bt[n][c]      = W0[t_c][n] - sum_K' Ar[c][K'] M_all[K'][n]
// The complex product runs as a real one: out (n x 2k interleaved re/im) = A_real (n x 2K,
// the interleaved storage of A) x B_real (2K x 2k) with, for b = B[K'][c],
//   B_real[2K'][2c] = Re b   B_real[2K'][2c+1] = Im b   B_real[2K'+1][2c] = -Im b   B_real[2K'+1][2c+1] = Re b.
// B is staged in LDS already in that operand form: Bp[kk][j] holds (B_real[2kk][j],
// B_real[2kk+1][j]) for real column j, so one 16-byte LDS read feeds the two MFMAs of a
// complex inner index with no VALU work.
// One wave owns RT x 16 output rows and all CT 16-column tiles.  Lane l feeds inner index
// 2*(l/16)+j of each 8-entry sub-chunk (two consecutive complex A entries per lane).
// Fragment layouts (16x16x4 f64, measured on gfx950 with a lane probe): A[i][k] in lane
// i + 16k, B[k][j] in lane j + 16k, D[i][j] in lane j + 16(i % 4), element i / 4.
// Memory pipeline: the A entries of the next DEPTH sub-chunks are in flight in a register
// ring.  Loads go through global-address-space pointers (counted on vmcnt only), are
// branch-free (row clamped to N-1, inner index clamped to K-1: the clamped entries meet zero
// rows of the staged B), and each staged chunk of KC inner entries is straight-line code with a
// scheduling barrier after every refill, so the waits before each MFMA group are exact.  The
// double-buffered variants load the next staged chunk into registers during the current one
// (its L2 latency hidden behind the MFMAs) and need one barrier per chunk instead of two.
typedef double d4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const double gdbl;
typedef __attribute__((address_space(1))) double gdbl_w;

__device__ inline zd ldg(const zd* p, size_t i) {
    gdbl* q = (gdbl*)(p + i);
    zd r;
    r.x = q[0];
    r.y = q[1];
    return r;
}

// MODE 0: colblk, 1: rowblk, 2: the PLAIN accept's V C^-1 (pngfopt.hpp:293-322), a colblk-shaped
// product with A = C^-1 read column by column (A[n][i] = C^-1[i][n], n-th column contiguous),
// inner index i over all N rows, B = V^T (V k x N, ld k: the staged-B layout of colblk) and the
// result written as VC (k x N, ld k) without W0.
enum { kSkinnyCol = 0, kSkinnyRow = 1, kSkinnyVC = 2 };

template <int CT, int MODE, int KC, int DEPTH, int WAVES, bool DB, int RT>
__device__ __forceinline__ void skinny_mfma(Dims d, DevTabs tb, const int32_t* __restrict__ t, int k,
                                            int K, const zd* __restrict__ small, int small_stride, zd* Bt,
                                            size_t bt_stride) {
    constexpr bool ROW = MODE == kSkinnyRow;
    constexpr int BW = CT * 8;   // complex columns held (>= k)
    constexpr int NJ = CT * 16;  // real columns
    constexpr int KS = 8, NS = KC / KS, NT = WAVES * 64;
    constexpr int NBUF = DB ? 2 : 1;
    constexpr int PER = (KC * BW + NT - 1) / NT;  // staged B entries per thread and chunk
    static_assert(NS % DEPTH == 0, "ring slots line up across chunks");
    __shared__ zd Bp[NBUF][KC][NJ + 2];
    const int f = blockIdx.y;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lr = lane & 15, lq = lane >> 4;
    // RT row tiles of 16 per wave: each staged B operand feeds RT times as many MFMAs
    const int n0 = (blockIdx.x * WAVES + wave) * 16 * RT;
    int nc[RT];  // clamped rows for loads
#pragma unroll
    for (int r = 0; r < RT; r++) nc[r] = min(n0 + 16 * r + lr, d.N - 1);
    const zd* A = MODE == kSkinnyVC ? tb.Cinv[f] : tb.Aall[f];
    const int lda = MODE == kSkinnyVC ? d.ldN : d.Kmax;
    const zd* M = tb.Mall[f];
    const zd* W0 = tb.W0[f];
    const zd* sm = small + (size_t)f * small_stride;
    const int Kc = K - 1;
    d4v acc[RT][CT];
#pragma unroll
    for (int r = 0; r < RT; r++)
#pragma unroll
        for (int c = 0; c < CT; c++) acc[r][c] = d4v{0.0, 0.0, 0.0, 0.0};
    auto ld = [&](int kb, zd (&a)[RT][2]) {
#pragma unroll
        for (int r = 0; r < RT; r++)
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int Kp = min(kb + 2 * lq + j, Kc);
                a[r][j] = ROW ? ldg(M, (size_t)Kp * d.ldN + nc[r]) : ldg(A, (size_t)nc[r] * lda + Kp);
            }
    };
    auto body = [&](const zd (&a)[RT][2], int kloc, int bsel) {
#pragma unroll
        for (int j = 0; j < 2; j++) {
            zd m[CT];
#pragma unroll
            for (int c = 0; c < CT; c++) m[c] = Bp[bsel][kloc + 2 * lq + j][c * 16 + lr];
#pragma unroll
            for (int r = 0; r < RT; r++)
#pragma unroll
                for (int c = 0; c < CT; c++)
                    acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r][j].x, m[c].x, acc[r][c], 0, 0, 0);
#pragma unroll
            for (int r = 0; r < RT; r++)
#pragma unroll
                for (int c = 0; c < CT; c++)
                    acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r][j].y, m[c].y, acc[r][c], 0, 0, 0);
        }
    };
    // B chunk k0..k0+KC-1 (rows past K and columns past k are zero) into registers, then
    // into LDS in operand form
    auto bload = [&](int k0, zd (&r)[PER]) {
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int e = threadIdx.x + u * NT;
            const int kk = e / BW, c = e - kk * BW;
            r[u] = zmk(0.0, 0.0);
            if (e < KC * BW && k0 + kk < K && c < k)
                r[u] = ldg(sm, ROW ? (size_t)c * K + k0 + kk : (size_t)(k0 + kk) * k + c);
        }
    };
    auto bstore = [&](const zd (&r)[PER], int bsel) {
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int e = threadIdx.x + u * NT;
            if (e >= KC * BW) continue;
            const int kk = e / BW, c = e - kk * BW;
            Bp[bsel][kk][2 * c] = zmk(r[u].x, -r[u].y);
            Bp[bsel][kk][2 * c + 1] = zmk(r[u].y, r[u].x);
        }
    };
    zd buf[DEPTH][RT][2];
    if (K > 0) {
#pragma unroll
        for (int q = 0; q < DEPTH; q++) ld(q * KS, buf[q]);
    }
    if (!DB) {
        for (int k0 = 0; k0 < K; k0 += KC) {
            __syncthreads();
            {
                zd r[PER];
                bload(k0, r);
                bstore(r, 0);
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < NS; i++) {
                body(buf[i % DEPTH], i * KS, 0);
                ld(k0 + (i + DEPTH) * KS, buf[i % DEPTH]);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    } else if (K > 0) {
        // double-buffered stage: the next chunk is loaded into registers while this one is
        // multiplied and written to the other buffer afterwards; one barrier per chunk
        {
            zd r[PER];
            bload(0, r);
            bstore(r, 0);
        }
        __syncthreads();
        int cur = 0;
        for (int k0 = 0; k0 < K; k0 += KC) {
            const bool more = k0 + KC < K;
            zd r[PER];
            if (more) bload(k0 + KC, r);
#pragma unroll
            for (int i = 0; i < NS; i++) {
                body(buf[i % DEPTH], i * KS, cur);
                ld(k0 + (i + DEPTH) * KS, buf[i % DEPTH]);
                __builtin_amdgcn_sched_barrier(0);
            }
            if (more) bstore(r, cur ^ 1);
            __syncthreads();
            cur ^= 1;
        }
    }
    // Epilogue: out = W0 entry - acc.  The t entries of this lane's columns are read once, and
    // the W0 entries of a group of EG column tiles are all loaded before any of their stores:
    // the compiler cannot tell that a store to A or Bt leaves t and W0 alone, so in program
    // order load-store-load-store every element waited for two dependent global loads.
    // (unconditional loads of in-range entries, so that they issue together; the change set
    // array always holds at least one entry)
    int tcol[CT];
#pragma unroll
    for (int c = 0; c < CT; c++) {
        const int cc = max(0, min((c * 16 + lr) >> 1, k - 1));
        tcol[c] = MODE != kSkinnyVC ? t[cc] : 0;
    }
    constexpr int EG = CT < 3 ? CT : 3;
#pragma unroll
    for (int c0 = 0; c0 < CT; c0 += EG) {
        double wv[RT][EG][4];
#pragma unroll
        for (int r = 0; r < RT; r++)
#pragma unroll
            for (int e = 0; e < EG; e++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int c = c0 + e;
                    const int col = c * 16 + lr, cc = col >> 1, q = col & 1;
                    const int row = n0 + 16 * r + lq + 4 * v;
                    wv[r][e][v] = 0.0;
                    if (MODE != kSkinnyVC && cc < k && row < d.N) {
                        // global-space pointers: the tables' entries would otherwise be flat accesses
                        gdbl* w = ROW ? (gdbl*)(W0 + (size_t)tcol[c] * d.ldN + row)
                                      : (gdbl*)(W0 + (size_t)row * d.ldN + tcol[c]);
                        wv[r][e][v] = w[q];
                    }
                }
#pragma unroll
        for (int r = 0; r < RT; r++)
#pragma unroll
            for (int e = 0; e < EG; e++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int c = c0 + e;
                    const int col = c * 16 + lr, cc = col >> 1, q = col & 1;
                    const int row = n0 + 16 * r + lq + 4 * v;
                    if (cc >= k || row >= d.N) continue;
                    const double a = acc[r][c][v];
                    if (MODE == kSkinnyVC)
                        ((gdbl_w*)Bt)[((size_t)f * bt_stride + (size_t)row * k + cc) * 2 + q] = a;
                    else if (!ROW)
                        ((gdbl_w*)A)[((size_t)row * d.Kmax + K + cc) * 2 + q] = wv[r][e][v] - a;
                    else
                        ((gdbl_w*)Bt)[((size_t)f * bt_stride + (size_t)row * d.gmax + cc) * 2 + q] = wv[r][e][v] - a;
                }
    }
}

template <int CT, bool ROW, int KC, int DEPTH, int WAVES, bool DB, int RT>
__global__ void __launch_bounds__(WAVES * 64) lr_skinny_mfma_kernel(Dims d, DevTabs tb, const int32_t* __restrict__ t,
                                                                    int k, int K, const zd* __restrict__ small,
                                                                    int small_stride, zd* Bt, size_t bt_stride,
                                                                    const int32_t* dyn) {
    if (dyn) {
        if (!dyn[kDynActive]) return;
        k = dyn[kDynK];
        K = dyn[kDynKK];
    }
    skinny_mfma<CT, ROW ? kSkinnyRow : kSkinnyCol, KC, DEPTH, WAVES, DB, RT>(d, tb, t, k, K, small, small_stride, Bt,
                                                                             bt_stride);
}

// VC (k x N, ld k, per frequency at f * vc_stride) = V C^-1 with V (k x N, ld k, at f * v_stride)
template <int CT, int KC, int DEPTH, int WAVES, bool DB, int RT>
__global__ void __launch_bounds__(WAVES * 64) plain_vcinv_mfma_kernel(Dims d, DevTabs tb, int k,
                                                                      const zd* __restrict__ V, int v_stride, zd* VC,
                                                                      size_t vc_stride) {
    skinny_mfma<CT, kSkinnyVC, KC, DEPTH, WAVES, DB, RT>(d, tb, nullptr, k, d.N, V, v_stride, VC, vc_stride);
}

// Tile-shape variants offered to the autotuner, per column-tile count CT: (KC inner entries
// per staged chunk, DEPTH sub-chunks in flight, waves per workgroup, double-buffered stage,
// RT 16-row tiles per wave).
// All of them issue the MFMAs of an output tile in the same inner-index order into one
// accumulator, so every variant gives bitwise the same result and the autotuner's choice
// cannot change a DBS run.
template <int CT, bool ROW, int KC, int DEPTH, int WAVES, bool DB, int RT>
static void launch_variant(hipStream_t st, const Dims& d, const DevTabs& tb, const int32_t* t, int k, int K,
                           const zd* small, int small_stride, zd* Bt, size_t bt_stride, const int32_t* dyn) {
    const int rows = 16 * WAVES * RT;
    dim3 grid((d.N + rows - 1) / rows, d.nf);
    hipLaunchKernelGGL((lr_skinny_mfma_kernel<CT, ROW, KC, DEPTH, WAVES, DB, RT>), grid, dim3(WAVES * 64), 0, st, d,
                       tb, t, k, K, small, small_stride, Bt, bt_stride, dyn);
}

template <bool ROW>
static void launch_skinny_mfma(hipStream_t st, const Dims& d, const DevTabs& tb, const int32_t* t, int k, int K,
                               const zd* small, int small_stride, zd* Bt, size_t bt_stride, int variant,
                               const int32_t* dyn) {
#define PNGF_SKV(CT, KC, DEPTH, W, DB, RT) \
    launch_variant<CT, ROW, KC, DEPTH, W, DB, RT>(st, d, tb, t, k, K, small, small_stride, Bt, bt_stride, dyn)
    if (k <= 24) {
        switch (variant) {
            case 1: PNGF_SKV(3, 32, 2, 4, true, 2); break;
            case 2: PNGF_SKV(3, 64, 2, 8, false, 1); break;
            case 3: PNGF_SKV(3, 16, 2, 4, true, 2); break;
            case 4: PNGF_SKV(3, 16, 2, 8, true, 1); break;
            case 5: PNGF_SKV(3, 32, 2, 4, true, 1); break;
            default: PNGF_SKV(3, 32, 2, 8, true, 1); break;
        }
    } else if (k <= 48) {
        if (variant % 2) PNGF_SKV(6, 16, 2, 8, false, 1);
        else PNGF_SKV(6, 16, 2, 8, true, 1);
    } else {
        if (variant % 2) PNGF_SKV(12, 16, 2, 4, false, 1);
        else PNGF_SKV(12, 16, 2, 8, false, 1);
    }
#undef PNGF_SKV
}

int skinny_variants(int k, bool c64) { return c64 ? (k <= 24 ? 4 : 1) : (k <= 24 ? 6 : 2); }

template <int CT, int KC, int DEPTH, int WAVES, bool DB, int RT>
static void launch_vc_variant(hipStream_t st, const Dims& d, const DevTabs& tb, int k, const zd* V, int v_stride,
                              zd* VC, size_t vc_stride) {
    const int rows = 16 * WAVES * RT;
    dim3 grid((d.N + rows - 1) / rows, d.nf);
    hipLaunchKernelGGL((plain_vcinv_mfma_kernel<CT, KC, DEPTH, WAVES, DB, RT>), grid, dim3(WAVES * 64), 0, st, d, tb,
                       k, V, v_stride, VC, vc_stride);
}

void launch_plain_vcinv(hipStream_t st, const Dims& d, const DevTabs& tb, int k, const zd* V, size_t v_stride,
                        zd* VC, size_t vc_stride, int variant) {
#define PNGF_VCV(CT, KC, DEPTH, W, DB, RT) \
    launch_vc_variant<CT, KC, DEPTH, W, DB, RT>(st, d, tb, k, V, (int)v_stride, VC, vc_stride)
    if (k <= 24) {
        switch (variant) {
            case 1: PNGF_VCV(3, 32, 2, 4, true, 2); break;
            case 2: PNGF_VCV(3, 64, 2, 8, false, 1); break;
            case 3: PNGF_VCV(3, 16, 2, 4, true, 2); break;
            case 5: PNGF_VCV(3, 32, 2, 4, true, 1); break;
            case 0: PNGF_VCV(3, 32, 2, 8, true, 1); break;
            default: PNGF_VCV(3, 16, 2, 8, true, 1); break;
        }
    } else if (k <= 48) {
        if (variant % 2) PNGF_VCV(6, 16, 2, 8, false, 1);
        else PNGF_VCV(6, 16, 2, 8, true, 1);
    } else {
        if (variant % 2) PNGF_VCV(12, 16, 2, 4, false, 1);
        else PNGF_VCV(12, 16, 2, 8, false, 1);
    }
#undef PNGF_VCV
}

// C^-1 -= CU M: the rank-k update of the PLAIN accept (pngfopt.hpp:335-358, zgemm with alpha = -1)
// with CU = C^-1[:, t] diag(s) (N x k, column c at c * ldN) and M = cap^-1 V C^-1 (k x N, row c at
// c * ldN), on FP64 MFMA (v_mfma_f64_16x16x4f64).  An output tile's 16 rows are 16 columns j of
// C^-1 and its 16 real columns are 8 complex rows i (re/im interleaved, the storage order), so
//   C^-1[i][j] -= sum_c M[c][j] CU[i][c]
// runs as the real product A_real (16 x 2k) x B_real (2k x 16) with A_real[j][2c + p] = component p
// of -M[c][j] and, for b = CU[i][c], B_real[2c][2i] = Re b, B_real[2c][2i+1] = Im b,
// B_real[2c+1][2i] = -Im b, B_real[2c+1][2i+1] = Re b.  The accumulator starts as the C^-1 tile,
// so the MFMAs leave C^-1 - CU M in it.  Lane l = lr + 16 lq: A entry (column lr, inner 4s + lq),
// B entry (inner 4s + lq, real column lr), D element v = column j0 + 4v + lq, complex row
// lr / 2, component lr % 2: each group of 16 lanes covers 128 contiguous bytes of one column.
// A workgroup of W waves owns 16 W columns and a range of rows, walked 8 TT rows at a time.  The
// step's CU rows are staged in LDS (double-buffered: one barrier per step; the row pitch KB + 1
// keeps the 16 lanes of a read on distinct banks), the A entries stay in registers for the
// whole range, and the next step's C^-1 tiles are loaded while the current ones are multiplied
// (TT output tiles per wave and step: fewer for large k, whose A entries take more registers).
template <int KB, int W, int TT>
__global__ void __launch_bounds__(W * 64) plain_rankk_kernel(Dims d, DevTabs tb, const zd* __restrict__ CU,
                                                             size_t cu_stride, const zd* __restrict__ M,
                                                             size_t m_stride, int k, int rows_per_wg) {
    constexpr int NS = KB / 2;  // MFMA steps of 4 real inner entries (2 complex)
    constexpr int RS = 8 * TT, KP = KB + 1, NT = W * 64;
    constexpr int PER = (RS * KB + NT - 1) / NT;
    __shared__ zd Bs[2][RS][KP];
    const int f = blockIdx.z;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lr = lane & 15, lq = lane >> 4;
    const int j0 = (blockIdx.x * W + wave) * 16;
    const int ib = blockIdx.y * rows_per_wg, ie = min(d.N, ib + rows_per_wg);
    const zd* CUf = CU + (size_t)f * cu_stride;
    const zd* Mf = M + (size_t)f * m_stride;
    gdbl_w* C = (gdbl_w*)tb.Cinv[f];
    double a[NS];
    {
        const int j = min(j0 + lr, d.N - 1);
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int c = 2 * s + (lq >> 1);
            double v = 0.0;
            if (c < k) {
                const zd m = ldg(Mf, (size_t)c * d.ldN + j);
                v = (lq & 1) ? m.y : m.x;
            }
            a[s] = -v;
        }
    }
    // B entry of this lane: component (p ^ q) of CU[i][c], negated for p = 1, q = 0
    const int bcomp = (lq ^ lr) & 1;
    const int bsign = ((lq & 1) && !(lr & 1)) ? (int)0x80000000 : 0;
    auto stage_load = [&](int i0, zd (&r)[PER]) {
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int e = threadIdx.x + u * NT;
            const int rr = e % RS, c = e / RS;  // consecutive threads: consecutive rows of one CU column
            r[u] = zmk(0.0, 0.0);
            if (e < RS * KB && c < k && i0 + rr < ie) r[u] = ldg(CUf, (size_t)c * d.ldN + i0 + rr);
        }
    };
    auto stage_store = [&](const zd (&r)[PER], int buf) {
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int e = threadIdx.x + u * NT;
            if (e < RS * KB) Bs[buf][e % RS][e / RS] = r[u];
        }
    };
    // C^-1 entries of the TT tiles x 4 elements of this lane for the step at row i0
    auto tile_index = [&](int i0, int tt, int v, bool& ok) {
        const int row = i0 + 8 * tt + (lr >> 1), col = j0 + 4 * v + lq;
        ok = row < ie && col < d.N;
        return ((size_t)col * d.ldN + row) * 2 + (lr & 1);
    };
    auto tile_load = [&](int i0, d4v (&x)[TT]) {
#pragma unroll
        for (int tt = 0; tt < TT; tt++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                bool ok;
                const size_t idx = tile_index(i0, tt, v, ok);
                x[tt][v] = ok ? C[idx] : 0.0;
            }
    };
    d4v cur[TT];
    tile_load(ib, cur);
    {
        zd r[PER];
        stage_load(ib, r);
        stage_store(r, 0);
    }
    __syncthreads();
    int buf = 0;
    for (int i0 = ib; i0 < ie; i0 += RS) {
        const bool more = i0 + RS < ie;
        d4v nxt[TT];
        zd r[PER];
        if (more) {
            tile_load(i0 + RS, nxt);
            stage_load(i0 + RS, r);
        }
        // B entries one inner step ahead of the MFMAs that take them
        const double* Bd = (const double*)&Bs[buf][0][0] + ((lr >> 1) * KP + (lq >> 1)) * 2 + bcomp;
        double bn[TT];
#pragma unroll
        for (int tt = 0; tt < TT; tt++) bn[tt] = Bd[8 * tt * KP * 2];
#pragma unroll
        for (int s = 0; s < NS; s++) {
            double bc[TT];
#pragma unroll
            for (int tt = 0; tt < TT; tt++) bc[tt] = bn[tt];
            if (s + 1 < NS)
#pragma unroll
                for (int tt = 0; tt < TT; tt++) bn[tt] = Bd[(8 * tt * KP + 2 * (s + 1)) * 2];
#pragma unroll
            for (int tt = 0; tt < TT; tt++) {
                const double b = __hiloint2double(__double2hiint(bc[tt]) ^ bsign, __double2loint(bc[tt]));
                cur[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b, cur[tt], 0, 0, 0);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int tt = 0; tt < TT; tt++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                bool ok;
                const size_t idx = tile_index(i0, tt, v, ok);
                if (ok) C[idx] = cur[tt][v];
            }
        if (more) stage_store(r, buf ^ 1);
        __syncthreads();
        buf ^= 1;
        if (more)
#pragma unroll
            for (int tt = 0; tt < TT; tt++) cur[tt] = nxt[tt];
    }
}

void launch_plain_rankk(hipStream_t st, const Dims& d, const DevTabs& tb, const zd* CU, size_t cu_stride,
                        const zd* M, size_t m_stride, int k) {
    constexpr int W = 8;
    const int rows_per_wg = 1024;
    dim3 grid((d.N + 16 * W - 1) / (16 * W), (d.N + rows_per_wg - 1) / rows_per_wg, d.nf);
#define PNGF_RKV(KB, TT)                                                                                          \
    hipLaunchKernelGGL((plain_rankk_kernel<KB, W, TT>), grid, dim3(W * 64), 0, st, d, tb, CU, cu_stride, M, m_stride, \
                       k, rows_per_wg)
    if (k <= 24) PNGF_RKV(24, 4);
    else if (k <= 48) PNGF_RKV(48, 4);
    else PNGF_RKV(96, 2);
#undef PNGF_RKV
}

// PLAIN score, split-N pass on FP64 MFMA (reference pngfopt.hpp:420-480, fused as in
// plain_partial_kernel): for the components n of one split,
//   X = [V ; Gobj] [C^-1[:, t] , Snew]     ((k + ne) x (k + 1) complex),
// V = (G - I)[t, :] (rows of the row-major G copy), Snew = S - C^-1[:, t] (s o Einc_t) formed on
// the fly.  The complex product runs as the real one X_real = A_real B_real with
// A_real[r][2n + p] = component p of A[r][n] (the storage order) and, for b = B[n][c],
// B_real[2n][2c] = Re b, B_real[2n][2c+1] = Im b, B_real[2n+1][2c] = -Im b,
// B_real[2n+1][2c+1] = Re b, in RT x CT tiles of v_mfma_f64_16x16x4f64 (RT = CT / 2 = 2 holds
// k + ne <= 32 rows and k + 1 <= 32 columns).  In a group of 8 components n0 .. n0 + 7 lane
// (lr, lq) supplies inner entries (n0 + 2 lq + e, p) at step 2e + p.  Its A operands of the group
// are one 32-byte piece of a row of G or Gobj; its B operand is the entry of column
// c = 8 ct + lr / 2 of C^-1 at component n0 + 2 lq + (lr & 1), the pair's other entry coming from
// lane lr ^ 1 (DPP).  All of it is loaded straight from HBM (each 16-lane quarter reads 128
// contiguous bytes per row or column), the next group's pieces in flight while the current one
// is multiplied.  The Snew column: each lane's share of sum_c C^-1[n][t_c] es_c at its component
// is summed over the column tiles, then over the 8 lanes of the same parity by an xor butterfly
// (every lane ends with the same bits), and S[n] minus it takes the operand slot of column k.  Each wave owns a quarter of the split; the four wave sums are
// added in LDS in wave order and written in plain_partial_kernel's layout.
// x from lane lane ^ 1 (DPP quad_perm [1, 0, 3, 2])
__device__ inline double dpp_swap1(double x) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), 0xB1, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), 0xB1, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

template <int RT, int CT>
__global__ void __launch_bounds__(256, 2) plain_partial_mfma_kernel(Dims d, DevTabs tb, DevVecs v, GroupTab gt,
        const int32_t* __restrict__ tile, const int8_t* __restrict__ P, const int32_t* __restrict__ owners,
        const int32_t* __restrict__ batch, int nsplit, zd* __restrict__ partial) {
    constexpr int NW = 4, NTL = RT * CT;
    __shared__ int t[96], sg[96], lp[96];
    __shared__ zd es[CT * 8];
    __shared__ int s_k;
    __shared__ double red[NTL][4][64];
    const int item = blockIdx.x, split = blockIdx.y;
    const int b = item / d.nf, f = item - b * d.nf;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lr = lane & 15, lq = lane >> 4, q = lr & 1;
    if (tid == 0) s_k = change_set(gt, batch[b], tile, P, owners, t, sg, lp);
    __syncthreads();
    const int k = s_k, ne = d.ne;
    if (tid < CT * 8) {
        zd e = zmk(0.0, 0.0);
        if (tid < k) {
            e = ldg(v.Einc + (size_t)f * d.ldN, t[tid]);
            if (sg[tid] < 0) e = zmk(-e.x, -e.y);
        }
        es[tid] = e;
    }
    __syncthreads();
    const zd* Gr = tb.Grow[f];
    const zd* Ci = tb.Cinv[f];
    const zd* Sv = v.S + (size_t)f * d.ldN;
    // operand rows and columns of this lane; absent ones read row t[0] (the same cache lines) and
    // are zeroed
    const int t0 = k > 0 ? t[0] : 0;  // a valid row for absent operands (k = 0: no change)
    const zd* arow[RT];
    int adiag[RT];
    bool aon[RT];
#pragma unroll
    for (int rt = 0; rt < RT; rt++) {
        const int r = 16 * rt + lr;
        aon[rt] = r < k + ne;
        adiag[rt] = r < k ? t[r] : -1;
        arow[rt] = r < k ? Gr + (size_t)t[r] * d.ldN
                         : (r < k + ne ? v.Gobj + ((size_t)f * ne + (r - k)) * d.ldN : Gr + (size_t)t0 * d.ldN);
    }
    const zd* bcol[CT];
    bool bon[CT];
    zd esc[CT];
#pragma unroll
    for (int ct = 0; ct < CT; ct++) {
        const int c = 8 * ct + (lr >> 1);
        bon[ct] = c < k;
        bcol[ct] = Ci + (size_t)(c < k ? t[c] : t0) * d.ldN;
        esc[ct] = es[c];
    }
    const int cts = (2 * k) >> 4;                     // column tile of Snew
    const bool snew_lane = (lr >> 1) == (k & 7);      // lanes of Snew's two real columns
    // this wave's components: a span of a multiple of 8
    const int nsw = nsplit * NW;
    const int span = (((d.N + nsw - 1) / nsw) + 7) & ~7;
    const int n_begin = (split * NW + wave) * span, n_end = min(d.N, n_begin + span);
    const int nlast = (d.N - 1) & ~1;                 // last even component: a pair load stays in ldN
    d4v acc[RT][CT];
#pragma unroll
    for (int rt = 0; rt < RT; rt++)
#pragma unroll
        for (int ct = 0; ct < CT; ct++) acc[rt][ct] = d4v{0.0, 0.0, 0.0, 0.0};
    // operands of one group of 8 components: A entries at both components of the lane's pair,
    // B and S at the component of the lane's parity (n0 + 2 lq + q); the other parity's B entry
    // comes from lane lr ^ 1 (same column) in the body
    struct Grp {
        zd a[RT][2], b[CT], s;
    };
    auto load = [&](int n0, Grp& g) {
        const int na = n0 + 2 * lq;
        const int ia = min(na, nlast);
        const bool va = na < n_end, vb = na + 1 < n_end, vq = q ? vb : va;
#pragma unroll
        for (int rt = 0; rt < RT; rt++) {
            zd x0 = ldg(arow[rt], ia), x1 = ldg(arow[rt], ia + 1);
            if (ia == adiag[rt]) x0.x -= 1.0;
            if (ia + 1 == adiag[rt]) x1.x -= 1.0;
            g.a[rt][0] = (aon[rt] && va) ? x0 : zmk(0.0, 0.0);
            g.a[rt][1] = (aon[rt] && vb) ? x1 : zmk(0.0, 0.0);
        }
#pragma unroll
        for (int ct = 0; ct < CT; ct++) {
            const zd x = ldg(bcol[ct], ia + q);
            g.b[ct] = (bon[ct] && vq) ? x : zmk(0.0, 0.0);
        }
        const zd sx = ldg(Sv, ia + q);
        g.s = vq ? sx : zmk(0.0, 0.0);
    };
    auto body = [&](Grp& g) {
        // Snew = S - sum_c C^-1[n][t_c] es_c at this lane's component n = n0 + 2 lq + q
        zd part = zmk(0.0, 0.0);
#pragma unroll
        for (int ct = 0; ct < CT; ct++) part = zfma(g.b[ct], esc[ct], part);
#pragma unroll
        for (int off = 2; off < 16; off <<= 1) {
            part.x += __shfl_xor(part.x, off, 64);
            part.y += __shfl_xor(part.y, off, 64);
        }
        const zd sn = zsub(g.s, part);
#pragma unroll
        for (int ct = 0; ct < CT; ct++)
            if (ct == cts && snew_lane) g.b[ct] = sn;
        double bb[CT][4];
#pragma unroll
        for (int ct = 0; ct < CT; ct++) {
            // entries at the pair's two components: this lane's and lane lr ^ 1's
            const zd own = g.b[ct];
            const zd oth = zmk(dpp_swap1(own.x), dpp_swap1(own.y));
            const zd x0 = q ? oth : own, x1 = q ? own : oth;
            // B_real entries of real column 2c + q at inner (n, p): p = 0 -> (Re, Im)[q], p = 1 ->
            // (-Im, Re)[q]
            bb[ct][0] = q ? x0.y : x0.x;
            bb[ct][1] = q ? x0.x : -x0.y;
            bb[ct][2] = q ? x1.y : x1.x;
            bb[ct][3] = q ? x1.x : -x1.y;
        }
        // steps (n_a, Re), (n_a, Im), (n_b, Re), (n_b, Im), each over all RT x CT accumulators
#pragma unroll
        for (int st = 0; st < 4; st++)
#pragma unroll
            for (int rt = 0; rt < RT; rt++) {
                const zd ax = g.a[rt][st >> 1];
                const double av = (st & 1) ? ax.y : ax.x;
#pragma unroll
                for (int ct = 0; ct < CT; ct++)
                    acc[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bb[ct][st], acc[rt][ct], 0, 0, 0);
            }
    };
    if (n_begin < n_end) {
        Grp cur;
        load(n_begin, cur);
        for (int n0 = n_begin; n0 < n_end; n0 += 8) {
            Grp nxt;
            const bool more = n0 + 8 < n_end;
            if (more) load(n0 + 8, nxt);
            body(cur);
            __builtin_amdgcn_sched_barrier(0);
            if (more) cur = nxt;
        }
    }
    // the four wave sums in wave order
    for (int w = 0; w < NW; w++) {
        if (wave == w)
#pragma unroll
            for (int rt = 0; rt < RT; rt++)
#pragma unroll
                for (int ct = 0; ct < CT; ct++)
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        double& x = red[rt * CT + ct][u][lane];
                        x = w == 0 ? acc[rt][ct][u] : x + acc[rt][ct][u];
                    }
        __syncthreads();
    }
    // D[i][j] of a tile sits in lane j + 16 (i % 4), element i / 4
    const int pstride = (d.gmax + ne) * (d.gmax + 1);
    zd* out = partial + ((size_t)item * nsplit + split) * pstride;
    for (int e = tid; e < (k + ne) * (k + 1); e += 256) {
        const int r = e / (k + 1), c = e - r * (k + 1);
        const int i = r & 15, j = (2 * c) & 15;
        const int tl = (r >> 4) * CT + ((2 * c) >> 4), ln = j + 16 * (i & 3);
        out[r * (d.gmax + 1) + c] = zmk(red[tl][i >> 2][ln], red[tl][i >> 2][ln + 1]);
    }
}

bool launch_plain_partial_mfma(hipStream_t st, const Dims& d, const DevTabs& t, const DevVecs& v, GroupTab gt,
                               const int32_t* tile, const int8_t* P, const int32_t* owners, const int32_t* batch,
                               int nitems, int nsplit, zd* partial) {
    if (d.gmax + d.ne > 32 || d.gmax + 1 > 32) return false;
    hipLaunchKernelGGL((plain_partial_mfma_kernel<2, 4>), dim3(nitems, nsplit), dim3(256), 0, st, d, t, v, gt, tile,
                       P, owners, batch, nsplit, partial);
    return true;
}

void launch_lr_colblk(hipStream_t st, const Dims& d, const DevTabs& tb, const int32_t* t, int k, int K,
                      const zd* Mc, int small_stride, int variant, const int32_t* dyn) {
    launch_skinny_mfma<false>(st, d, tb, t, k, K, Mc, small_stride, nullptr, 0, variant, dyn);
}
void launch_lr_rowblk(hipStream_t st, const Dims& d, const DevTabs& tb, const int32_t* t, int k, int K,
                      const zd* Ar, int small_stride, zd* Bt, size_t bt_stride, int variant, const int32_t* dyn) {
    launch_skinny_mfma<true>(st, d, tb, t, k, K, Ar, small_stride, Bt, bt_stride, variant, dyn);
}

// The same skinny products in complex64 on FP32 MFMA (v_mfma_f32_16x16x4f32, FP32 accumulation):
// the complex product as the real one A_real x B_real of the FP64 kernel above, the staged B in
// operand form, the register ring of A entries and the double-buffered stage, at 8 bytes per
// complex.  Fragment layouts (16x16x4 f32, measured on gfx950 with a lane probe): A[i][k] in lane
// i + 16k, B[k][j] in lane j + 16k, D[i][j] in lane j + 16(i / 4), element i % 4 (the FP64
// instruction's D layout differs: lane j + 16(i % 4), element i / 4).
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const float gflt;
typedef __attribute__((address_space(1))) float gflt_w;

__device__ inline zf ldgf(const zf* p, size_t i) {
    gflt* q = (gflt*)(p + i);
    zf r;
    r.x = q[0];
    r.y = q[1];
    return r;
}

template <int CT, bool ROW, int KC, int DEPTH, int WAVES, bool DB, int RT>
__global__ void __launch_bounds__(WAVES * 64) lr_skinny_mfma32_kernel(Dims d, DevTabsT<zf> tb,
                                                                      const int32_t* __restrict__ t, int k, int K,
                                                                      const zf* __restrict__ small, int small_stride,
                                                                      zf* Bt, size_t bt_stride, const int32_t* dyn) {
    if (dyn) {
        if (!dyn[kDynActive]) return;
        k = dyn[kDynK];
        K = dyn[kDynKK];
    }
    constexpr int BW = CT * 8;   // complex columns held (>= k)
    constexpr int NJ = CT * 16;  // real columns
    constexpr int KS = 8, NS = KC / KS, NT = WAVES * 64;
    constexpr int NBUF = DB ? 2 : 1;
    constexpr int PER = (KC * BW + NT - 1) / NT;
    static_assert(NS % DEPTH == 0, "ring slots line up across chunks");
    __shared__ zf Bp[NBUF][KC][NJ + 2];
    const int f = blockIdx.y;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lr = lane & 15, lq = lane >> 4;
    const int n0 = (blockIdx.x * WAVES + wave) * 16 * RT;
    int nc[RT];
#pragma unroll
    for (int r = 0; r < RT; r++) nc[r] = min(n0 + 16 * r + lr, d.N - 1);
    const zf* A = tb.Aall[f];
    const zf* M = tb.Mall[f];
    const zf* W0 = tb.W0[f];
    const zf* sm = small + (size_t)f * small_stride;
    const int Kc = K - 1;
    f4v acc[RT][CT];
#pragma unroll
    for (int r = 0; r < RT; r++)
#pragma unroll
        for (int c = 0; c < CT; c++) acc[r][c] = f4v{0.f, 0.f, 0.f, 0.f};
    auto ld = [&](int kb, zf (&a)[RT][2]) {
#pragma unroll
        for (int r = 0; r < RT; r++)
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int Kp = min(kb + 2 * lq + j, Kc);
                a[r][j] = ROW ? ldgf(M, (size_t)Kp * d.ldN + nc[r]) : ldgf(A, (size_t)nc[r] * d.Kmax + Kp);
            }
    };
    auto body = [&](const zf (&a)[RT][2], int kloc, int bsel) {
#pragma unroll
        for (int j = 0; j < 2; j++) {
            zf m[CT];
#pragma unroll
            for (int c = 0; c < CT; c++) m[c] = Bp[bsel][kloc + 2 * lq + j][c * 16 + lr];
#pragma unroll
            for (int r = 0; r < RT; r++)
#pragma unroll
                for (int c = 0; c < CT; c++)
                    acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[r][j].x, m[c].x, acc[r][c], 0, 0, 0);
#pragma unroll
            for (int r = 0; r < RT; r++)
#pragma unroll
                for (int c = 0; c < CT; c++)
                    acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[r][j].y, m[c].y, acc[r][c], 0, 0, 0);
        }
    };
    auto bload = [&](int k0, zf (&r)[PER]) {
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int e = threadIdx.x + u * NT;
            const int kk = e / BW, c = e - kk * BW;
            r[u] = zmkf(0.f, 0.f);
            if (e < KC * BW && k0 + kk < K && c < k)
                r[u] = ldgf(sm, ROW ? (size_t)c * K + k0 + kk : (size_t)(k0 + kk) * k + c);
        }
    };
    auto bstore = [&](const zf (&r)[PER], int bsel) {
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int e = threadIdx.x + u * NT;
            if (e >= KC * BW) continue;
            const int kk = e / BW, c = e - kk * BW;
            Bp[bsel][kk][2 * c] = zmkf(r[u].x, -r[u].y);
            Bp[bsel][kk][2 * c + 1] = zmkf(r[u].y, r[u].x);
        }
    };
    zf buf[DEPTH][RT][2];
    if (K > 0) {
#pragma unroll
        for (int q = 0; q < DEPTH; q++) ld(q * KS, buf[q]);
    }
    if (!DB) {
        for (int k0 = 0; k0 < K; k0 += KC) {
            __syncthreads();
            {
                zf r[PER];
                bload(k0, r);
                bstore(r, 0);
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < NS; i++) {
                body(buf[i % DEPTH], i * KS, 0);
                ld(k0 + (i + DEPTH) * KS, buf[i % DEPTH]);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    } else if (K > 0) {
        {
            zf r[PER];
            bload(0, r);
            bstore(r, 0);
        }
        __syncthreads();
        int cur = 0;
        for (int k0 = 0; k0 < K; k0 += KC) {
            const bool more = k0 + KC < K;
            zf r[PER];
            if (more) bload(k0 + KC, r);
#pragma unroll
            for (int i = 0; i < NS; i++) {
                body(buf[i % DEPTH], i * KS, cur);
                ld(k0 + (i + DEPTH) * KS, buf[i % DEPTH]);
                __builtin_amdgcn_sched_barrier(0);
            }
            if (more) bstore(r, cur ^ 1);
            __syncthreads();
            cur ^= 1;
        }
    }
    // epilogue as in the FP64 kernel: t read once, a group's W0 entries loaded before its stores
    int tcol[CT];
#pragma unroll
    for (int c = 0; c < CT; c++) tcol[c] = t[max(0, min((c * 16 + lr) >> 1, k - 1))];
    constexpr int EG = CT < 3 ? CT : 3;
#pragma unroll
    for (int c0 = 0; c0 < CT; c0 += EG) {
        float wv[RT][EG][4];
#pragma unroll
        for (int r = 0; r < RT; r++)
#pragma unroll
            for (int e = 0; e < EG; e++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int c = c0 + e;
                    const int col = c * 16 + lr, cc = col >> 1, q = col & 1;
                    const int row = n0 + 16 * r + 4 * lq + v;
                    wv[r][e][v] = 0.f;
                    if (cc < k && row < d.N) {
                        gflt* w = ROW ? (gflt*)(W0 + (size_t)tcol[c] * d.ldN + row)
                                      : (gflt*)(W0 + (size_t)row * d.ldN + tcol[c]);
                        wv[r][e][v] = w[q];
                    }
                }
#pragma unroll
        for (int r = 0; r < RT; r++)
#pragma unroll
            for (int e = 0; e < EG; e++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int c = c0 + e;
                    const int col = c * 16 + lr, cc = col >> 1, q = col & 1;
                    const int row = n0 + 16 * r + 4 * lq + v;
                    if (cc >= k || row >= d.N) continue;
                    const float a = acc[r][c][v];
                    if (!ROW)
                        ((gflt_w*)A)[((size_t)row * d.Kmax + K + cc) * 2 + q] = wv[r][e][v] - a;
                    else
                        ((gflt_w*)Bt)[((size_t)f * bt_stride + (size_t)row * d.gmax + cc) * 2 + q] = wv[r][e][v] - a;
                }
    }
}

template <int CT, bool ROW, int KC, int DEPTH, int WAVES, bool DB, int RT>
static void launch_variant32(hipStream_t st, const Dims& d, const DevTabsT<zf>& tb, const int32_t* t, int k, int K,
                             const zf* small, int small_stride, zf* Bt, size_t bt_stride, const int32_t* dyn) {
    const int rows = 16 * WAVES * RT;
    dim3 grid((d.N + rows - 1) / rows, d.nf);
    hipLaunchKernelGGL((lr_skinny_mfma32_kernel<CT, ROW, KC, DEPTH, WAVES, DB, RT>), grid, dim3(WAVES * 64), 0, st, d,
                       tb, t, k, K, small, small_stride, Bt, bt_stride, dyn);
}

template <bool ROW>
static void launch_skinny_mfma32(hipStream_t st, const Dims& d, const DevTabsT<zf>& tb, const int32_t* t, int k,
                                 int K, const zf* small, int small_stride, zf* Bt, size_t bt_stride, int variant,
                                 const int32_t* dyn) {
#define PNGF_SKV32(CT, KC, DEPTH, W, DB, RT) \
    launch_variant32<CT, ROW, KC, DEPTH, W, DB, RT>(st, d, tb, t, k, K, small, small_stride, Bt, bt_stride, dyn)
    if (k <= 24) {
        switch (variant) {
            case 1: PNGF_SKV32(3, 16, 2, 8, true, 1); break;
            case 2: PNGF_SKV32(3, 32, 2, 4, true, 2); break;
            case 3: PNGF_SKV32(3, 64, 2, 8, true, 1); break;
            default: PNGF_SKV32(3, 32, 2, 8, true, 1); break;
        }
    } else if (k <= 48) {
        PNGF_SKV32(6, 16, 2, 8, true, 1);
    } else {
        PNGF_SKV32(12, 16, 2, 4, false, 1);
    }
#undef PNGF_SKV32
}

void launch_lr_colblk(hipStream_t st, const Dims& d, const DevTabsT<zf>& tb, const int32_t* t, int k, int K,
                      const zf* Mc, int small_stride, int variant, const int32_t* dyn) {
    launch_skinny_mfma32<false>(st, d, tb, t, k, K, Mc, small_stride, nullptr, 0, variant, dyn);
}

void launch_lr_rowblk(hipStream_t st, const Dims& d, const DevTabsT<zf>& tb, const int32_t* t, int k, int K,
                      const zf* Ar, int small_stride, zf* Bt, size_t bt_stride, int variant, const int32_t* dyn) {
    launch_skinny_mfma32<true>(st, d, tb, t, k, K, Ar, small_stride, Bt, bt_stride, variant, dyn);
}

// out[c][n] = sgn_c * sum_l capinv[c][l] B[n][l]   (B row n: k entries at stride ldb;
// B = b_tab[f] + b_col0 when b_tab is given, else B + f * b_stride)
// PLAIN:   M = cap^-1 VCinv           (pngfopt.hpp:276-333; no sign)
// LOWRANK: M_new = diag(s) cap^-1 W[t,:]  (W[t,:]^T = W[:,t], the new A_all columns, when W
//          is symmetric)
// k <= 32: thread e of the workgroup's 32 x k outputs (c = e / 32, n = e % 32) sums over l with
// capinv read from global memory (its rows are shared by the workgroup's threads).
template <class Z>
__global__ __launch_bounds__(256) void apply_capinv_small_kernel(Dims d, const Z* __restrict__ acc, int acc_stride,
                                                                 const Z* __restrict__ B, size_t b_stride,
                                                                 const Z* const* b_tab, int b_col0, int ldb, int k,
                                                                 const int32_t* __restrict__ s, Z* const* out_tab,
                                                                 Z* out_single, size_t out_stride, int out_row0,
                                                                 int ld_out) {
    __shared__ Z Bs[32][33];
    const int f = blockIdx.y, tid = threadIdx.x;
    const int n0 = blockIdx.x * 32;
    const Z* Bf = b_tab ? b_tab[f] + b_col0 : B + (size_t)f * b_stride;
    const Z* ci = acc + (size_t)f * acc_stride;
    for (int e = tid; e < 32 * k; e += 256) {
        const int nn = e / k, l = e - nn * k;
        const int n = n0 + nn;
        Bs[nn][l] = n < d.N ? Bf[(size_t)n * ldb + l] : zc<Z>(0, 0);
    }
    __syncthreads();
    Z* out = out_tab ? out_tab[f] : out_single + (size_t)f * out_stride;
    for (int e = tid; e < 32 * k; e += 256) {
        const int c = e / 32, nn = e - c * 32;
        const int n = n0 + nn;
        if (n >= d.N) continue;
        Z sacc = zc<Z>(0, 0);
        for (int l = 0; l < k; l++) sacc = zfma(ci[c * k + l], Bs[nn][l], sacc);
        if (s && s[c] < 0) sacc = zc<Z>(-sacc.x, -sacc.y);
        out[(size_t)(out_row0 + c) * ld_out + n] = sacc;
    }
}

// k > 32: a workgroup takes 32 rows n (B rows staged in LDS) and every c; thread (tn, tcg) keeps the 4 x 3
// outputs n = tn + 8 a, c = tcg + 32 b in registers, reading capinv 16 columns at a time from an
// LDS copy, so that each l costs 7 LDS reads for 12 updates.  Every output's sum runs over l in
// ascending order from zero.  k <= 96.
template <class Z>
__global__ __launch_bounds__(256) void apply_capinv_kernel(Dims d, const Z* __restrict__ acc, int acc_stride,
                                                           const Z* __restrict__ B, size_t b_stride,
                                                           const Z* const* b_tab, int b_col0, int ldb, int k,
                                                           const int32_t* __restrict__ s, Z* const* out_tab,
                                                           Z* out_single, size_t out_stride, int out_row0,
                                                           int ld_out) {
    constexpr int NA = 4, NB = 3, LC = 16;
    __shared__ Z Bs[32][97];
    __shared__ Z Cs[96][LC + 1];
    const int f = blockIdx.y, tid = threadIdx.x;
    const int n0 = blockIdx.x * 32;
    const Z* Bf = b_tab ? b_tab[f] + b_col0 : B + (size_t)f * b_stride;
    const Z* ci = acc + (size_t)f * acc_stride;
    for (int e = tid; e < 32 * k; e += 256) {
        const int nn = e / k, l = e - nn * k;
        const int n = n0 + nn;
        Bs[nn][l] = n < d.N ? Bf[(size_t)n * ldb + l] : zc<Z>(0, 0);
    }
    const int tn = tid % 8, tcg = tid / 8;
    Z sacc[NA][NB];
#pragma unroll
    for (int a = 0; a < NA; a++)
#pragma unroll
        for (int b = 0; b < NB; b++) sacc[a][b] = zc<Z>(0, 0);
    for (int l0 = 0; l0 < k; l0 += LC) {
        const int lc = min(LC, k - l0);
        __syncthreads();
        for (int e = tid; e < k * LC; e += 256) {
            const int c = e / LC, l = e - c * LC;
            Cs[c][l] = l < lc ? ci[c * k + l0 + l] : zc<Z>(0, 0);
        }
        __syncthreads();
        for (int l = 0; l < lc; l++) {
            Z bv[NA], cv[NB];
#pragma unroll
            for (int a = 0; a < NA; a++) bv[a] = Bs[tn + 8 * a][l0 + l];
#pragma unroll
            for (int b = 0; b < NB; b++) cv[b] = Cs[min(tcg + 32 * b, k - 1)][l];
#pragma unroll
            for (int a = 0; a < NA; a++)
#pragma unroll
                for (int b = 0; b < NB; b++) sacc[a][b] = zfma(cv[b], bv[a], sacc[a][b]);
        }
    }
    Z* out = out_tab ? out_tab[f] : out_single + (size_t)f * out_stride;
#pragma unroll
    for (int a = 0; a < NA; a++)
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int n = n0 + tn + 8 * a, c = tcg + 32 * b;
            if (n < d.N && c < k) {
                Z v = sacc[a][b];
                if (s && s[c] < 0) v = zc<Z>(-v.x, -v.y);
                out[(size_t)(out_row0 + c) * ld_out + n] = v;
            }
        }
}

template <class Z>
void launch_apply_capinv(hipStream_t st, const Dims& d, const Z* acc, int acc_stride, const Z* Bt,
                         size_t bt_stride, int ldb, int k, const int32_t* s, Z* const* out_tab,
                         Z* out_single, size_t out_stride, int out_row0, int ld_out, const Z* const* b_tab,
                         int b_col0) {
    dim3 grid((d.N + 31) / 32, d.nf);
    if (k <= 32)
        hipLaunchKernelGGL(apply_capinv_small_kernel<Z>, grid, dim3(256), 0, st, d, acc, acc_stride, Bt, bt_stride,
                           b_tab, b_col0, ldb, k, s, out_tab, out_single, out_stride, out_row0, ld_out);
    else
        hipLaunchKernelGGL(apply_capinv_kernel<Z>, grid, dim3(256), 0, st, d, acc, acc_stride, Bt, bt_stride, b_tab,
                           b_col0, ldb, k, s, out_tab, out_single, out_stride, out_row0, ld_out);
}

// q += A_new w ;  R[r] -= Rt[r] M_new   (O(kN) parts of the LOWRANK accept).  A workgroup of 128
// takes 32 components n: their rows of A_new (k entries each, 169 KB apart in A_all) are staged
// in LDS with consecutive lanes on consecutive entries, then thread n < 32 forms q[n] and threads
// 32.. the R[r][n] (M_new rows read across consecutive n).  Each sum runs over c in ascending
// order from the old value.
template <class Z>
__global__ __launch_bounds__(128) void lr_small_kernel(Dims d, DevTabsT<Z> tb, DevVecsT<Z> v,
                                                       const Z* __restrict__ acc, int acc_stride, int k, int K) {
    constexpr int RW = 32;
    __shared__ Z As[RW][97];
    const int f = blockIdx.y, tid = threadIdx.x;
    const int n0 = blockIdx.x * RW;
    const Z* st = acc + (size_t)f * acc_stride;
    const Z* w = st + 2 * k * k + k;
    const Z* Rt = st + 2 * k * k + 2 * k + 4;
    const Z* A = tb.Aall[f];
    for (int e = tid; e < RW * k; e += 128) {
        const int r = e / k, c = e - r * k;
        const int n = n0 + r;
        As[r][c] = n < d.N ? A[(size_t)n * d.Kmax + K + c] : zc<Z>(0, 0);
    }
    __syncthreads();
    if (tid < RW) {
        const int n = n0 + tid;
        if (n < d.N) {
            Z qa = v.q[(size_t)f * d.ldN + n];
            for (int c = 0; c < k; c++) qa = zfma(As[tid][c], w[c], qa);
            v.q[(size_t)f * d.ldN + n] = qa;
        }
    } else {
        const Z* M = tb.Mall[f] + (size_t)K * d.ldN;
        for (int e = tid - RW; e < d.ne * RW; e += 128 - RW) {
            const int r = e / RW, n = n0 + (e - r * RW);
            if (n >= d.N) continue;
            Z* R = v.R + ((size_t)f * d.ne + r) * d.ldN;
            Z ra = R[n];
            for (int c = 0; c < k; c++) ra = zfms(Rt[r * k + c], M[(size_t)c * d.ldN + n], ra);
            R[n] = ra;
        }
    }
}

template <class Z>
void launch_lr_small(hipStream_t st, const Dims& d, const DevTabsT<Z>& tb, const DevVecsT<Z>& v, const Z* acc,
                     int acc_stride, int k, int K) {
    dim3 grid((d.N + 31) / 32, d.nf);
    hipLaunchKernelGGL(lr_small_kernel<Z>, grid, dim3(128), 0, st, d, tb, v, acc, acc_stride, k, K);
}

// D_h -= A_new[comp_h, :] M_new[:, comp_h] for every group h (groups of up to 96 components: the
// accept path of flips with k > 32; smaller ones fold this into lr_finish_kernel).  Thread (ta, tb)
// of a 16 x 16 grid keeps the 6 x 6 entries of rows ta + 16 i and columns tb + 16 j in registers,
// so that each step c reads 6 entries of A_new and 6 of M_new from LDS for 36 updates (one entry
// per thread would read two per update).  Every entry's sum runs over c in ascending order from
// zero, 16 columns of A_new and rows of M_new staged at a time.
template <class Z>
__global__ void __launch_bounds__(256) lr_blocks_tile_kernel(Dims d, DevTabsT<Z> tb, DevVecsT<Z> v, GroupTab gt,
                                                             int k, int K) {
    constexpr int T = 6, G = 16;
    __shared__ Z Ac[16][97];  // Ac[c][a] = A_new[comp a][c0 + c]
    __shared__ Z Mc[16][97];  // Mc[c][b] = M_new[c0 + c][comp b]
    const int h = blockIdx.x, f = blockIdx.y, tid = threadIdx.x;
    const int nc = gt.ncomp[h];
    if (nc == 0) return;
    const int ta = tid / G, tbb = tid % G;
    const int32_t* cl = gt.comps + (size_t)h * gt.gmax;
    const Z* A = tb.Aall[f];
    const Z* M = tb.Mall[f];
    Z* D = v.Dblk + ((size_t)f * d.ngroups + h) * gt.gmax * gt.gmax;
    Z acc[T][T];
#pragma unroll
    for (int i = 0; i < T; i++)
#pragma unroll
        for (int j = 0; j < T; j++) acc[i][j] = zc<Z>(0, 0);
    for (int c0 = 0; c0 < k; c0 += 16) {
        const int cl_ = min(16, k - c0);
        for (int e = tid; e < nc * 16; e += 256) {
            const int a = e / 16, c = e - a * 16;
            Ac[c][a] = c < cl_ ? A[(size_t)cl[a] * d.Kmax + K + c0 + c] : zc<Z>(0, 0);
        }
        for (int e = tid; e < 16 * nc; e += 256) {
            const int c = e / nc, bb = e - c * nc;
            Mc[c][bb] = c < cl_ ? M[(size_t)(K + c0 + c) * d.ldN + cl[bb]] : zc<Z>(0, 0);
        }
        __syncthreads();
        for (int c = 0; c < cl_; c++) {
            Z av[T], mv[T];
#pragma unroll
            for (int i = 0; i < T; i++) av[i] = Ac[c][ta + G * i];
#pragma unroll
            for (int j = 0; j < T; j++) mv[j] = Mc[c][tbb + G * j];
#pragma unroll
            for (int i = 0; i < T; i++)
#pragma unroll
                for (int j = 0; j < T; j++) acc[i][j] = zfma(av[i], mv[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < T; i++)
#pragma unroll
        for (int j = 0; j < T; j++) {
            const int a = ta + G * i, bb = tbb + G * j;
            if (a < nc && bb < nc) {
                Z* p = &D[a * gt.gmax + bb];
                *p = zsub(*p, acc[i][j]);
            }
        }
}

template <class Z>
void launch_lr_blocks(hipStream_t st, const Dims& d, const DevTabsT<Z>& tb, const DevVecsT<Z>& v, GroupTab gt, int k,
                      int K) {
    hipLaunchKernelGGL(lr_blocks_tile_kernel<Z>, dim3(d.ngroups, d.nf), dim3(256), 0, st, d, tb, v, gt, k, K);
}

// LOWRANK accept epilogue for k <= 32 in one launch.  A block takes a list of up to 32
// components ("rows") and does
//   rows part:  M_new[c][n] = s_c sum_l capinv'[c][l] S[n][l]  (S[n][:] = W[t, n]: the new A_all
//               columns when W is symmetric, else the row strip Bt), written to M_all;
//               q[n] += sum_c A_new[n][c] w[c];   R[r][n] -= sum_c Rt[r][c] M_new[c][n]
//   block part: D_h[a][b] -= sum_c A_new[cl_a][c] M_new[c][cl_b] for the group h whose
//               components are the rows
// Default layout of the grid: blocks [0, nchunk) take 32 consecutive components (rows part),
// blocks [nchunk, nchunk + ngroups) one group each (block part, M_new recomputed at its
// components).  Partition layout (the candidate groups' component lists are disjoint, as in the
// disjoint layout's raster groups): block h < ngroups takes group h and does both parts, the
// remaining blocks take 32 of the components outside every group (`left`).  Same operations in
// the same order as apply_capinv, lr_small and lr_blocks.
// global-space stores of the (re, im) pair i of a complex array (a table entry is otherwise a
// flat pointer, whose stores the LDS waits of the kernel below would also wait for)
__device__ inline void gstore2(zd* base, size_t i, double x, double y) {
    gdbl_w* p = (gdbl_w*)base;
    p[2 * i] = x;
    p[2 * i + 1] = y;
}
__device__ inline void gstore2(zf* base, size_t i, float x, float y) {
    gflt_w* p = (gflt_w*)base;
    p[2 * i] = x;
    p[2 * i + 1] = y;
}
__device__ inline zd ldgz(const zd* p, size_t i) { return ldg(p, i); }
__device__ inline zf ldgz(const zf* p, size_t i) { return ldgf(p, i); }

// SYMA (symmetric W): the M_new rows take the capacitance block's LDS once it is dead, so a
// workgroup holds two G x (G + 1) arrays instead of three, and 8 workgroups of 4 waves fit a CU
// (19.3 KB of LDS and 64 VGPRs at G = 24) instead of 5.
template <class Z, int G, bool SYMA>
__global__ void __launch_bounds__(256, 8) lr_finish_kernel(Dims d, DevTabsT<Z> tb, DevVecsT<Z> v, GroupTab gt,
                                                        const Z* __restrict__ acc, int acc_stride,
                                                        const Z* __restrict__ Bt, size_t bt_stride,
                                                        const int32_t* __restrict__ s, int k, int K, int nchunk,
                                                        const int32_t* dyn, const int32_t* __restrict__ left,
                                                        int nleft, int ngroups_part, FlipUpdate fu) {
    if (dyn) {
        if (!dyn[kDynActive]) return;
        k = dyn[kDynK];
        K = dyn[kDynKK];
    }
    if (fu.P && blockIdx.x == 0 && blockIdx.y == 0) {
        // the flip's components are distinct, and so are its tiles: one thread each
        for (int c = threadIdx.x; c < k; c += 256) fu.P[fu.t[c]] = (int8_t)(fu.P[fu.t[c]] + s[c]);
        for (int i = threadIdx.x; i < dyn[kDynNt]; i += 256) fu.tile[fu.gtiles[i]] ^= 1;
    }
    __shared__ Z Xs[G][G + 1];
    __shared__ Z As[G][G + 1];
    __shared__ Z Ss2[SYMA ? 1 : G][G + 1];
    // row strip (two-strip path), then M_new[c][row]; in the symmetric path the storage of Xs,
    // which nothing reads after the first product
    Z (*Ss)[G + 1] = SYMA ? Xs : Ss2;
    __shared__ int rows[G];
    const int f = blockIdx.y, tid = threadIdx.x, bx = blockIdx.x;
    bool do_rows, do_block;
    int h = -1, nrow, r0 = 0;
    if (left) {
        do_rows = true;
        do_block = bx < ngroups_part;
        if (do_block) {
            h = bx;
            nrow = gt.ncomp[h];
        } else {
            r0 = (bx - ngroups_part) * G;
            nrow = min(G, nleft - r0);
        }
    } else {
        do_rows = bx < nchunk;
        do_block = !do_rows;
        if (do_block) {
            h = bx - nchunk;
            nrow = gt.ncomp[h];
        } else {
            r0 = bx * G;
            nrow = min(G, d.N - r0);
        }
    }
    const Z* stf = acc + (size_t)f * acc_stride;
    if (d.lam && bx == 0) {
        // the accept's M rows, kept as Lam[K + c] = s_c cap^-1[c][:] (see DevTabs)
        Z* L = tb.Lam[f] + (size_t)K * d.gmax;
        for (int e = tid; e < k * k; e += 256) {
            const int c = e / k, l = e - c * k;
            const Z x = stf[e];
            gstore2(L, (size_t)c * d.gmax + l, s[c] < 0 ? -x.x : x.x, s[c] < 0 ? -x.y : x.y);
        }
        if (f == 0 && tid < k) {
            tb.fstart[K + tid] = K;
            tb.flen[K + tid] = k;
        }
    }
    if (nrow <= 0) return;
    for (int e = tid; e < k * k; e += 256) {
        const int r = e / k, c = e - r * k;
        Xs[r][c] = stf[e];
    }
    if (tid < nrow) rows[tid] = h >= 0 ? gt.comps[(size_t)h * gt.gmax + tid] : (left ? left[r0 + tid] : r0 + tid);
    __syncthreads();
    const Z* A = tb.Aall[f];
    const Z* S = Bt + (size_t)f * bt_stride;
    for (int e = tid; e < nrow * k; e += 256) {
        const int rr = e / k, c = e - rr * k;
        const int n = rows[rr];
        As[rr][c] = ldgz(A, (size_t)n * d.Kmax + K + c);
        if (!d.sym) Ss[rr][c] = S[(size_t)n * d.gmax + c];
    }
    __syncthreads();
    Z m[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int e = tid + 256 * u;
        m[u] = zc<Z>(0, 0);
        if (e < k * nrow) {
            const int c = e / nrow, rr = e - c * nrow;
            Z sacc = zc<Z>(0, 0);
            for (int l = 0; l < k; l++) sacc = zfma(Xs[c][l], d.sym ? As[rr][l] : Ss[rr][l], sacc);
            m[u] = s[c] < 0 ? zc<Z>(-sacc.x, -sacc.y) : sacc;
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int e = tid + 256 * u;
        if (e < k * nrow) {
            const int c = e / nrow, rr = e - c * nrow;
            Ss[c][rr] = m[u];
        }
    }
    __syncthreads();
    if (do_rows) {
        if (!d.lam) {
            Z* M = tb.Mall[f];
            for (int e = tid; e < k * nrow; e += 256) {
                const int c = e / nrow, rr = e - c * nrow;
                const Z v = Ss[c][rr];
                gstore2(M, (size_t)(K + c) * d.ldN + rows[rr], v.x, v.y);
            }
        }
        const Z* w = stf + 2 * k * k + k;
        const Z* Rt = stf + 2 * k * k + 2 * k + 4;
        if (tid < nrow) {
            Z* qp = v.q + (size_t)f * d.ldN + rows[tid];
            Z qa = *qp;
            for (int c = 0; c < k; c++) qa = zfma(As[tid][c], w[c], qa);
            *qp = qa;
        } else if (tid >= 64 && tid < 64 + 32 * d.ne) {
            const int r = (tid - 64) >> 5, rr = (tid - 64) & 31;
            if (rr < nrow) {
                Z* rp = v.R + ((size_t)f * d.ne + r) * d.ldN + rows[rr];
                Z ra = *rp;
                for (int c = 0; c < k; c++) ra = zfms(Rt[r * k + c], Ss[c][rr], ra);
                *rp = ra;
            }
        }
    }
    if (do_block) {
        Z* Dh = v.Dblk + ((size_t)f * d.ngroups + h) * gt.gmax * gt.gmax;
        for (int e = tid; e < nrow * nrow; e += 256) {
            const int a = e / nrow, b = e - a * nrow;
            Z sacc = zc<Z>(0, 0);
            for (int c = 0; c < k; c++) sacc = zfma(As[a][c], Ss[c][b], sacc);
            Dh[a * gt.gmax + b] = zsub(Dh[a * gt.gmax + b], sacc);
        }
    }
}

template <class Z>
void launch_lr_finish(hipStream_t st, const Dims& d, const DevTabsT<Z>& tb, const DevVecsT<Z>& v, GroupTab gt,
                      const Z* acc, int acc_stride, const Z* Bt, size_t bt_stride, const int32_t* s, int k,
                      int K, const int32_t* dyn, const int32_t* left, int nleft, int ngroups_part, FlipUpdate fu) {
    // G: components per workgroup (the group size bound); k <= 24 halves the LDS footprint
    if (d.gmax <= 24) {
        const int nchunk = (d.N + 23) / 24;
        const int nblk = left ? ngroups_part + (nleft + 23) / 24 : nchunk + d.ngroups;
        if (d.sym)
            hipLaunchKernelGGL((lr_finish_kernel<Z, 24, true>), dim3(nblk, d.nf), dim3(256), 0, st, d, tb, v, gt, acc,
                               acc_stride, Bt, bt_stride, s, k, K, nchunk, dyn, left, nleft, ngroups_part, fu);
        else
            hipLaunchKernelGGL((lr_finish_kernel<Z, 24, false>), dim3(nblk, d.nf), dim3(256), 0, st, d, tb, v, gt, acc,
                               acc_stride, Bt, bt_stride, s, k, K, nchunk, dyn, left, nleft, ngroups_part, fu);
    } else {
        const int nchunk = (d.N + 31) / 32;
        const int nblk = left ? ngroups_part + (nleft + 31) / 32 : nchunk + d.ngroups;
        if (d.sym)
            hipLaunchKernelGGL((lr_finish_kernel<Z, 32, true>), dim3(nblk, d.nf), dim3(256), 0, st, d, tb, v, gt, acc,
                               acc_stride, Bt, bt_stride, s, k, K, nchunk, dyn, left, nleft, ngroups_part, fu);
        else
            hipLaunchKernelGGL((lr_finish_kernel<Z, 32, false>), dim3(nblk, d.nf), dim3(256), 0, st, d, tb, v, gt, acc,
                               acc_stride, Bt, bt_stride, s, k, K, nchunk, dyn, left, nleft, ngroups_part, fu);
    }
}

__global__ void mask_update_kernel(int8_t* P, int32_t* tile, const int32_t* t, const int32_t* s, int k,
                                   const int32_t* gtiles, int ntiles) {
    // the components of a flip are distinct, and so are its tiles: one thread each
    for (int c = threadIdx.x; c < k; c += blockDim.x) P[t[c]] = (int8_t)(P[t[c]] + s[c]);
    for (int i = threadIdx.x; i < ntiles; i += blockDim.x) tile[gtiles[i]] ^= 1;
}

void launch_mask_update(hipStream_t st, int8_t* P, int32_t* tile, const int32_t* t, const int32_t* s, int k,
                        const int32_t* gtiles, int ntiles) {
    hipLaunchKernelGGL(mask_update_kernel, dim3(1), dim3(128), 0, st, P, tile, t, s, k, gtiles, ntiles);
}

// ---------------------------------------------------- device-driven loop -----
// Raster first-improvement iterations without a host round trip.  Per iteration: the scoring
// kernel (lr_eval_kernel) reads the batch size from ls->nb, and its last workgroup selects,
// writes the accept record and the next batch; the accept kernels read k, K and "accept
// pending" from ls->dyn; the finish kernel applies the flip to the mask and the tile map, and
// the next iteration's loop control adds its k to the factor count.
// Size and start of the next batch: the candidates after the current position, cyclically,
// at most bs of them (doubling after each non-improving batch), none once the run is done.
__device__ void loop_next_batch(LoopState* ls, int* B_out, int* pos_out, int* ncl_out) {
    int64_t B = 0;
    if (!ls->done) {
        B = ls->ncl - ls->nonimp;
        if (ls->max_evals > 0) B = min(B, ls->max_evals - ls->flip_count);
        B = min(B, (int64_t)min(ls->bs, ls->bmax));
        B = max(B, (int64_t)0);
    }
    ls->nb = (int)B;
    *B_out = (int)B;
    *pos_out = ls->pos;
    *ncl_out = ls->ncl;
}

__global__ void loop_batch_kernel(LoopState* ls, const int32_t* __restrict__ cl, int32_t* batch) {
    __shared__ int s_B, s_pos, s_ncl;
    const int tid = threadIdx.x;
    if (tid == 0) loop_next_batch(ls, &s_B, &s_pos, &s_ncl);
    __syncthreads();
    for (int i = tid; i < s_B; i += blockDim.x) batch[i] = cl[(s_pos + i) % s_ncl];
}

void launch_loop_batch(hipStream_t st, LoopState* ls, const int32_t* cl, int32_t* batch) {
    hipLaunchKernelGGL(loop_batch_kernel, dim3(1), dim3(256), 0, st, ls, cl, batch);
}

// The selection (ls->sel, ls->F_sel) turned into the accept record the accept kernels read,
// or into the non-improving bookkeeping; one wave (lanes 0..63).  Same rules as the host loop
// in pngf_dbs_run.
__device__ void loop_after_select_wave(const Dims& d, LoopState* ls, const GroupTab& gt, const int32_t* tile,
                                       const int8_t* P, const int32_t* owners, const int32_t* batch,
                                       const int32_t* info, const LoopCtl& c) {
    const int tid = threadIdx.x;
    // the previous iteration's accept (its kernels have all run) joins the factor count
    if (tid == 0 && ls->dyn[kDynActive]) {
        ls->dyn[kDynKK] += ls->dyn[kDynK];
        ls->dyn[kDynActive] = 0;
    }
    if (ls->done) return;
    const int nb = ls->nb, sel = ls->sel;
    // a singular capacitance matrix in the batch ends the run with an error, as in the host loop
    bool bad = false;
    for (int i = tid; i < nb * d.nf; i += 64) bad |= info[i] != 0;
    if (__ballot(bad)) {
        if (tid == 0) {
            ls->status = kLoopSingular;
            ls->done = 1;
            ls->dyn[kDynActive] = 0;
        }
        return;
    }
    if (sel >= 0) {
        const int g = batch[sel];
        const int k = change_set_wave(gt, g, tile, P, owners, c.t, c.s, c.lp);
        if (tid == 0) {
            if (ls->nacc >= ls->acc_cap) {
                ls->status = kLoopFull;
                ls->done = 1;
                ls->dyn[kDynActive] = 0;
                return;
            }
            c.acc_group[ls->nacc] = g;
            c.acc_eval[ls->nacc] = ls->flip_count + sel;
            c.acc_F[ls->nacc] = ls->F_sel;
            ls->nacc++;
            ls->flip_count += sel + 1;
            ls->spec += nb - sel - 1;
            ls->F_best = ls->F_sel;
            ls->pos = (ls->pos + sel + 1) % ls->ncl;
            ls->nonimp = 0;
            ls->bs = ls->bmin;
            const int nt = gt.ntiles[g];
            for (int i = 0; i < nt; i++) c.gtiles[i] = gt.tiles[4 * g + i];
            ls->g = g;
            ls->dyn[kDynK] = k;
            ls->dyn[kDynNt] = nt;
            ls->dyn[kDynActive] = 1;
        }
    } else if (tid == 0) {
        ls->flip_count += nb;
        ls->nonimp += nb;
        ls->pos = (int)((ls->pos + nb) % ls->ncl);
        ls->bs = min(2 * ls->bs, 1 << 30);
        ls->dyn[kDynActive] = 0;
    }
    if (tid == 0) {
        ls->batches += nb > 0 ? 1 : 0;
        if (ls->nonimp >= ls->ncl || (ls->max_evals > 0 && ls->flip_count >= ls->max_evals)) ls->done = 1;
    }
}

// --------------------------------------------------------- PLAIN accept -----
// V (k x N column-major, ld k) = (G - I)[t, :]  (pngfopt.hpp:420-432)
__global__ void build_V_kernel(Dims d, DevTabs tb, const int32_t* __restrict__ t, int k, zd* V, size_t v_stride) {
    const int f = blockIdx.y;
    const zd* Gr = tb.Grow[f];
    zd* Vf = V + (size_t)f * v_stride;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < (size_t)k * d.N;
         e += (size_t)gridDim.x * blockDim.x) {
        const int n = (int)(e / k), i = (int)(e - (size_t)n * k);
        zd val = Gr[(size_t)t[i] * d.ldN + n];
        if (n == t[i]) val.x -= 1.0;
        Vf[e] = val;
    }
}

void launch_build_V(hipStream_t st, const Dims& d, const DevTabs& tb, const int32_t* t, int k, zd* V,
                    size_t v_stride) {
    hipLaunchKernelGGL(build_V_kernel, dim3(1024, d.nf), dim3(256), 0, st, d, tb, t, k, V, v_stride);
}

// CinvU (N x k column-major, ld ldN) = C^-1[:, t] diag(s)   (pngfopt.hpp:260-274)
__global__ void snapshot_cinvu_kernel(Dims d, DevTabs tb, const int32_t* __restrict__ t,
                                      const int32_t* __restrict__ s, int k, zd* CU, size_t cu_stride) {
    const int f = blockIdx.y, c = blockIdx.z;
    const zd* Ci = tb.Cinv[f] + (size_t)t[c] * d.ldN;
    zd* out = CU + (size_t)f * cu_stride + (size_t)c * d.ldN;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < d.N; n += gridDim.x * blockDim.x) {
        zd val = Ci[n];
        out[n] = s[c] > 0 ? val : zmk(-val.x, -val.y);
    }
}

void launch_snapshot_cinvu(hipStream_t st, const Dims& d, const DevTabs& tb, const int32_t* t,
                           const int32_t* s, int k, zd* CU, size_t cu_stride) {
    hipLaunchKernelGGL(snapshot_cinvu_kernel, dim3((d.N + 255) / 256, d.nf, k), dim3(256), 0, st, d, tb, t, s, k,
                       CU, cu_stride);
}

// S' = S - CinvU (Einc_t + y)  ;  R' = R - RU M  with RU = R[:,t] diag(s);  Eobj = newEobj.
// (exact Woodbury forms of compute_R_matrix(0), pngfopt.hpp:237-250)
// M is row-major k x ldN (M = cap^-1 V C^-1).
__global__ void plain_small_kernel(Dims d, DevVecs v, const zd* __restrict__ acc, int acc_stride,
                                   const zd* __restrict__ CU, size_t cu_stride, const zd* __restrict__ M,
                                   size_t m_stride, const int32_t* __restrict__ t, const zd* __restrict__ rut,
                                   int k) {
    const int f = blockIdx.y;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const zd* st = acc + (size_t)f * acc_stride;
    const zd* y = st + 2 * k * k;
    const zd* einc = v.Einc + (size_t)f * d.ldN;
    const zd* ru = rut + (size_t)f * d.ne * 96;
    if (n < d.N) {
        const zd* cu = CU + (size_t)f * cu_stride;
        zd sa = v.S[(size_t)f * d.ldN + n];
        for (int c = 0; c < k; c++) sa = zfms(cu[(size_t)c * d.ldN + n], zadd(einc[t[c]], y[c]), sa);
        v.S[(size_t)f * d.ldN + n] = sa;
        const zd* Mf = M + (size_t)f * m_stride;
        for (int r = 0; r < d.ne; r++) {
            zd* R = v.R + ((size_t)f * d.ne + r) * d.ldN;
            zd ra = R[n];
            for (int c = 0; c < k; c++) ra = zfms(ru[r * 96 + c], Mf[(size_t)c * d.ldN + n], ra);
            R[n] = ra;
        }
    }
}

__global__ void plain_rut_kernel(Dims d, DevVecs v, const zd* acc, int acc_stride, const int32_t* t,
                                 const int32_t* s, int k, zd* rut) {
    const int f = blockIdx.x, tid = threadIdx.x;
    for (int e = tid; e < d.ne * k; e += blockDim.x) {
        const int r = e / k, c = e - r * k;
        zd val = v.R[((size_t)f * d.ne + r) * d.ldN + t[c]];
        rut[(size_t)f * d.ne * 96 + r * 96 + c] = s[c] > 0 ? val : zmk(-val.x, -val.y);
    }
    if (tid < d.ne) v.Eobj[f * d.ne + tid] = acc[(size_t)f * acc_stride + 2 * k * k + 2 * k + tid];
}

void launch_plain_small(hipStream_t st, const Dims& d, const DevVecs& v, const zd* acc, int acc_stride,
                        const zd* CU, size_t cu_stride, const zd* M, size_t m_stride, const int32_t* t,
                        const int32_t* s, int k, zd* rut) {
    // RU snapshot must precede the R update
    hipLaunchKernelGGL(plain_rut_kernel, dim3(d.nf), dim3(64), 0, st, d, v, acc, acc_stride, t, s, k, rut);
    hipLaunchKernelGGL(plain_small_kernel, dim3((d.N + 255) / 256, d.nf), dim3(256), 0, st, d, v, acc, acc_stride,
                       CU, cu_stride, M, m_stride, t, rut, k);
}

// ------------------------------------------------------- ranker features ---
// Inputs of the learned ranker for candidate group g at the current state
// (LOWRANK state).  Row layout (din = 51 + 8 nf floats):
//   [0, 49)   metal bits of the 7x7 tile neighbourhood of the primary tile (0 outside)
//   49, 50    primary tile x, y scaled to [-1, 1]
//   51 + 8f + {0..7}: Re/Im of four complex terms at frequency f:
//       c1 = mean_c cap_cc                      (capacitance diagonal, cap = I + W[t,t] diag(s))
//       c2 = zfac sum_c R0[t_c] dPE_c           (first-order change of the port voltage)
//       c3 = zfac sum_c R0[t_c] s_c rhs_c / cap_cc   (its Jacobi correction)
//       c4 = Zin_f / z0                          (current input impedance)
__global__ void features_kernel(Dims d, DevVecs v, GroupTab gt, const int32_t* __restrict__ tile,
                                const int8_t* __restrict__ P, const int32_t* __restrict__ owners,
                                const int32_t* __restrict__ batch, int Tx, int Ty, ObjParams op, float* out) {
    __shared__ int t[96], sg[96], lp[96];
    __shared__ int s_k;
    const int b = blockIdx.x, tid = threadIdx.x;
    const int g = batch[b];
    const int din = 51 + 8 * d.nf;
    float* row = out + (size_t)b * din;
    if (tid == 0) s_k = change_set(gt, g, tile, P, owners, t, sg, lp);
    __syncthreads();
    const int k = s_k;
    const int prim = gt.tiles[4 * g];
    const int ty = prim / Tx, tx = prim - ty * Tx;
    for (int e = tid; e < 49; e += blockDim.x) {
        const int x = tx + e % 7 - 3, y = ty + e / 7 - 3;
        row[e] = (x >= 0 && x < Tx && y >= 0 && y < Ty) ? (float)tile[y * Tx + x] : 0.f;
    }
    if (tid == 0) {
        row[49] = Tx > 1 ? 2.f * tx / (Tx - 1) - 1.f : 0.f;
        row[50] = Ty > 1 ? 2.f * ty / (Ty - 1) - 1.f : 0.f;
    }
    for (int f = tid; f < d.nf; f += blockDim.x) {
        const zd* D = v.Dblk + ((size_t)f * d.ngroups + g) * d.gmax * d.gmax;
        const zd* q = v.q + (size_t)f * d.ldN;
        const zd* einc = v.Einc + (size_t)f * d.ldN;
        const zd* R0 = v.R + (size_t)f * d.ne * d.ldN;
        zd c1 = zmk(0.0, 0.0), c2 = zmk(0.0, 0.0), c3 = zmk(0.0, 0.0);
        for (int c = 0; c < k; c++) {
            const zd dc = D[lp[c] * d.gmax + lp[c]];
            const zd cap = sg[c] > 0 ? zmk(1.0 + dc.x, dc.y) : zmk(1.0 - dc.x, -dc.y);
            c1 = zadd(c1, cap);
            zd rhs = q[t[c]];
            for (int c2i = 0; c2i < k; c2i++) {
                const zd e = einc[t[c2i]];
                const zd dpe = sg[c2i] > 0 ? zmk(-e.x, -e.y) : e;
                rhs = zfma(D[lp[c] * d.gmax + lp[c2i]], dpe, rhs);
            }
            const zd e = einc[t[c]];
            const zd dpe = sg[c] > 0 ? zmk(-e.x, -e.y) : e;
            c2 = zfma(R0[t[c]], dpe, c2);
            zd y = zdiv(rhs, cap);
            if (sg[c] < 0) y = zmk(-y.x, -y.y);
            c3 = zfma(R0[t[c]], y, c3);
        }
        if (k > 0) c1 = zscale(c1, 1.0 / k);
        c2 = zscale(c2, op.zfac);
        c3 = zscale(c3, op.zfac);
        const zd u = zadd(v.Eobj[f * d.ne], v.fc[f].vinc);
        const zd c4 = zmk(-u.x * op.zfac / op.z0, -u.y * op.zfac / op.z0);
        float* o = row + 51 + 8 * f;
        o[0] = (float)c1.x; o[1] = (float)c1.y; o[2] = (float)c2.x; o[3] = (float)c2.y;
        o[4] = (float)c3.x; o[5] = (float)c3.y; o[6] = (float)c4.x; o[7] = (float)c4.y;
    }
}

void launch_features(hipStream_t st, const Dims& d, const DevVecs& v, GroupTab gt, const int32_t* tile,
                     const int8_t* P, const int32_t* owners, const int32_t* batch, int nb, int Tx, int Ty,
                     const ObjParams& op, float* out) {
    hipLaunchKernelGGL(features_kernel, dim3(nb), dim3(64), 0, st, d, v, gt, tile, P, owners, batch, Tx, Ty, op, out);
}

// Labels: |S11_f(newEobj)|^2 - |S11_f(Eobj)|^2 for every (candidate, frequency).
__global__ void labels_kernel(Dims d, DevVecs v, ObjParams op, const zd* __restrict__ eobj_out, int nb, float* out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nb * d.nf) return;
    const int b = e / d.nf, f = e - b * d.nf;
    auto s11sq = [&](zd E0) {
        const zd u = zadd(E0, v.fc[f].vinc);
        const zd zin = zmk(-u.x * op.zfac, -u.y * op.zfac);
        return zabs2(zdiv(zmk(zin.x - op.z0, zin.y), zmk(zin.x + op.z0, zin.y)));
    };
    out[e] = (float)(s11sq(eobj_out[((size_t)b * d.nf + f) * d.ne]) - s11sq(v.Eobj[f * d.ne]));
}

void launch_labels(hipStream_t st, const Dims& d, const DevVecs& v, const ObjParams& op, const zd* eobj_out, int nb,
                   float* out) {
    hipLaunchKernelGGL(labels_kernel, dim3((nb * d.nf + 255) / 256), dim3(256), 0, st, d, v, op, eobj_out, nb, out);
}

// ------------------------------------------------------------------ init -----
// Synthetic G: dyadic Green's function of a homogeneous eps_r medium, softened
// (R^2 -> R^2 + a^2), times -i (f / f_ref) and a complex scale.
//   G_mn = scale * g(R) [ t1 (u_m.u_n) + t2 (u_m.R)(u_n.R) / R^2 ],
//   g = exp(-i k R) / (4 pi R),  t1 = 1 - i/(kR) - 1/(kR)^2,  t2 = -1 + 3i/(kR) + 3/(kR)^2
// pos[m] = (x, y, orientation) in metres (orientation 0 = x, 1 = y).
__global__ void synth_g_kernel(int N, int ldN, const double* __restrict__ pos, double kr, zd scale, double soft2,
                               int row_major, zd* out) {
    __shared__ double pc[32][3];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 8;
    const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
    if (ty == 0 && c0 + tx < N) { pc[tx][0] = pos[3 * (c0 + tx)]; pc[tx][1] = pos[3 * (c0 + tx) + 1]; pc[tx][2] = pos[3 * (c0 + tx) + 2]; }
    __syncthreads();
    const int r = r0 + ty, c = c0 + tx;
    if (r >= N || c >= N) return;
    double pr[3] = {pos[3 * r], pos[3 * r + 1], pos[3 * r + 2]};
    zd v = zmul(synth_entry(pr, pc[tx], kr, soft2), scale);
    // element (r, c); row-major puts row r contiguous, column-major puts column c contiguous
    if (row_major) out[(size_t)r * ldN + c] = v;
    else out[(size_t)c * ldN + r] = v;
}

void launch_synth_g(hipStream_t st, int N, int ldN, const double* pos, double kr, zd scale, double soft,
                    int row_major, zd* out) {
    dim3 grid((N + 31) / 32, (N + 7) / 8);
    hipLaunchKernelGGL(synth_g_kernel, grid, dim3(32, 8), 0, st, N, ldN, pos, kr, scale, soft * soft, row_major, out);
}

// C = I - P + P G, column-major (pngfopt.hpp:196-233)
template <class Z>
__global__ void build_C_kernel(int N, int ldN, const int8_t* __restrict__ P, const Z* __restrict__ G, Z* C) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x, col = blockIdx.y;
    if (row >= N) return;
    const size_t i = (size_t)col * ldN + row;
    C[i] = P[row] ? G[i] : zc<Z>(row == col ? 1 : 0, 0);
}

template <class Z>
void launch_build_C(hipStream_t st, int N, int ldN, const int8_t* P, const Z* Gcm, Z* C) {
    hipLaunchKernelGGL(build_C_kernel<Z>, dim3((N + 255) / 256, N), dim3(256), 0, st, N, ldN, P, Gcm, C);
}

template <class Z>
__global__ void add_diag_kernel(int N, int ldN, Z* A, double v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) A[(size_t)i * ldN + i].x += (typename zreal<Z>::type)v;
}

template <class Z>
void launch_add_diag(hipStream_t st, int N, int ldN, Z* A, double v) {
    hipLaunchKernelGGL(add_diag_kernel<Z>, dim3((N + 255) / 256), dim3(256), 0, st, N, ldN, A, v);
}

__global__ void transpose_kernel(int N, int ldN, const zd* __restrict__ in, zd* __restrict__ out) {
    __shared__ zd tileb[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    for (int j = threadIdx.y; j < 32; j += 8) {
        const int r = bx + threadIdx.x, c = by + j;
        if (r < N && c < N) tileb[j][threadIdx.x] = in[(size_t)c * ldN + r];
    }
    __syncthreads();
    for (int j = threadIdx.y; j < 32; j += 8) {
        const int r = bx + j, c = by + threadIdx.x;
        if (r < N && c < N) out[(size_t)r * ldN + c] = tileb[threadIdx.x][j];
    }
}

void launch_transpose(hipStream_t st, int N, int ldN, const zd* in, zd* out) {
    dim3 grid((N + 31) / 32, (N + 31) / 32);
    hipLaunchKernelGGL(transpose_kernel, grid, dim3(32, 8), 0, st, N, ldN, in, out);
}

// out[0] = max |G_ij - G_ji|, out[1] = max |G_ij| over the column-major N x N matrix
// (|z| = |Re z| + |Im z|), kept as the bit patterns of non-negative doubles, whose order is
// that of the unsigned integers.  out must hold zeros before the launch.
__global__ void __launch_bounds__(256) asym_kernel(int N, int ldN, const zd* __restrict__ G,
                                                   unsigned long long* out) {
    __shared__ zd t1[32][33], t2[32][33];
    __shared__ double red[2][4];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    for (int j = threadIdx.y; j < 32; j += 8) {
        const int r1 = bx + threadIdx.x, c1 = by + j;  // t1[j][x] = G(bx + x, by + j)
        t1[j][threadIdx.x] = (r1 < N && c1 < N) ? G[(size_t)c1 * ldN + r1] : zmk(0.0, 0.0);
        const int r2 = by + threadIdx.x, c2 = bx + j;  // t2[j][x] = G(by + x, bx + j)
        t2[j][threadIdx.x] = (r2 < N && c2 < N) ? G[(size_t)c2 * ldN + r2] : zmk(0.0, 0.0);
    }
    __syncthreads();
    double dmax = 0.0, amax = 0.0;
    for (int j = threadIdx.y; j < 32; j += 8) {
        const zd a = t1[j][threadIdx.x], b = t2[threadIdx.x][j];  // G(bx+x, by+j), G(by+j, bx+x)
        dmax = fmax(dmax, zabs1(zsub(a, b)));
        amax = fmax(amax, zabs1(a));
    }
    for (int off = 32; off > 0; off >>= 1) {
        dmax = fmax(dmax, __shfl_xor(dmax, off, 64));
        amax = fmax(amax, __shfl_xor(amax, off, 64));
    }
    const int tid = threadIdx.y * 32 + threadIdx.x;
    if ((tid & 63) == 0) { red[0][tid >> 6] = dmax; red[1][tid >> 6] = amax; }
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < 4; w++) { dmax = fmax(dmax, red[0][w]); amax = fmax(amax, red[1][w]); }
        atomicMax(&out[0], (unsigned long long)__double_as_longlong(dmax));
        atomicMax(&out[1], (unsigned long long)__double_as_longlong(amax));
    }
}

void launch_asym(hipStream_t st, int N, int ldN, const zd* G, unsigned long long* out) {
    dim3 grid((N + 31) / 32, (N + 31) / 32);
    hipLaunchKernelGGL(asym_kernel, grid, dim3(32, 8), 0, st, N, ldN, G, out);
}

// 64-bit mixer (splitmix64 finaliser) of the perturbation below.
__host__ __device__ inline unsigned long long mix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// G(i, j) += amp (u1 + i u2), u1, u2 uniform in [-1, 1) from the top 53 bits of
// mix64(seed * 2^40 + (i N + j) * 2 + {0, 1}): a deterministic perturbation with G(i, j) and
// G(j, i) drawn independently, i.e. a non-reciprocal one.  Column-major, ld ldN.
__global__ void __launch_bounds__(256) perturb_kernel(int N, int ldN, zd* G, double amp, unsigned long long seed) {
    // product and sum rounded separately, as on the host (the default contraction would fuse them)
#pragma clang fp contract(off)
    for (int col = blockIdx.y; col < N; col += gridDim.y)
        for (int r = blockIdx.x * 256 + threadIdx.x; r < N; r += gridDim.x * 256) {
            const unsigned long long key = (seed << 40) + 2ull * ((unsigned long long)r * N + col);
            const double u1 = (double)(mix64(key) >> 11) * 0x1.0p-52 - 1.0;
            const double u2 = (double)(mix64(key + 1) >> 11) * 0x1.0p-52 - 1.0;
            zd& g = G[(size_t)col * ldN + r];
            g.x = g.x + amp * u1;
            g.y = g.y + amp * u2;
        }
}

void launch_perturb(hipStream_t st, int N, int ldN, zd* G, double amp, unsigned long long seed) {
    const int gx = (N + 255) / 256 < 8 ? (N + 255) / 256 : 8;
    dim3 grid(gx, N < 4096 ? N : 4096);
    hipLaunchKernelGGL(perturb_kernel, grid, dim3(256), 0, st, N, ldN, G, amp, seed);
}

// *flag |= 1 if any entry of the column-major N x N matrix is NaN or infinite (rows >= N of
// each column are padding and are not read).  grid.y strides over the columns.
__global__ void __launch_bounds__(256) nonfinite_kernel(int N, int ldN, const zd* __restrict__ G, int32_t* flag) {
    int bad = 0;
    for (int col = blockIdx.y; col < N; col += gridDim.y)
        for (int r = blockIdx.x * 256 + threadIdx.x; r < N; r += gridDim.x * 256) {
            const zd v = G[(size_t)col * ldN + r];
            bad |= !(isfinite(v.x) && isfinite(v.y));
        }
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

void launch_nonfinite(hipStream_t st, int N, int ldN, const zd* G, int32_t* flag) {
    const int gx = (N + 255) / 256 < 8 ? (N + 255) / 256 : 8;
    dim3 grid(gx, N < 4096 ? N : 4096);
    hipLaunchKernelGGL(nonfinite_kernel, grid, dim3(256), 0, st, N, ldN, G, flag);
}

// W = (W + W^T) / 2 in place (N x N, leading dimension ldN).  The workgroup of tile pair
// (bx, by), bx <= by, owns both tiles: it stages them in LDS and writes both back, and the two
// entries of a pair get the same value (addition commutes), so W ends exactly symmetric.
template <class Z>
__global__ void __launch_bounds__(256) symmetrize_kernel(int N, int ldN, Z* W) {
    typedef typename zreal<Z>::type R;
    if (blockIdx.x > blockIdx.y) return;
    __shared__ Z t1[32][33], t2[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    for (int j = threadIdx.y; j < 32; j += 8) {
        const int r1 = bx + j, c1 = by + threadIdx.x;  // t1[j][x] = W[bx + j][by + x]
        t1[j][threadIdx.x] = (r1 < N && c1 < N) ? W[(size_t)r1 * ldN + c1] : zc<Z>(0, 0);
        const int r2 = by + j, c2 = bx + threadIdx.x;  // t2[j][x] = W[by + j][bx + x]
        t2[j][threadIdx.x] = (r2 < N && c2 < N) ? W[(size_t)r2 * ldN + c2] : zc<Z>(0, 0);
    }
    __syncthreads();
    for (int j = threadIdx.y; j < 32; j += 8) {
        const int r1 = bx + j, c1 = by + threadIdx.x;
        const Z a = t1[j][threadIdx.x], b = t2[threadIdx.x][j];  // W[r1][c1], W[c1][r1]
        const Z m = zc<Z>(R(0.5) * (a.x + b.x), R(0.5) * (a.y + b.y));
        if (r1 < N && c1 < N) W[(size_t)r1 * ldN + c1] = m;
        const int r2 = by + j, c2 = bx + threadIdx.x;
        const Z a2 = t2[j][threadIdx.x], b2 = t1[threadIdx.x][j];  // W[r2][c2], W[c2][r2]
        const Z m2 = zc<Z>(R(0.5) * (a2.x + b2.x), R(0.5) * (a2.y + b2.y));
        if (r2 < N && c2 < N) W[(size_t)r2 * ldN + c2] = m2;
    }
}

template <class Z>
void launch_symmetrize(hipStream_t st, int N, int ldN, Z* W) {
    dim3 grid((N + 31) / 32, (N + 31) / 32);
    hipLaunchKernelGGL(symmetrize_kernel<Z>, grid, dim3(32, 8), 0, st, N, ldN, W);
}

// Gobj row 0 = sum_p G(p, :),  Einc = sum_q G(:, q)   (pngfopt.hpp:584-595)
template <class Z>
__global__ void port_sums_kernel(int N, int ldN, const Z* __restrict__ G, const int32_t* __restrict__ port, int np,
                                 Z* gobj0, Z* einc) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    Z a = zc<Z>(0, 0), b = zc<Z>(0, 0);
    for (int i = 0; i < np; i++) {
        const int p = port[i];
        a = zadd(a, G[(size_t)n * ldN + p]);
        b = zadd(b, G[(size_t)p * ldN + n]);
    }
    gobj0[n] = a;
    einc[n] = b;
}

template <class Z>
void launch_port_sums(hipStream_t st, int N, int ldN, const Z* Gcm, const int32_t* port, int np, Z* gobj0,
                      Z* einc) {
    hipLaunchKernelGGL(port_sums_kernel<Z>, dim3((N + 255) / 256), dim3(256), 0, st, N, ldN, Gcm, port, np, gobj0,
                       einc);
}

template <class Z>
__global__ void pe_kernel(int N, const int8_t* __restrict__ P, const Z* __restrict__ einc, Z* pe) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    Z e = einc[n];
    pe[n] = P[n] ? zc<Z>(-e.x, -e.y) : zc<Z>(0, 0);
}

template <class Z>
void launch_pe(hipStream_t st, int N, const int8_t* P, const Z* einc, Z* pe) {
    hipLaunchKernelGGL(pe_kernel<Z>, dim3((N + 255) / 256), dim3(256), 0, st, N, P, einc, pe);
}

// out[r] = sum_n A[r][n] x[n]   (rows of length N at stride ldN; fixed-order tree)
template <class Z>
__global__ void dot_rows_kernel(int N, int ldN, const Z* __restrict__ A, const Z* __restrict__ x, zd* out) {
    __shared__ Z red[256];
    const int r = blockIdx.x, tid = threadIdx.x;
    Z s = zc<Z>(0, 0);
    for (int n = tid; n < N; n += 256) s = zfma(A[(size_t)r * ldN + n], x[n], s);
    red[tid] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off) red[tid] = zadd(red[tid], red[tid + off]);
        __syncthreads();
    }
    if (tid == 0) out[r] = zwide(red[0]);
}

template <class Z>
void launch_dot_rows(hipStream_t st, int N, int ldN, int rows, const Z* A, const Z* x, zd* out) {
    hipLaunchKernelGGL(dot_rows_kernel<Z>, dim3(rows), dim3(256), 0, st, N, ldN, A, x, out);
}

// out = (complex64) in, rounded to nearest, over the N x N column-major matrix (ld ldN)
__global__ void cast_c64_kernel(int N, int ldN, const zd* __restrict__ in, zf* out) {
    for (int col = blockIdx.y; col < N; col += gridDim.y)
        for (int r = blockIdx.x * 256 + threadIdx.x; r < N; r += gridDim.x * 256) {
            const zd v = in[(size_t)col * ldN + r];
            out[(size_t)col * ldN + r] = zmkf((float)v.x, (float)v.y);
        }
}

void launch_cast_c64(hipStream_t st, int N, int ldN, const zd* in, zf* out) {
    const int gx = (N + 255) / 256 < 8 ? (N + 255) / 256 : 8;
    dim3 grid(gx, N < 4096 ? N : 4096);
    hipLaunchKernelGGL(cast_c64_kernel, grid, dim3(256), 0, st, N, ldN, in, out);
}


// Current W over an arbitrary group's components (flips outside the driver's
// group table): D[a][b] = W0[c_a][c_b] - sum_K' A_all[c_a][K'] M_all[K'][c_b].
template <class Z>
__global__ void lr_block_current_kernel(Dims d, DevTabsT<Z> tb, GroupTab gt, int g, int K, Z* Dblk) {
    const int f = blockIdx.y;
    const int nc = gt.ncomp[g];
    const int32_t* cl = gt.comps + (size_t)g * gt.gmax;
    const Z* W0 = tb.W0[f];
    const Z* A = tb.Aall[f];
    Z* D = Dblk + ((size_t)f * d.ngroups + g) * gt.gmax * gt.gmax;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nc * nc; e += gridDim.x * blockDim.x) {
        const int a = e / nc, b = e - a * nc;
        Z s = W0[(size_t)cl[a] * d.ldN + cl[b]];
        for (int kk = 0; kk < K; kk++) s = zfms(A[(size_t)cl[a] * d.Kmax + kk], m_entry(tb, d, f, kk, cl[b]), s);
        D[a * gt.gmax + b] = s;
    }
}

template <class Z>
void launch_lr_block_current(hipStream_t st, const Dims& d, const DevTabsT<Z>& tb, GroupTab gt, int g, int K, Z* Dblk) {
    hipLaunchKernelGGL(lr_block_current_kernel<Z>, dim3(36, d.nf), dim3(256), 0, st, d, tb, gt, g, K, Dblk);
}

// D[g][a][b] = W0[comp_g[a]][comp_g[b]]
template <class Z>
__global__ void extract_blocks_kernel(Dims d, const Z* __restrict__ W0, GroupTab gt, Z* Dblk) {
    const int g = blockIdx.x;
    const int nc = gt.ncomp[g];
    const int32_t* cl = gt.comps + (size_t)g * gt.gmax;
    Z* D = Dblk + (size_t)g * gt.gmax * gt.gmax;
    for (int e = threadIdx.x; e < nc * nc; e += blockDim.x) {
        const int a = e / nc, b = e - a * nc;
        D[a * gt.gmax + b] = W0[(size_t)cl[a] * d.ldN + cl[b]];
    }
}

template <class Z>
void launch_extract_blocks(hipStream_t st, const Dims& d, const Z* W0, GroupTab gt, Z* Dblk_f) {
    hipLaunchKernelGGL(extract_blocks_kernel<Z>, dim3(d.ngroups), dim3(256), 0, st, d, W0, gt, Dblk_f);
}

// ------------------------------------------------- explicit instantiations -----
// complex128 (every method) and complex64 (the LOWRANK state of the single-precision mode)
#define PNGF_INST(Z)                                                                                             \
    template void launch_lr_eval<Z>(hipStream_t, const Dims&, const DevVecsT<Z>&, GroupTab, const int32_t*,        \
                                    const int8_t*, const int32_t*, const int32_t*, int, zd*, int32_t*, Z*,        \
                                    const int32_t*);                                                              \
    template void launch_lr_accept_prep<Z>(hipStream_t, const Dims&, const DevVecsT<Z>&, int, const int32_t*,      \
                                           const int32_t*, const int32_t*, int, Z*, int, int32_t*, Z*,            \
                                           const int32_t*, const int32_t*);                                       \
    template void launch_lr_gather<Z>(hipStream_t, const Dims&, const DevTabsT<Z>&, const int32_t*, int, int, Z*, \
                                      Z*, int, const int32_t*);                                                   \
    template void launch_lr_build_m<Z>(hipStream_t, const Dims&, const DevTabsT<Z>&, int, int, int, Z*, int);      \
    template void launch_apply_capinv<Z>(hipStream_t, const Dims&, const Z*, int, const Z*, size_t, int, int,      \
                                         const int32_t*, Z* const*, Z*, size_t, int, int, const Z* const*, int);   \
    template void launch_lr_small<Z>(hipStream_t, const Dims&, const DevTabsT<Z>&, const DevVecsT<Z>&, const Z*,  \
                                     int, int, int);                                                              \
    template void launch_lr_finish<Z>(hipStream_t, const Dims&, const DevTabsT<Z>&, const DevVecsT<Z>&, GroupTab, \
                                      const Z*, int, const Z*, size_t, const int32_t*, int, int, const int32_t*,  \
                                      const int32_t*, int, int, FlipUpdate);                                      \
    template void launch_lr_blocks<Z>(hipStream_t, const Dims&, const DevTabsT<Z>&, const DevVecsT<Z>&, GroupTab, \
                                      int, int);                                                                  \
    template void launch_build_C<Z>(hipStream_t, int, int, const int8_t*, const Z*, Z*);                          \
    template void launch_add_diag<Z>(hipStream_t, int, int, Z*, double);                                          \
    template void launch_symmetrize<Z>(hipStream_t, int, int, Z*);                                                 \
    template void launch_port_sums<Z>(hipStream_t, int, int, const Z*, const int32_t*, int, Z*, Z*);               \
    template void launch_pe<Z>(hipStream_t, int, const int8_t*, const Z*, Z*);                                     \
    template void launch_dot_rows<Z>(hipStream_t, int, int, int, const Z*, const Z*, zd*);                         \
    template void launch_lr_block_current<Z>(hipStream_t, const Dims&, const DevTabsT<Z>&, GroupTab, int, int, Z*); \
    template void launch_extract_blocks<Z>(hipStream_t, const Dims&, const Z*, GroupTab, Z*);
PNGF_INST(zd)
PNGF_INST(zf)
#undef PNGF_INST

}  // namespace pngf
