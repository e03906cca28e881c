/*
 * pngf_gpu.h -- C ABI of the MI355X (gfx950) implementation of the pngf-opt
 * direct-binary-search (DBS) Woodbury hot path.
 *
 * One handle holds every frequency of one design problem (the reference keeps
 * one PNGF_Optimizer object per frequency, pngf-opt/main.cpp:166,224; here the
 * frequency axis is a batch axis on the device).  All functions return an int
 * status (PNGF_OK = 0, negative on error) and never exit or throw; the text of
 * the last error is available from pngf_last_error().  Host arrays passed in are
 * copied during the call and never retained.  A handle is used by one host
 * thread at a time and owns one HIP stream; all its device work is ordered on
 * that stream.
 *
 * Complex numbers are pairs of doubles (re, im), layout-compatible with
 * std::complex<double> and numpy complex128.
 */
#ifndef PNGF_GPU_H
#define PNGF_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pngf_handle pngf_handle;
typedef struct { double re, im; } pngf_c128;

/* status codes */
#define PNGF_OK            0
#define PNGF_ERR_ARG      -1  /* bad argument (null pointer, index out of range, ...) */
#define PNGF_ERR_IO       -2  /* G file missing or unreadable */
#define PNGF_ERR_SIZE     -3  /* G file size does not match (N+ne-1)^2 complex128 */
#define PNGF_ERR_SINGULAR -4  /* C or a capacitance matrix is singular (LU info > 0) */
#define PNGF_ERR_HIP      -5  /* HIP / rocBLAS / rocSOLVER error, or no GPU */
#define PNGF_ERR_NOMEM    -6  /* device allocation failed or exceeds mem_limit */
#define PNGF_ERR_STATE    -7  /* call out of order (e.g. eval before init_state) */
#define PNGF_ERR_NONFINITE -8 /* G contains NaN or Inf */

/* component layout of the design region */
#define PNGF_LAYOUT_SHARED   0 /* reference Yee grid, shared tile edges (pngfopt.hpp:86-94,624-664) */
#define PNGF_LAYOUT_DISJOINT 1 /* 24 components owned by each tile + separate port components */

/* objective reduced over frequencies */
#define PNGF_OBJ_REF 0 /* sum_f 200|S11_f|^2 + (15 - D_f)^2 (main.cpp:277-289); needs num_eobj = 3 */
#define PNGF_OBJ_S11 1 /* mean_f |S11_f|^2 */

/* evaluation / update method */
#define PNGF_METHOD_PLAIN   0 /* reference formulation: G rows and C^-1 columns per candidate
                                 (pngfopt.hpp:379-539), rank-k update of C^-1 on accept (256-367) */
#define PNGF_METHOD_LOWRANK 1 /* state W = (G-I) C^-1 kept as W0 - A M (deferred rank-k factors);
                                 candidates scored from k x k blocks of W in O(k^3) */

/* accept rule inside the raster driver */
#define PNGF_ORDER_FIRST 0 /* first improving flip in visit order (the reference loop) */
#define PNGF_ORDER_BEST  1 /* best improving flip of a full exact sweep (steepest descent); a
                              different accepted sequence, reported beside the reference one */

/* LOWRANK accept: for a complex-symmetric G (reciprocal medium) W = (G-I)(I+P(G-I))^-1 is
   symmetric too, so the row strip W[t,:] that an accept needs is the transposed column strip
   and one of the two N x K x k products is skipped */
#define PNGF_SYM_AUTO 0 /* use it when every G is symmetric to 1e-14 of its largest entry */
#define PNGF_SYM_OFF  1 /* always form both strips */
#define PNGF_SYM_PROJECT 2 /* opt-in, not the reference's results: a G whose max|G - G^T| is at most
                              1e-5 of its largest entry (acmefdfd's single-precision MUMPS output,
                              widened, is symmetric to ~1e-7) is replaced by (G + G^T) / 2 before
                              the state is built, and the accept then uses W = W^T; a G further from
                              symmetric keeps both strips.  The results are those of the projected G. */

/* LOWRANK deferred factors W = W0 - A M: how the M side is held */
#define PNGF_FACTORS_AUTO    0 /* stored M, unless device memory caps its capacity below the wanted
                                  number of accepts and PNGF_FACTORS_REBUILT applies */
#define PNGF_FACTORS_STORED  1 /* M stored: one A column and one M row per factor column */
#define PNGF_FACTORS_REBUILT 2 /* symmetric W, flips of at most 32 components: M rows rebuilt from A
                                  and the k x k Lam' of their accept (M = Lam' A^T), so a factor
                                  column costs one A column: twice the capacity, a slower gather */

/* who drives the raster first-improvement loop (same accepted sequence either way) */
#define PNGF_LOOP_AUTO 0 /* the device: batch, selection and accept bookkeeping run in kernels and the
                            host synchronises once per 16 iterations (LOWRANK, no training log, no
                            history, flips of at most 32 components); otherwise the host */
#define PNGF_LOOP_HOST 1 /* the host: one synchronisation per scored batch */

/* arithmetic of the optimizer state (LOWRANK) */
#define PNGF_PREC_C128 0 /* complex128 throughout (the reference's default build) */
#define PNGF_PREC_C64  1 /* complex64, as the reference's USE_SINGLE_PRECISION build (main.cpp:58-72): G cast
                            to c64 at init (pngfopt.hpp:128-156), C inverted by cgetrf + cgetri, W, the factors,
                            the blocks, q and R in c64, capacitance solves in FP32, the strips on FP32 MFMA with
                            FP32 accumulation; Eobj (ne scalars per frequency) and the objective stay FP64 */

/* DBS visit order */
#define PNGF_DRIVER_REF    0 /* MagicHat order, mirror-symmetric 2/4-tile flips (main.cpp:338-459) */
#define PNGF_DRIVER_RASTER 1 /* raster order, single-tile flips, cyclic until one full cycle
                                has no improving flip */

typedef struct {
    int32_t layout;       /* PNGF_LAYOUT_* */
    int32_t tiles_x;      /* tiles in x (reference num_tiles_x) */
    int32_t tiles_y;      /* tiles in y */
    int32_t cells_x;      /* Yee cells per tile in x (reference num_cells_per_tile_x; 3) */
    int32_t cells_y;      /* Yee cells per tile in y (3) */
    int32_t num_eobj;     /* rows of Gobj: 1 (port voltage) or 3 (+ two NF2FF rows) */
    int32_t max_flips;    /* tiles per flip: 1 for the raster driver, 4 for the reference driver */
    int32_t num_freq;     /* frequencies held by the handle */
    int32_t method;       /* PNGF_METHOD_* */
    int32_t objective;    /* PNGF_OBJ_* */
    int32_t driver;       /* PNGF_DRIVER_*: fixes the candidate groups (flip sets) */
    int32_t max_deferred; /* LOWRANK: accepted flips kept as factors before they are folded
                             into W0 (0 = N / flip size, the point beyond which folding pays,
                             capped by free device memory) */
    int32_t device;       /* HIP device ordinal */
    int32_t tile_variant; /* LOWRANK skinny-product tile shape: 0 = autotune at init, v+1 = use v */
    double dx, dy, dz;    /* Yee cell size used by the port impedance (main.cpp:157-160) */
    double z0;            /* reference impedance for S11 (50 ohm, main.cpp:279) */
    int64_t mem_limit;    /* refuse device allocations beyond this many bytes (0 = no limit) */
    double synth_soft;    /* synthetic G only: softening radius in cells (0 = 1.0); smaller values
                             give better-conditioned C */
    int32_t symmetry;     /* LOWRANK: PNGF_SYM_AUTO, PNGF_SYM_OFF or PNGF_SYM_PROJECT */
    int32_t factors;      /* LOWRANK: PNGF_FACTORS_* */
    int32_t precision;    /* PNGF_PREC_* (PNGF_PREC_C64 needs PNGF_METHOD_LOWRANK) */
} pngf_config;

/* Creates a handle; selects the device and allocates nothing large yet.
   Replaces the PNGF_Optimizer constructor / init bookkeeping (pngfopt.hpp:66-119). */
int pngf_create(const pngf_config* cfg, pngf_handle** out);
int pngf_destroy(pngf_handle* h);
const char* pngf_last_error(const pngf_handle* h);
/* Number of design components N for a layout (pngfopt.hpp:86-94 for SHARED). */
int pngf_layout_size(int32_t layout, int32_t tiles_x, int32_t tiles_y, int32_t cells_x, int32_t cells_y);

/* Frequency of slot f in Hz: k0 = 2 pi f / c0 of the directivity term (main.cpp:162-164, one
   PNGF_Optimizer per frequency there) and the synthetic G's frequency. */
int pngf_set_frequency(pngf_handle* h, int32_t f, double hz);

/* Reads acmefdfd's raw G file for frequency slot f: M x M complex128, column-major,
   M = N + num_eobj - 1, no header.  G = file[0:N,0:N], Gobj rows 1.. = file[N.., 0:N].
   Replaces PNGF_Optimizer::init's fread (pngfopt.hpp:121-184); adds a size check. */
int pngf_load_g(pngf_handle* h, int32_t f, const char* path);
/* Same contents from a host buffer in the file layout.  No counterpart in the reference, which
   reads files only. */
int pngf_set_g(pngf_handle* h, int32_t f, const pngf_c128* file_buf);
/* Synthetic stand-in for acmefdfd's G (dyadic Green's function of an eps_r = 3.5
   medium sampled at the component positions, softened self term), generated on the
   device for slot f at its frequency.  See pngf_synth_g_host for the definition.  No
   counterpart in the reference, where acmefdfd (FDFD + MUMPS, host) produces G. */
int pngf_synth_g(pngf_handle* h, int32_t f);
/* Adds a deterministic non-reciprocal perturbation to the G of slot f (before pngf_init_state):
   G(i, j) += rel * max|G| * (u1 + i u2), with u1, u2 uniform in [-1, 1) drawn from a hash of
   (seed, i, j) (|z| = |Re z| + |Im z|; u = (mix64(seed * 2^40 + 2 (i N + j) + c) >> 11) * 2^-52 - 1,
   c = 0, 1, mix64 the splitmix64 finaliser).  A stand-in for acmefdfd's G, whose single-precision
   MUMPS Schur complement (acmefdfd/apf_solver.hpp:188-198) is reciprocal only to rounding.  No
   counterpart in the reference. */
int pngf_perturb_g(pngf_handle* h, int32_t f, double rel, int32_t seed);
/* Host version of the synthetic generator, written in the acmefdfd file layout
   (M x M column-major).  The device version evaluates the same formula.  No counterpart in
   the reference. */
int pngf_synth_g_host(int32_t layout, int32_t tiles_x, int32_t tiles_y, int32_t cells_x,
                      int32_t cells_y, int32_t num_eobj, double hz, double soft_cells,
                      pngf_c128* file_buf);

/* reference: PNGF_Optimizer::set_tile_map (pngfopt.hpp:669-674) */
int pngf_set_tile_map(pngf_handle* h, const int32_t* tile_map);
/* reference: PNGF_Optimizer::set_source_port (pngfopt.hpp:559-618); returns
   PNGF_ERR_ARG where the reference calls exit(0). */
int pngf_set_source_port(pngf_handle* h, int32_t tile_x, int32_t tile_y, int32_t y_offset, int32_t y_size);

/* compute_C_matrix + LinAlg::invertMatrix + compute_PEinc + compute_R_matrix(1) for every
   frequency (pngfopt.hpp:196-250, linalg.hpp:34-85, main.cpp:244-252).  LOWRANK also
   forms W = (G-I) C^-1 and releases G and C^-1. */
int pngf_init_state(pngf_handle* h);

/* reference: compute_updated_Eobj (pngfopt.hpp:379-539) at every frequency, plus the
   objective of main.cpp:401-423.  new_eobj: [num_freq][num_eobj] (may be null); F may be null. */
int pngf_eval(pngf_handle* h, const int32_t* tiles, int32_t ntiles, pngf_c128* new_eobj, double* F);
/* Scores ncand flips against the same state in one batch.  tiles: [ncand][max_flips],
   ntiles: [ncand].  F_out: [ncand]; eobj_out: [ncand][num_freq][num_eobj] (may be null).
   The batched form of compute_updated_Eobj (pngfopt.hpp:379-539) and main.cpp:401-423, which
   the reference calls one flip and one frequency at a time. */
int pngf_eval_batch(pngf_handle* h, const int32_t* tiles, const int32_t* ntiles, int32_t ncand,
                    double* F_out, pngf_c128* eobj_out);
/* First improvement over a visit sequence against the current state: the sequential loop of
   main.cpp:338-459 without its accept (`if (F < F_best)`, strict, main.cpp:429), scored in
   batches.  tiles/ntiles as in pngf_eval_batch.  *accepted_idx = index of the first flip with
   F < F_best, or -1; *F_new = its objective (F_best if none; may be null); *evaluated = flips up
   to and including it, or ncand (the reference's flip_count increments; may be null).  The flip
   is not committed: pngf_accept does that. */
int pngf_sweep(pngf_handle* h, const int32_t* tiles, const int32_t* ntiles, int32_t ncand, double F_best,
               int32_t* accepted_idx, double* F_new, int64_t* evaluated);
/* Commits a flip: update_C_inv + compute_PEinc + compute_R_matrix(0) at every frequency
   (pngfopt.hpp:256-367, 542-552, 237-250).  The flip is named explicitly instead of
   relying on the last compute_updated_Eobj call. */
int pngf_accept(pngf_handle* h, const int32_t* tiles, int32_t ntiles);

/* Current state.  Any pointer may be null.  eobj: [num_freq][num_eobj]; vinc: [num_freq];
   ln: [num_freq][2] (LN1inc, LN2inc); s11: [num_freq]; F: objective.  Replaces main.cpp's
   reads of the public members Eobj, Vinc, LN1inc, LN2inc (main.cpp:271-305, pngfopt.hpp:21-50). */
int pngf_get_state(pngf_handle* h, pngf_c128* eobj, pngf_c128* vinc, pngf_c128* ln,
                   pngf_c128* s11, double* F);
/* reference: the public member tile_map (pngfopt.hpp:21-50), T x T, x fastest. */
int pngf_get_tile_map(pngf_handle* h, int32_t* tile_map);
/* C^-1 of slot f, column-major N x N (PLAIN only: LOWRANK does not keep it).  The reference's
   public member C, which holds C^-1 after main.cpp:246. */
int pngf_get_cinv(pngf_handle* h, int32_t f, pngf_c128* out);
/* Current W = (G-I) C^-1 of slot f, row-major N x N (LOWRANK only; folds pending factors).  No
   counterpart in the reference, which keeps C^-1 instead. */
int pngf_get_w(pngf_handle* h, int32_t f, pngf_c128* out);

/* G of slot f, column-major N x N (PLAIN only: LOWRANK releases G after init).  The reference's
   member G (pngfopt.hpp:21-50). */
int pngf_get_g(pngf_handle* h, int32_t f, pngf_c128* out);

/* Device-side snapshot of the mutable optimizer state (C^-1 or W, S or q, R, Eobj,
   mask, tile map) and return to it.  No counterpart in the reference (which keeps no
   checkpoint, main.cpp:469-472); used to repeat a DBS run from one start and to
   resume.  The first call allocates one more copy of the N x N state per frequency. */
int pngf_checkpoint(pngf_handle* h);
int pngf_restore(pngf_handle* h);

typedef struct {
    int64_t max_evals;     /* stop after this many evaluations (0 = run to convergence) */
    int32_t seed;          /* MagicHat seed for PNGF_DRIVER_REF (reference default -514213645) */
    int32_t max_accepts;   /* capacity of the arrays below */
    int32_t* acc_tiles;    /* [max_accepts] primary tile of each accepted flip (may be null) */
    int64_t* acc_eval;     /* [max_accepts] flip_count at which it was accepted (may be null) */
    double* acc_F;         /* [max_accepts] objective after each accept (may be null) */
    int32_t batch_min;     /* PLAIN: first speculative batch after an accept (0 = default) */
    int32_t batch_max;     /* PLAIN: largest speculative batch (0 = default) */
    int32_t rng_discard;   /* PNGF_DRIVER_REF: outputs of the hat's mt19937 already consumed before
                              the loop (main.cpp:195 draws (T/2+1)^2 bits for INIT_RANDOM; 0 for
                              the default INIT_EMPTY) */
    int32_t order;         /* PNGF_ORDER_FIRST (reference semantics, main.cpp:429) or PNGF_ORDER_BEST */
    int32_t loop;          /* PNGF_LOOP_AUTO or PNGF_LOOP_HOST */
    int32_t resume;        /* 1: continue from the driver position in driver_state (first-improvement
                              order) instead of the start of the visit order */
    /* Optional training log (LOWRANK): for every scored candidate, its ranker inputs
       (pngf_ranker_dims din floats) and the exact change of |S11_f|^2 at every frequency
       (dout floats), appended until log_cap rows. */
    float* log_x;
    float* log_y;
    int64_t log_cap;
    /* Optional optimization history, the records of the reference's Logger (logger.hpp:64-73;
       main.cpp:304-328, 401-426): record 0 is the start state (tile -1, decision -1), then one
       record per committed evaluation in order.  hist_tile = primary tile of the flip,
       hist_dec = 1 if F < F_best, hist_perf[i * (2 + 3 nf) + ...] = F, F_best before this
       evaluation, then per frequency Re S11, Im S11 and the directivity D (0 for
       PNGF_OBJ_S11).  Written until hist_cap records (pngf_dbs_run, first-improvement order;
       the loop then runs host-driven). */
    int32_t* hist_tile;
    int32_t* hist_dec;
    double* hist_perf;
    int64_t hist_cap;
    /* Optional driver position, PNGF_DRIVER_STATE_WORDS words (may be null): written when the run
       returns (first-improvement order), read first when resume = 1.  Raster driver: the place in
       the visit cycle and the count of non-improving flips since the last accept; reference
       driver: the MagicHat generator (std::mt19937 state), its remaining draws and the last tile
       tried (magichat.hpp:84-107, main.cpp:343-353).  With the tile map, this is all a run needs
       to continue on a handle rebuilt by pngf_init_state; the reference keeps no checkpoint
       (main.cpp:469-472). */
    uint32_t* driver_state;
} pngf_dbs_params;

#define PNGF_DRIVER_STATE_WORDS 2048

typedef struct {
    int64_t flip_count;        /* committed evaluations (main.cpp:458) */
    int64_t accepts;
    int64_t speculative_evals; /* evaluations scored after the accepted index (wasted) */
    int64_t batches;
    double F_start, F_final;
    double seconds;            /* wall time of the DBS loop (main.cpp:333-464 window) */
    double accept_seconds;     /* device time spent in accepts (event-timed) */
    int32_t flushes;           /* LOWRANK: times the deferred factors were folded into W0 */
    int32_t converged;         /* 1 if stopped because no flip improves */
    int64_t logged;            /* rows written to log_x / log_y */
    int64_t ranked_scored;     /* ranked driver: candidates scored by the network */
    int64_t full_sweeps;       /* ranked driver: exact full sweeps (hit-rate fallbacks) */
    int64_t ranked_steps;      /* ranked driver: steps that followed the network's order */
    int64_t ranked_hits;       /* ranked driver: those whose first top_k chunk held an improving flip */
    int64_t hist_count;        /* history records written (pngf_dbs_params.hist_*) */
} pngf_dbs_result;

/* Runs the DBS loop of main.cpp:338-459 on the device with the driver chosen at create.
   Every accepted flip is the first improving one in visit order, as in the sequential
   reference loop; candidates are scored in speculative batches against frozen state. */
int pngf_dbs_run(pngf_handle* h, const pngf_dbs_params* p, pngf_dbs_result* r);

/* Runs pngf_dbs_run on n independent handles concurrently (host worker threads; each
   handle's work stays on its own streams, so the device overlaps the instances).  For
   many DBS instances of uneven size on one GPU (reference: separate processes per run,
   main.cpp:173-175).  Returns the first non-zero status. */
int pngf_dbs_run_many(pngf_handle** hs, int32_t n, const pngf_dbs_params* ps, pngf_dbs_result* rs);
/* Same with an explicit number of host worker threads: the handles are taken largest
   (N x frequencies) first by `workers` threads, each running one handle at a time.
   pngf_dbs_run_many uses 4. */
int pngf_dbs_run_pool(pngf_handle** hs, int32_t n, const pngf_dbs_params* ps, pngf_dbs_result* rs,
                      int32_t workers);

/* ---- learned flip ranker (BF16 MLP, hand-written HIP forward / backward / Adam) ----
   (this block, pngf_mlp_* through pngf_dbs_run_ranked, has no counterpart in the reference)
   A din -> dh -> dh -> dout network with GELU, FP32 master weights, BF16 MFMA products
   with FP32 accumulation.  No counterpart in the reference: it ranks candidate flips so
   that exact scoring runs on the top-K only. */
typedef struct pngf_mlp pngf_mlp;
int pngf_mlp_create(int32_t din, int32_t dh, int32_t dout, int32_t seed, pngf_mlp** out);
int pngf_mlp_destroy(pngf_mlp* m);
const char* pngf_mlp_last_error(const pngf_mlp* m);
int pngf_mlp_num_params(const pngf_mlp* m);
/* parameters: W1[din][dh] b1[dh] W2[dh][dh] b2[dh] W3[dh][dout] b3[dout], FP32 */
int pngf_mlp_get_params(pngf_mlp* m, float* out);
int pngf_mlp_set_params(pngf_mlp* m, const float* in);
int pngf_mlp_predict(pngf_mlp* m, const float* X, int32_t n, float* out);
/* The forward pass of pngf_mlp_predict for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the
   reference.  X_dev [n][din] and out_dev [n][dout] are device pointers on the network's device;
   the work is enqueued on `stream` (a hipStream_t; null: the network's own stream) and the call
   returns without waiting.  It allocates nothing: the activations live in a fixed per-device
   scratch array of 2^20 floats, so 2 n dh must not exceed 1048576 (PNGF_ERR_ARG otherwise).
   Calls that overlap on one GPU share that array and race. */
int pngf_mlp_predict_dev(pngf_mlp* m, const float* X_dev, int32_t n, float* out_dev, void* stream);
/* gradient of mean((f(X) - Y)^2) for one batch, no update */
int pngf_mlp_grad(pngf_mlp* m, const float* X, const float* Y, int32_t n, float* grads, float* loss);
int pngf_mlp_train(pngf_mlp* m, const float* X, const float* Y, int32_t n, int32_t epochs, int32_t batch,
                   float lr, int32_t seed, float* loss);

/* Ranker inputs for the current state (LOWRANK, raster driver): din = 51 + 8 num_freq,
   dout = num_freq.  tiles: candidate tiles (single-tile flips). */
int pngf_ranker_dims(pngf_handle* h, int32_t* din, int32_t* dout);
int pngf_features(pngf_handle* h, const int32_t* tiles, int32_t ncand, float* out);

typedef struct {
    int32_t top_k;          /* candidates scored exactly per step */
    int32_t window;         /* steps over which the hit rate is measured */
    float min_hit_rate;     /* below it, steps fall back to exact full sweeps */
    int32_t reserved;
    const float* x_mean;    /* [din] input standardisation */
    const float* x_std;
    const float* y_mean;    /* [dout] output de-standardisation */
    const float* y_std;
} pngf_rank_params;
/* Ranked DBS (S11 objective, raster candidates): each step the network scores every
   candidate and the candidates are scored exactly in FP64 in that order, top_k at a time;
   the best improving flip of the first chunk that holds one is accepted.  A step whose
   chunks run through every candidate without an improving one ends the run (converged: an
   exact local minimum).  When fewer than min_hit_rate of the last `window` steps found an
   improving flip in their first chunk, the next step is an exact full sweep instead
   (accept its best).  flip_count counts exact evaluations only. */
int pngf_dbs_run_ranked(pngf_handle* h, pngf_mlp* m, const pngf_rank_params* rp, const pngf_dbs_params* p,
                        pngf_dbs_result* r);

/* Per-kernel timing collected with HIP events on the handle's stream (enable first).  No
   counterpart in the reference (pngf_profile, pngf_kernel_stats, pngf_autotune_result). */
#define PNGF_K_EVAL       0 /* candidate scoring (both methods) */
#define PNGF_K_COLBLK     1 /* LOWRANK accept: W[:,t] from W0 and the factors */
#define PNGF_K_ROWBLK     2 /* LOWRANK accept: W[t,:] from W0 and the factors */
#define PNGF_K_ACC_SMALL  3 /* accept: O(kN) updates of q / S, R, blocks */
#define PNGF_K_VCINV      4 /* PLAIN accept: V C^-1 */
#define PNGF_K_RANKK      5 /* PLAIN accept: rank-k update of C^-1;  LOWRANK: fold of factors */
#define PNGF_K_COUNT      6
typedef struct {
    int64_t launches;
    double total_ms;       /* summed event-timed duration */
    double bytes;          /* summed algorithmic bytes */
    double flops;          /* summed algorithmic flops */
} pngf_kstat;
int pngf_profile(pngf_handle* h, int32_t enable);
/* Tile shape chosen for the LOWRANK skinny products by the autotuner that pngf_init_state
   runs on the handle's own buffers; ms[i] = measured time of variant i (may be null). */
int pngf_autotune_result(pngf_handle* h, int32_t* chosen, int32_t* nvariants, double* ms);
int pngf_kernel_stats(pngf_handle* h, int32_t which, pngf_kstat* out);
/* After pngf_init_state: *used = 1 if the LOWRANK accept uses W = W^T (PNGF_SYM_AUTO and G
   symmetric, or PNGF_SYM_PROJECT and every G projected), *rel_asym = max over frequencies of
   max|G_ij - G_ji| / max|G_ij| (|z| = |Re z| + |Im z|) of G as given, before any projection.
   Either pointer may be null. */
int pngf_symmetry(pngf_handle* h, int32_t* used, double* rel_asym);
/* After pngf_init_state (LOWRANK): *columns = factor columns the handle holds before it folds them
   into W0, *rebuilt = 1 if M rows are rebuilt from A and Lam' (PNGF_FACTORS_REBUILT in effect).
   Either pointer may be null.  No counterpart in the reference, which folds every accept. */
int pngf_factor_capacity(pngf_handle* h, int32_t* columns, int32_t* rebuilt);

#ifdef __cplusplus
}
#endif
#endif /* PNGF_GPU_H */
