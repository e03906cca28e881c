// C ABI of the gfx950 DBS Woodbury path (include/pngf_gpu.h): handle, G input,
// one-time initial state (rocSOLVER / rocBLAS), candidate scoring, accepts and the
// DBS driver loop.  All device work of a handle is ordered on its own stream.
#include "kernels.h"
#include "layout.h"
#include "../../include/pngf_gpu.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <random>
#include <sstream>
#include <string>
#include <vector>
#include <sys/stat.h>
#include <thread>

using namespace pngf;

namespace {

typedef rocblas_double_complex rbz;

// iterations of the device-driven DBS loop per host synchronisation (and per captured graph)
constexpr int kLoopChunkMax = 16;
// rows of M rebuilt per zgemm when the factors are kept as A and Lam and are folded
constexpr int kMChunk = 768;

struct Timed {
    hipEvent_t a, b;
    int which;
    double bytes, flops;
};

// Draw-without-replacement visit order of the reference driver
// (pngf-opt/magichat.hpp:53-107, seed != 0 branch): std::mt19937 seeded with the
// int seed, uniform_int_distribution<> re-parameterised on every draw.
struct Hat {
    std::mt19937 gen;
    std::uniform_int_distribution<> dis;
    int total = 0, left = 0;
    std::vector<int> idx;
    void init(int seed, int n) {
        gen.seed(seed);
        dis = std::uniform_int_distribution<>(0, n - 1);
        total = left = n;
        idx.resize(n);
        for (int i = 0; i < n; i++) idx[i] = i;
    }
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

size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

// pinned host staging (int32 slots): accept change set, selection, batch ids, LU infos
constexpr int kPinSel = 512, kPinBatch = 1024, kPinInfo = 16384, kPinInts = 65536;

}  // namespace

struct pngf_handle {
    pngf_config cfg{};
    // complex64 state (PNGF_PREC_C64): every zd*-typed buffer of the LOWRANK state below then holds
    // complex64 entries, esz bytes each; Eobj, eobj_out and the objective stay complex128
    bool c64 = false;
    size_t esz = sizeof(zd);
    Layout L;
    Groups grp;
    int ngroups_drv = 0, adhoc = -1;
    std::map<std::vector<int>, int> group_of;  // sorted tile set -> driver group
    std::string err;
    hipStream_t st = nullptr;
    rocblas_handle rb = nullptr;
    Dims d{};
    ObjParams op{};
    std::vector<double> freqs;
    std::vector<int32_t> tile;
    std::vector<int8_t> P;
    std::vector<int32_t> owners;
    bool tile_set = false, inited = false;
    std::vector<int> g_loaded;
    std::vector<FreqConst> fc;
    std::vector<std::vector<zd>> gobj_extra;  // [f][(ne-1)*N]
    std::vector<double> pos;                  // [N][3] metres, for the synthetic G
    zd synth_scale{};                         // complex normalisation of the synthetic G
    double synth_beta = 0.0;                  // NF2FF row scale of the synthetic G
    // device: common
    int32_t *d_tile = nullptr, *d_owners = nullptr, *d_gtiles = nullptr, *d_gnt = nullptr, *d_gcomps = nullptr,
            *d_gnc = nullptr, *d_batch = nullptr, *d_small = nullptr, *d_sel = nullptr, *d_info = nullptr;
    int8_t* d_P = nullptr;
    double* d_pos = nullptr;
    double *d_F = nullptr, *d_Fbest = nullptr, *d_Fsel = nullptr;
    zd* d_eobj_out = nullptr;
    zd* d_pe = nullptr;
    // per frequency big buffers (host arrays of device pointers)
    std::vector<zd*> Gcm, Grow, Cinv, W0, Aall, Mall;
    zd** d_tabs = nullptr;  // device copy: [6][nf]
    // factors kept as A and Lam (no M): symmetric W and groups of <= kLdsGmax components
    bool lam = false;
    std::vector<zd*> Lam;
    int32_t* d_fstart = nullptr;
    int32_t* d_flen = nullptr;
    zd* d_mchunk = nullptr;  // rows of M rebuilt for a fold (kMChunk x ldN)
    DevTabs tabs{};
    DevVecs vecs{};
    // scratch
    zd *d_partial = nullptr, *d_acc = nullptr, *d_Mc = nullptr, *d_Ar = nullptr, *d_Bt = nullptr, *d_V = nullptr,
       *d_VC = nullptr, *d_CU = nullptr, *d_M = nullptr, *d_rut = nullptr;
    zd* d_lu = nullptr;  // per-block capacitance scratch when gmax > kLdsGmax
    float* d_feat = nullptr;  // ranker inputs [ngroups][din]
    float* d_lab = nullptr;   // ranker labels [ngroups][nf]
    int acc_stride = 0;
    size_t small_stride = 0, bt_stride = 0, v_stride = 0;
    int nsplit_items = 2048;
    size_t partial_elems = 0;
    int K = 0, Kmax = 0, flushes = 0;
    int skv = 0;                     // tile-shape variant of the skinny products (autotuned)
    hipStream_t st2 = nullptr;       // side stream: the capacitance solve of an accept
    bool sym_proj_ok = true;         // PNGF_SYM_PROJECT: every frequency's G was projected
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool sym = false;                // LOWRANK accept takes W[t,:] = W[:,t]^T (G symmetric)
    double asym = 0.0;               // max over frequencies of max|G_ij - G_ji| / max|G_ij|
    unsigned long long* d_asym = nullptr;
    int32_t* d_flag = nullptr;       // non-finite entries found in G (init)
    std::vector<double> tune_ms;     // measured time per variant (colblk [+ rowblk]), ms
    // pinned host staging
    int32_t* h_pin_i = nullptr;   // batch ids / small ints
    double* h_pin_d = nullptr;
    // device-driven raster loop (PNGF_LOOP_AUTO)
    LoopState* d_loop = nullptr;
    LoopState* h_loop = nullptr;  // pinned
    int32_t *d_cl = nullptr, *d_accg = nullptr;
    int64_t* d_acce = nullptr;
    double* d_accF = nullptr;
    int acc_alloc = 0;
    // kLoopChunk iterations of that loop captured once and replayed (all their arguments are
    // fixed device pointers and sizes: the loop state lives in device memory)
    // partition layout of the accept epilogue: the candidate groups' component lists are
    // disjoint; d_left = the components outside every group
    bool part = false;
    int32_t* d_left = nullptr;
    // device-driven loop with profiling on: events around each iteration's column strip and the
    // iteration's (pending, k, K) record, read after every chunk
    hipEvent_t lev[2 * kLoopChunkMax] = {};
    hipEvent_t levr[2 * kLoopChunkMax] = {};  // the same around the row strip (non-symmetric W)
    int32_t* d_iterlog = nullptr;
    int32_t* h_iterlog = nullptr;
    int nleft = 0;
    hipGraph_t loop_graph = nullptr;
    hipGraphExec_t loop_exec = nullptr;
    bool graph_failed = false;
    // profiling
    bool prof = false;
    std::vector<Timed> pending;
    std::vector<hipEvent_t> pool;
    pngf_kstat kst[PNGF_K_COUNT]{};
    double acc_seconds = 0.0;
    // checkpoint (device copies of the mutable state)
    bool ckpt = false;
    std::vector<zd*> ck_big;
    zd *ck_vec = nullptr, *ck_D = nullptr;
    int8_t* ck_P = nullptr;
    std::vector<int32_t> ck_tile;
    std::vector<int8_t> ck_Ph;
    // memory accounting
    size_t mem_used = 0;
    std::vector<std::pair<void*, size_t>> allocs;
};

namespace {

int fail(pngf_handle* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    return code;
}

#define HC(h, x)                                                                                  \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) return fail(h, PNGF_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define RB(h, x)                                                                                  \
    do {                                                                                          \
        rocblas_status s_ = (x);                                                                  \
        if (s_ != rocblas_status_success)                                                         \
            return fail(h, PNGF_ERR_HIP, std::string(#x) + ": rocblas status " + std::to_string((int)s_)); \
    } while (0)
#define CK(x)                      \
    do {                           \
        int r_ = (x);              \
        if (r_ != PNGF_OK) return r_; \
    } while (0)

int dalloc(pngf_handle* h, void** p, size_t bytes) {
    bytes = std::max<size_t>(bytes, 256);
    if (h->cfg.mem_limit > 0 && h->mem_used + bytes > (size_t)h->cfg.mem_limit)
        return fail(h, PNGF_ERR_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes exceeds mem_limit");
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(h, PNGF_ERR_NOMEM, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    h->mem_used += bytes;
    h->allocs.push_back({*p, bytes});
    return PNGF_OK;
}

template <class T>
int dalloc_t(pngf_handle* h, T** p, size_t n) { return dalloc(h, (void**)p, n * sizeof(T)); }
// n complex entries of the handle's state precision (esz bytes each)
int zalloc(pngf_handle* h, zd** p, size_t n) { return dalloc(h, (void**)p, n * h->esz); }

void dfree(pngf_handle* h, void* p) {
    if (!p) return;
    for (size_t i = 0; i < h->allocs.size(); i++)
        if (h->allocs[i].first == p) {
            h->mem_used -= h->allocs[i].second;
            h->allocs.erase(h->allocs.begin() + i);
            break;
        }
    (void)hipFree(p);
}

// ------------------------------------------------------------- profiling ---
void prof_begin(pngf_handle* h, Timed& t, hipStream_t s = nullptr) {
    if (!h->prof) return;
    if (h->pool.size() < 2) {
        for (int i = 0; i < 16; i++) { hipEvent_t e; if (hipEventCreate(&e) == hipSuccess) h->pool.push_back(e); }
    }
    t.a = h->pool.back(); h->pool.pop_back();
    t.b = h->pool.back(); h->pool.pop_back();
    (void)hipEventRecord(t.a, s ? s : h->st);
}
void prof_end(pngf_handle* h, Timed& t, int which, double bytes, double flops, hipStream_t s = nullptr) {
    if (!h->prof) return;
    (void)hipEventRecord(t.b, s ? s : h->st);
    t.which = which; t.bytes = bytes; t.flops = flops;
    h->pending.push_back(t);
}
// Resolve recorded event pairs (call after the stream has been synchronised).
void prof_collect(pngf_handle* h) {
    for (auto& t : h->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            auto& k = h->kst[t.which];
            k.launches++; k.total_ms += ms; k.bytes += t.bytes; k.flops += t.flops;
        }
        h->pool.push_back(t.a); h->pool.push_back(t.b);
    }
    h->pending.clear();
}

// ----------------------------------------------------------------- groups ---
void build_groups(pngf_handle* h) {
    Layout& L = h->L;
    Groups& G = h->grp;
    G = Groups();
    G.gmax = L.per_tile * std::max(1, h->cfg.max_flips);
    h->group_of.clear();
    if (h->cfg.driver == PNGF_DRIVER_RASTER) {
        std::vector<int> c;
        for (int t = 0; t < L.Tx * L.Ty; t++) {
            c.clear(); L.tile_components(t, c);
            bool touches = false;
            for (int i : c) for (int p : L.port) if (i == p) touches = true;
            int tt[1] = {t};
            G.add(L, tt, touches ? 0 : 1, t);
            if (!touches) h->group_of[{t}] = G.n - 1;
        }
    } else {
        int nhat = ((L.Tx - 1) / 2 + 1) * ((L.Ty - 1) / 2 + 1) - 1;
        for (int r = 0; r < nhat; r++) {
            int tt[4];
            int nt = ref_flip_set(r, L.Tx, L.Ty, tt);
            G.add(L, tt, nt, nt ? tt[0] : -1);
            if (nt) {
                std::vector<int> key(tt, tt + nt);
                std::sort(key.begin(), key.end());
                h->group_of[key] = G.n - 1;
            }
        }
    }
    h->ngroups_drv = G.n;
    int none[1] = {0};
    G.add(L, none, 0, -1);  // ad-hoc slot
    h->adhoc = G.n - 1;
}

GroupTab gtab(pngf_handle* h) {
    GroupTab g;
    g.tiles = h->d_gtiles; g.ntiles = h->d_gnt; g.comps = h->d_gcomps; g.ncomp = h->d_gnc; g.gmax = h->grp.gmax;
    return g;
}

// The LOWRANK state of a handle seen with the complex type of its precision (Z = zf for
// PNGF_PREC_C64, whose buffers hold complex64 entries behind the zd-typed pointers).
template <class Z>
struct LRView {
    DevTabsT<Z> tb;
    DevVecsT<Z> v;
    Z *acc, *lu, *Mc, *Ar, *Bt, *mchunk;
};
template <class Z>
LRView<Z> lrview(pngf_handle* h) {
    return LRView<Z>{tabs_as<Z>(h->tabs), vecs_as<Z>(h->vecs), (Z*)h->d_acc, (Z*)h->d_lu,
                     (Z*)h->d_Mc,         (Z*)h->d_Ar,         (Z*)h->d_Bt,  (Z*)h->d_mchunk};
}

// Host change set of group g against the host mirrors (same rule as the device).
int host_change_set(pngf_handle* h, int g, int* t, int* s, int* lp) {
    const Groups& G = h->grp;
    const int nt = G.ntiles[g];
    const int32_t* gtl = &G.tiles[4 * g];
    int k = 0;
    for (int a = 0; a < G.ncomp[g]; a++) {
        const int i = G.comps[(size_t)g * G.gmax + a];
        int nv = 0;
        for (int hh = 0; hh < 2; hh++) {
            int o = h->owners[2 * i + hh];
            if (o < 0) continue;
            int v = h->tile[o];
            for (int q = 0; q < nt; q++) v ^= (gtl[q] == o) ? 1 : 0;
            nv |= v;
        }
        if (nv != h->P[i]) { t[k] = i; s[k] = nv - h->P[i]; lp[k] = a; k++; }
    }
    return k;
}

// Writes group `g` (the ad-hoc slot) for an arbitrary tile set and uploads it.
int set_adhoc(pngf_handle* h, const int32_t* tiles, int nt) {
    Groups& G = h->grp;
    const int g = h->adhoc;
    std::vector<int> c;
    for (int i = 0; i < nt; i++) h->L.tile_components(tiles[i], c);
    std::sort(c.begin(), c.end());
    c.erase(std::unique(c.begin(), c.end()), c.end());
    if ((int)c.size() > G.gmax) return fail(h, PNGF_ERR_ARG, "flip touches more components than max_flips allows");
    for (int i = 0; i < 4; i++) G.tiles[4 * g + i] = i < nt ? tiles[i] : -1;
    G.ntiles[g] = nt;
    G.ncomp[g] = (int)c.size();
    G.prim[g] = tiles[0];
    for (int i = 0; i < G.gmax; i++) G.comps[(size_t)g * G.gmax + i] = i < (int)c.size() ? c[i] : -1;
    HC(h, hipMemcpyAsync(h->d_gtiles + 4 * g, &G.tiles[4 * g], 4 * sizeof(int32_t), hipMemcpyHostToDevice, h->st));
    HC(h, hipMemcpyAsync(h->d_gnt + g, &G.ntiles[g], sizeof(int32_t), hipMemcpyHostToDevice, h->st));
    HC(h, hipMemcpyAsync(h->d_gnc + g, &G.ncomp[g], sizeof(int32_t), hipMemcpyHostToDevice, h->st));
    HC(h, hipMemcpyAsync(h->d_gcomps + (size_t)g * G.gmax, &G.comps[(size_t)g * G.gmax], G.gmax * sizeof(int32_t),
                         hipMemcpyHostToDevice, h->st));
    HC(h, hipStreamSynchronize(h->st));
    if (h->cfg.method == PNGF_METHOD_LOWRANK && h->inited) {
        if (h->c64)
            launch_lr_block_current(h->st, h->d, tabs_as<zf>(h->tabs), gtab(h), g, h->K, (zf*)h->vecs.Dblk);
        else
            launch_lr_block_current(h->st, h->d, h->tabs, gtab(h), g, h->K, h->vecs.Dblk);
    }
    return PNGF_OK;
}

// Maps a tile set to a group (driver table, or the ad-hoc slot).
int group_for(pngf_handle* h, const int32_t* tiles, int nt, int* g) {
    if (!tiles || nt < 1 || nt > h->cfg.max_flips) return fail(h, PNGF_ERR_ARG, "bad flip tile list");
    std::vector<int> key(tiles, tiles + nt);
    for (int t : key)
        if (t < 0 || t >= h->L.Tx * h->L.Ty) return fail(h, PNGF_ERR_ARG, "tile index out of range");
    std::sort(key.begin(), key.end());
    for (size_t i = 1; i < key.size(); i++)
        if (key[i] == key[i - 1]) return fail(h, PNGF_ERR_ARG, "repeated tile in flip");
    auto it = h->group_of.find(key);
    if (it != h->group_of.end()) { *g = it->second; return PNGF_OK; }
    CK(set_adhoc(h, tiles, nt));
    *g = h->adhoc;
    return PNGF_OK;
}

// ------------------------------------------------------------ synthetic G ---
const double kFref = 30.0e9;
const double kEpsR = 3.5;
const double kSoftDefault = 1.0;  // softening radius in cells
// largest relative asymmetry of G for which the LOWRANK accept treats W as symmetric
const double kSymTol = 1e-14;
// PNGF_SYM_PROJECT: the largest asymmetry (relative to G's largest entry) that is projected away
const double kSymProjectTol = 1e-5;

void synth_positions(const Layout& L, double dx, double dy, std::vector<double>& pos) {
    pos.resize((size_t)L.N * 3);
    for (int i = 0; i < L.N; i++) {
        double x, y; int o;
        L.position(i, x, y, o);
        pos[3 * i] = x * dx; pos[3 * i + 1] = y * dy; pos[3 * i + 2] = o;
    }
}

double kr_of(double hz) { return 2.0 * kPi * hz * std::sqrt(kEpsR) / kC0; }

// Normalisation: at 30 GHz the port block of G sums to Vt = -(50 + 30i) / zfac,
// i.e. the port alone sees Zin = 50 + 30i ohm; every frequency f carries an
// extra factor f / 30 GHz.  The NF2FF rows (num_eobj = 3) are beta (f/30 GHz)
// times the x and y orientation of each component, with beta chosen so that the
// port alone has directivity 1.5 at 30 GHz.
void synth_norm(const Layout& L, const std::vector<double>& pos, double dx, double dy, double dz, double soft,
                zd& scale, double& beta) {
    const double zfac = dx * 0.5 / (2.0 * dy * dz);
    const double kr = kr_of(kFref), soft2 = (soft * dx) * (soft * dx);
    zd sum = zmk(0.0, 0.0);
    for (int p : L.port)
        for (int q : L.port) sum = zadd(sum, synth_entry(&pos[3 * p], &pos[3 * q], kr, soft2));
    const zd vt = zmk(-50.0 / zfac, -30.0 / zfac);
    scale = zdiv(vt, sum);
    const double pfac = dx * 0.5 * (2.0 * dy * dz) * 0.5;
    const double k0 = 2.0 * kPi * kFref / kC0;
    const double np = (double)L.port.size();
    const double c = 4.0 * kPi * (k0 * k0 / (32 * kPi * kPi * kEta0));
    beta = std::sqrt(1.5 * (-vt.x * pfac) / (c * np * np));
}

// ------------------------------------------------------------------ init ---
int check_ready(pngf_handle* h) {
    if (!h->inited) return fail(h, PNGF_ERR_STATE, "pngf_init_state has not completed");
    return PNGF_OK;
}

int upload_tabs(pngf_handle* h) {
    const int nf = h->d.nf;
    std::vector<zd*> all(6 * nf, nullptr);
    for (int f = 0; f < nf; f++) {
        all[0 * nf + f] = h->W0[f]; all[1 * nf + f] = h->Aall[f]; all[2 * nf + f] = h->Mall[f];
        all[3 * nf + f] = h->Grow[f]; all[4 * nf + f] = h->Cinv[f];
        all[5 * nf + f] = h->Lam.empty() ? nullptr : h->Lam[f];
    }
    HC(h, hipMemcpy(h->d_tabs, all.data(), all.size() * sizeof(zd*), hipMemcpyHostToDevice));
    h->tabs.W0 = h->d_tabs; h->tabs.Aall = h->d_tabs + nf; h->tabs.Mall = h->d_tabs + 2 * nf;
    h->tabs.Grow = h->d_tabs + 3 * nf; h->tabs.Cinv = h->d_tabs + 4 * nf;
    h->tabs.Lam = h->d_tabs + 5 * nf;
    h->tabs.fstart = h->d_fstart;
    h->tabs.flen = h->d_flen;
    return PNGF_OK;
}

int alloc_gcm(pngf_handle* h, int f) {
    if (!h->Gcm[f]) CK(dalloc_t(h, &h->Gcm[f], (size_t)h->d.ldN * h->d.N));
    return PNGF_OK;
}

// A NaN or an infinity in G gives NaN scores, which never compare below F_best: the run would
// stop as "converged" on garbage.  G is refused instead (the reference does not check).
int check_finite(pngf_handle* h, int f) {
    hipStream_t st = h->st;
    HC(h, hipMemsetAsync(h->d_flag, 0, sizeof(int32_t), st));
    launch_nonfinite(st, h->d.N, h->d.ldN, h->Gcm[f], h->d_flag);
    HC(h, hipMemcpyAsync(h->h_pin_i, h->d_flag, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HC(h, hipStreamSynchronize(st));
    int bad = h->h_pin_i[0];
    for (const zd& v : h->gobj_extra[f])
        if (!std::isfinite(v.x) || !std::isfinite(v.y)) bad = 1;
    if (bad) return fail(h, PNGF_ERR_NONFINITE, "G of frequency slot " + std::to_string(f) + " holds NaN or Inf");
    return PNGF_OK;
}

// *asym_f = max|G_ij - G_ji| / max|G_ij| of slot f's (complex128) G; 1 unless the symmetric accept may apply
int measure_asym(pngf_handle* h, int f, double* asym_f) {
    *asym_f = 1.0;
    if (h->cfg.symmetry == PNGF_SYM_OFF) return PNGF_OK;
    hipStream_t st = h->st;
    unsigned long long m[2] = {0, 0};
    HC(h, hipMemsetAsync(h->d_asym, 0, 2 * sizeof(unsigned long long), st));
    launch_asym(st, h->d.N, h->d.ldN, h->Gcm[f], h->d_asym);
    HC(h, hipMemcpyAsync(m, h->d_asym, sizeof(m), hipMemcpyDeviceToHost, st));
    HC(h, hipStreamSynchronize(st));
    double dm, am;
    std::memcpy(&dm, &m[0], sizeof(double));
    std::memcpy(&am, &m[1], sizeof(double));
    *asym_f = am > 0.0 ? dm / am : 0.0;
    h->asym = std::max(h->asym, *asym_f);
    return PNGF_OK;
}

// PNGF_SYM_PROJECT: G measured as given (the reported asymmetry) and, within kSymProjectTol,
// replaced in place by (G + G^T) / 2 before anything is built from it; *asym_f is then the
// projected G's asymmetry (0).  A frequency left unprojected keeps both strips for the handle.
int project_g(pngf_handle* h, int f, double* asym_f) {
    CK(measure_asym(h, f, asym_f));
    if (*asym_f <= kSymProjectTol) {
        launch_symmetrize(h->st, h->d.N, h->d.ldN, h->Gcm[f]);
        HC(h, hipGetLastError());
        *asym_f = 0.0;
    } else {
        h->sym_proj_ok = false;
    }
    return PNGF_OK;
}

// complex64 state (the reference's USE_SINGLE_PRECISION): G cast to c64 as the reference casts the
// file (pngfopt.hpp:128-156), then Gobj row 0 and Einc (pngfopt.hpp:584-595), C and its inverse by
// cgetrf + cgetri (linalg.hpp:49-84), PEinc, R = Gobj C^-1, Eobj = R PEinc, W = (G - I) C^-1, q and the
// blocks, all in c64 (Eobj widened to c128)
int init_freq_c64(pngf_handle* h, int f) {
    typedef rocblas_float_complex rbc;
    const int N = h->d.N, ld = h->d.ldN, ne = h->d.ne;
    hipStream_t st = h->st;
    const rbc one = {1.f, 0.f}, zero = {0.f, 0.f};
    CK(check_finite(h, f));
    double asym_f = 1.0;
    if (h->cfg.symmetry == PNGF_SYM_PROJECT) CK(project_g(h, f, &asym_f));
    else CK(measure_asym(h, f, &asym_f));
    zd* Gbuf = nullptr;
    CK(zalloc(h, &Gbuf, (size_t)ld * N));
    zf* Gf = (zf*)Gbuf;
    launch_cast_c64(st, N, ld, h->Gcm[f], Gf);
    HC(h, hipStreamSynchronize(st));
    dfree(h, h->Gcm[f]);
    h->Gcm[f] = nullptr;
    zd* Cbuf = nullptr;
    CK(zalloc(h, &Cbuf, (size_t)ld * N));
    zf* C = (zf*)Cbuf;
    launch_build_C(st, N, ld, h->d_P, (const zf*)Gf, C);
    zf* gobj = (zf*)h->vecs.Gobj + (size_t)f * ne * ld;
    zf* einc = (zf*)h->vecs.Einc + (size_t)f * ld;
    launch_port_sums(st, N, ld, (const zf*)Gf, h->d_small + 400, (int)h->L.port.size(), gobj, einc);
    std::vector<zf> row(N);
    for (int r = 1; r < ne; r++) {
        for (int n = 0; n < N; n++) {
            const zd v = h->gobj_extra[f][(size_t)(r - 1) * N + n];
            row[n] = zmkf((float)v.x, (float)v.y);
        }
        HC(h, hipMemcpy(gobj + (size_t)r * ld, row.data(), N * sizeof(zf), hipMemcpyHostToDevice));
    }
    int* ipiv = nullptr;
    CK(dalloc_t(h, &ipiv, N));
    RB(h, rocsolver_cgetrf(h->rb, N, N, (rbc*)C, ld, ipiv, h->d_info));
    int info = 0;
    HC(h, hipMemcpyAsync(&info, h->d_info, sizeof(int), hipMemcpyDeviceToHost, st));
    HC(h, hipStreamSynchronize(st));
    if (info != 0) { dfree(h, ipiv); dfree(h, Cbuf); dfree(h, Gbuf); return fail(h, PNGF_ERR_SINGULAR, "C is singular (cgetrf info " + std::to_string(info) + ")"); }
    RB(h, rocsolver_cgetri(h->rb, N, (rbc*)C, ld, ipiv, h->d_info));
    HC(h, hipMemcpyAsync(&info, h->d_info, sizeof(int), hipMemcpyDeviceToHost, st));
    HC(h, hipStreamSynchronize(st));
    dfree(h, ipiv);
    if (info != 0) { dfree(h, Cbuf); dfree(h, Gbuf); return fail(h, PNGF_ERR_SINGULAR, "C is singular (cgetri info " + std::to_string(info) + ")"); }
    zf* pe = (zf*)h->d_pe;
    launch_pe(st, N, h->d_P, (const zf*)einc, pe);
    zf* R = (zf*)h->vecs.R + (size_t)f * ne * ld;
    RB(h, rocblas_cgemm(h->rb, rocblas_operation_transpose, rocblas_operation_none, N, ne, N, &one, (rbc*)C, ld,
                        (rbc*)gobj, ld, &zero, (rbc*)R, ld));
    launch_dot_rows(st, N, ld, ne, (const zf*)R, (const zf*)pe, h->vecs.Eobj + (size_t)f * ne);
    launch_add_diag(st, N, ld, Gf, -1.0);
    zd* Wbuf = nullptr;
    CK(zalloc(h, &Wbuf, (size_t)ld * N));
    zf* W = (zf*)Wbuf;
    RB(h, rocblas_cgemm(h->rb, rocblas_operation_transpose, rocblas_operation_transpose, N, N, N, &one, (rbc*)C, ld,
                        (rbc*)Gf, ld, &zero, (rbc*)W, ld));
    if (asym_f <= kSymTol) launch_symmetrize(st, N, ld, W);
    RB(h, rocblas_cgemv(h->rb, rocblas_operation_transpose, N, N, &one, (rbc*)W, ld, (rbc*)pe, 1, &zero,
                        (rbc*)((zf*)h->vecs.q + (size_t)f * ld), 1));
    launch_extract_blocks(st, h->d, (const zf*)W, gtab(h), (zf*)h->vecs.Dblk + (size_t)f * h->d.ngroups * h->d.gmax * h->d.gmax);
    HC(h, hipStreamSynchronize(st));
    dfree(h, Gbuf);
    dfree(h, Cbuf);
    h->W0[f] = Wbuf;
    HC(h, hipGetLastError());
    // Vinc, LN1inc, LN2inc summed in single precision (pngfopt.hpp:599-617 in c64), kept as c128
    std::vector<zf> e(h->L.port.size());
    for (size_t i = 0; i < h->L.port.size(); i++)
        HC(h, hipMemcpyAsync(&e[i], einc + h->L.port[i], sizeof(zf), hipMemcpyDeviceToHost, st));
    HC(h, hipStreamSynchronize(st));
    zf vinc = zmkf(0.f, 0.f), ln1 = zmkf(0.f, 0.f), ln2 = zmkf(0.f, 0.f);
    for (size_t i = 0; i < e.size(); i++) vinc = zadd(vinc, e[i]);
    if (ne > 2)
        for (int p : h->L.port) {
            const zd a = h->gobj_extra[f][p], b = h->gobj_extra[f][(size_t)N + p];
            ln1 = zadd(ln1, zmkf((float)a.x, (float)a.y));
            ln2 = zadd(ln2, zmkf((float)b.x, (float)b.y));
        }
    FreqConst fcv;
    fcv.vinc = zwide(vinc);
    fcv.ln1 = zwide(ln1);
    fcv.ln2 = zwide(ln2);
    fcv.k0 = 2.0 * kPi * h->freqs[f] / kC0;
    h->fc[f] = fcv;
    return PNGF_OK;
}

int init_freq(pngf_handle* h, int f) {
    if (h->c64) return init_freq_c64(h, f);
    const int N = h->d.N, ld = h->d.ldN, ne = h->d.ne;
    hipStream_t st = h->st;
    const rbz one = {1.0, 0.0}, zero = {0.0, 0.0};
    zd* G = h->Gcm[f];
    CK(check_finite(h, f));
    double asym_f = 1.0;
    if (h->cfg.symmetry == PNGF_SYM_PROJECT) CK(project_g(h, f, &asym_f));
    zd* C = nullptr;
    CK(dalloc_t(h, &C, (size_t)ld * N));
    launch_build_C(st, N, ld, h->d_P, G, C);
    zd* gobj = h->vecs.Gobj + (size_t)f * ne * ld;
    zd* einc = h->vecs.Einc + (size_t)f * ld;
    launch_port_sums(st, N, ld, G, h->d_small + 400, (int)h->L.port.size(), gobj, einc);
    for (int r = 1; r < ne; r++)
        HC(h, hipMemcpyAsync(gobj + (size_t)r * ld, h->gobj_extra[f].data() + (size_t)(r - 1) * N, N * sizeof(zd),
                             hipMemcpyHostToDevice, st));
    // C^-1 in place (linalg.hpp:49-84: zgetrf + zgetri).  A design with no metal component has
    // C = I (build_C wrote the identity), which is its own inverse: nothing to factor.
    const bool all_void = std::find(h->P.begin(), h->P.end(), (int8_t)1) == h->P.end();
    if (!all_void) {
        int* ipiv = nullptr;
        CK(dalloc_t(h, &ipiv, N));
        RB(h, rocsolver_zgetrf(h->rb, N, N, (rbz*)C, ld, ipiv, h->d_info));
        int info = 0;
        HC(h, hipMemcpyAsync(&info, h->d_info, sizeof(int), hipMemcpyDeviceToHost, st));
        HC(h, hipStreamSynchronize(st));
        if (info != 0) { dfree(h, ipiv); dfree(h, C); return fail(h, PNGF_ERR_SINGULAR, "C is singular (zgetrf info " + std::to_string(info) + ")"); }
        RB(h, rocsolver_zgetri(h->rb, N, (rbz*)C, ld, ipiv, h->d_info));
        HC(h, hipMemcpyAsync(&info, h->d_info, sizeof(int), hipMemcpyDeviceToHost, st));
        HC(h, hipStreamSynchronize(st));
        dfree(h, ipiv);
        if (info != 0) { dfree(h, C); return fail(h, PNGF_ERR_SINGULAR, "C is singular (zgetri info " + std::to_string(info) + ")"); }
    }
    // PEinc, R = Gobj C^-1 (row-major ne x ld: R^T = C^-T Gobj^T), Eobj = R PEinc
    launch_pe(st, N, h->d_P, einc, h->d_pe);
    zd* R = h->vecs.R + (size_t)f * ne * ld;
    RB(h, rocblas_zgemm(h->rb, rocblas_operation_transpose, rocblas_operation_none, N, ne, N, &one, (rbz*)C, ld,
                        (rbz*)gobj, ld, &zero, (rbz*)R, ld));
    launch_dot_rows(st, N, ld, ne, R, h->d_pe, h->vecs.Eobj + (size_t)f * ne);
    if (h->cfg.method == PNGF_METHOD_PLAIN) {
        RB(h, rocblas_zgemv(h->rb, rocblas_operation_none, N, N, &one, (rbz*)C, ld, (rbz*)h->d_pe, 1, &zero,
                            (rbz*)(h->vecs.S + (size_t)f * ld), 1));
        zd* Gr = nullptr;
        CK(dalloc_t(h, &Gr, (size_t)ld * N));
        launch_transpose(st, N, ld, G, Gr);
        HC(h, hipStreamSynchronize(st));
        dfree(h, G);
        h->Gcm[f] = nullptr;
        h->Grow[f] = Gr;
        h->Cinv[f] = C;
    } else {
        if (h->cfg.symmetry != PNGF_SYM_PROJECT) CK(measure_asym(h, f, &asym_f));
        // W = (G - I) C^-1, stored row-major: W^T = C^-T (G - I)^T
        launch_add_diag(st, N, ld, G, -1.0);
        zd* W = nullptr;
        CK(dalloc_t(h, &W, (size_t)ld * N));
        RB(h, rocblas_zgemm(h->rb, rocblas_operation_transpose, rocblas_operation_transpose, N, N, N, &one, (rbz*)C,
                            ld, (rbz*)G, ld, &zero, (rbz*)W, ld));
        // symmetric G: W is symmetric, and the computed one is made exactly so (its asymmetry is
        // rounding, cond(C) eps) before the blocks and q are taken from it
        if (asym_f <= kSymTol) launch_symmetrize(st, N, ld, W);
        RB(h, rocblas_zgemv(h->rb, rocblas_operation_transpose, N, N, &one, (rbz*)W, ld, (rbz*)h->d_pe, 1, &zero,
                            (rbz*)(h->vecs.q + (size_t)f * ld), 1));
        launch_extract_blocks(st, h->d, W, gtab(h), h->vecs.Dblk + (size_t)f * h->d.ngroups * h->d.gmax * h->d.gmax);
        HC(h, hipStreamSynchronize(st));
        dfree(h, G);
        dfree(h, C);
        h->Gcm[f] = nullptr;
        h->W0[f] = W;
    }
    HC(h, hipGetLastError());
    // host constants: Vinc, LN1inc, LN2inc (pngfopt.hpp:599-617), k0
    std::vector<zd> e(h->L.port.size());
    for (size_t i = 0; i < h->L.port.size(); i++)
        HC(h, hipMemcpyAsync(&e[i], einc + h->L.port[i], sizeof(zd), hipMemcpyDeviceToHost, st));
    HC(h, hipStreamSynchronize(st));
    FreqConst fcv;
    fcv.vinc = zmk(0.0, 0.0); fcv.ln1 = zmk(0.0, 0.0); fcv.ln2 = zmk(0.0, 0.0);
    for (size_t i = 0; i < e.size(); i++) fcv.vinc = zadd(fcv.vinc, e[i]);
    if (ne > 2)
        for (int p : h->L.port) {
            fcv.ln1 = zadd(fcv.ln1, h->gobj_extra[f][p]);
            fcv.ln2 = zadd(fcv.ln2, h->gobj_extra[f][(size_t)N + p]);
        }
    fcv.k0 = 2.0 * kPi * h->freqs[f] / kC0;
    h->fc[f] = fcv;
    return PNGF_OK;
}

// ------------------------------------------------------------ autotuner ---
// Times every tile-shape variant of the skinny products (both, or the column strip alone
// when W is symmetric) on this handle's own
// buffers (N, frequencies, a representative factor depth K) and keeps the fastest.
// The factor buffers are zero at this point, so the timed products read finite data;
// the columns they write are rewritten by the first accepts that reach them.
int autotune_skinny(pngf_handle* h) {
    const Dims& d = h->d;
    int g0 = -1;
    for (int g = 0; g < h->ngroups_drv; g++)
        if (h->grp.ntiles[g] > 0) { g0 = g; break; }
    if (g0 < 0) return PNGF_OK;
    const int k = h->grp.ncomp[g0];
    const int K = std::max(0, std::min(h->Kmax - k, 64 * k));
    for (int i = 0; i < k; i++) h->h_pin_i[i] = h->grp.comps[(size_t)g0 * d.gmax + i];
    HC(h, hipMemcpy(h->d_small, h->h_pin_i, k * sizeof(int32_t), hipMemcpyHostToDevice));
    const int nv = skinny_variants(k, h->c64);
    const LRView<zf> w32 = lrview<zf>(h);
    hipEvent_t e0, e1;
    HC(h, hipEventCreate(&e0));
    HC(h, hipEventCreate(&e1));
    h->tune_ms.assign(nv, 0.0);
    double best = 1e30;
    for (int v = 0; v < nv; v++) {
        for (int rep = 0; rep < 3; rep++) {
            if (rep == 1) HC(h, hipEventRecord(e0, h->st));
            if (h->c64) {
                launch_lr_colblk(h->st, d, w32.tb, h->d_small, k, K, w32.Mc, (int)h->small_stride, v);
                if (!h->sym)
                    launch_lr_rowblk(h->st, d, w32.tb, h->d_small, k, K, w32.Ar, (int)h->small_stride, w32.Bt,
                                     h->bt_stride, v);
                continue;
            }
            launch_lr_colblk(h->st, d, h->tabs, h->d_small, k, K, h->d_Mc, (int)h->small_stride, v);
            if (!h->sym)
                launch_lr_rowblk(h->st, d, h->tabs, h->d_small, k, K, h->d_Ar, (int)h->small_stride, h->d_Bt,
                                 h->bt_stride, v);
        }
        HC(h, hipEventRecord(e1, h->st));
        HC(h, hipEventSynchronize(e1));
        float ms = 0.f;
        HC(h, hipEventElapsedTime(&ms, e0, e1));
        h->tune_ms[v] = ms / 2.0;
        if (ms < best) { best = ms; h->skv = v; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HC(h, hipGetLastError());
    return PNGF_OK;
}

// ------------------------------------------------------------------ eval ---
int eval_launch(pngf_handle* h, const int32_t* groups, int nb) {
    hipStream_t st = h->st;
    for (int i = 0; i < nb; i++) h->h_pin_i[kPinBatch + i] = groups[i];
    HC(h, hipMemcpyAsync(h->d_batch, h->h_pin_i + kPinBatch, nb * sizeof(int32_t), hipMemcpyHostToDevice, st));
    Timed tm{};
    prof_begin(h, tm);
    const Dims& d = h->d;
    double bytes = 0, flops = 0;
    const double kk = (double)d.gmax;
    if (h->cfg.method == PNGF_METHOD_LOWRANK) {
        if (h->c64)
            launch_lr_eval(st, d, vecs_as<zf>(h->vecs), gtab(h), h->d_tile, h->d_P, h->d_owners, h->d_batch, nb,
                           h->d_eobj_out, h->d_info, (zf*)h->d_lu);
        else
            launch_lr_eval(st, d, h->vecs, gtab(h), h->d_tile, h->d_P, h->d_owners, h->d_batch, nb, h->d_eobj_out,
                           h->d_info, h->d_lu);
        bytes = (double)nb * d.nf * (double)h->esz * (kk * kk + kk * (d.ne + 2));
        flops = (double)nb * d.nf * ((8.0 / 3.0) * kk * kk * kk + 16.0 * kk * kk);
    } else {
        const int items = nb * d.nf;
        int nsplit = std::max(1, std::min(64, h->nsplit_items / std::max(1, items)));
        nsplit = std::min(nsplit, std::max(1, d.N / 64));
        launch_plain_eval(st, d, h->tabs, h->vecs, gtab(h), h->d_tile, h->d_P, h->d_owners, h->d_batch, nb, nsplit,
                          h->d_partial, h->d_eobj_out, h->d_info, nullptr, 0, h->d_lu);
        bytes = (double)items * 16.0 * d.N * (2 * kk + 1 + d.ne);
        flops = (double)items * (8.0 * d.N * (kk * kk + 2 * kk + d.ne) + (8.0 / 3.0) * kk * kk * kk);
    }
    prof_end(h, tm, PNGF_K_EVAL, bytes, flops);
    launch_objective(st, h->op, h->vecs.fc, h->d_eobj_out, nb, h->d_F, h->d_Fbest, h->d_sel, h->d_Fsel);
    HC(h, hipGetLastError());
    return PNGF_OK;
}

// ---------------------------------------------------------------- accept ---
// out -= A_all[:, :K] M[:K, :] for frequency f (out: N x ldN row-major, W0 or a copy of it);
// with the factors kept as A and Lam, M is rebuilt kMChunk rows at a time.
int fold_into_c64(pngf_handle* h, int f, zf* out) {
    typedef rocblas_float_complex rbc;
    const rbc one = {1.f, 0.f}, mone = {-1.f, 0.f};
    const Dims& d = h->d;
    const LRView<zf> w = lrview<zf>(h);
    if (!h->lam) {
        RB(h, rocblas_cgemm(h->rb, rocblas_operation_none, rocblas_operation_none, d.N, d.N, h->K, &mone,
                            (rbc*)h->Mall[f], d.ldN, (rbc*)h->Aall[f], d.Kmax, &one, (rbc*)out, d.ldN));
        return PNGF_OK;
    }
    for (int kk0 = 0; kk0 < h->K; kk0 += kMChunk) {
        const int nr = std::min(kMChunk, h->K - kk0);
        launch_lr_build_m(h->st, d, w.tb, f, kk0, nr, w.mchunk, d.ldN);
        RB(h, rocblas_cgemm(h->rb, rocblas_operation_none, rocblas_operation_none, d.N, d.N, nr, &mone,
                            (rbc*)w.mchunk, d.ldN, (rbc*)((zf*)h->Aall[f] + kk0), d.Kmax, &one, (rbc*)out, d.ldN));
    }
    return PNGF_OK;
}

int fold_into(pngf_handle* h, int f, zd* out) {
    if (h->c64) return fold_into_c64(h, f, (zf*)out);
    const rbz one = {1.0, 0.0}, mone = {-1.0, 0.0};
    const Dims& d = h->d;
    if (!h->lam) {
        RB(h, rocblas_zgemm(h->rb, rocblas_operation_none, rocblas_operation_none, d.N, d.N, h->K, &mone,
                            (rbz*)h->Mall[f], d.ldN, (rbz*)h->Aall[f], d.Kmax, &one, (rbz*)out, d.ldN));
        return PNGF_OK;
    }
    for (int kk0 = 0; kk0 < h->K; kk0 += kMChunk) {
        const int nr = std::min(kMChunk, h->K - kk0);
        launch_lr_build_m(h->st, d, h->tabs, f, kk0, nr, h->d_mchunk, d.ldN);
        RB(h, rocblas_zgemm(h->rb, rocblas_operation_none, rocblas_operation_none, d.N, d.N, nr, &mone,
                            (rbz*)h->d_mchunk, d.ldN, (rbz*)(h->Aall[f] + kk0), d.Kmax, &one, (rbz*)out, d.ldN));
    }
    return PNGF_OK;
}

int flush_factors(pngf_handle* h) {
    if (h->K == 0) return PNGF_OK;
    const Dims& d = h->d;
    Timed tm{};
    prof_begin(h, tm);
    for (int f = 0; f < d.nf; f++) {
        CK(fold_into(h, f, h->W0[f]));
        if (h->sym) {
            if (h->c64) launch_symmetrize(h->st, d.N, d.ldN, (zf*)h->W0[f]);
            else launch_symmetrize(h->st, d.N, d.ldN, h->W0[f]);
        }
    }
    const double eb2 = 2.0 * (double)h->esz;
    prof_end(h, tm, PNGF_K_RANKK, (double)d.nf * (eb2 * d.N * (double)d.N + eb2 * d.N * (double)h->K),
             (double)d.nf * 8.0 * h->K * (double)d.N * d.N);
    h->K = 0;
    h->flushes++;
    return PNGF_OK;
}

// The LOWRANK accept of group g (change set t, s, lp already on the device) in the state's precision.
template <class Z>
int accept_lowrank(pngf_handle* h, int g, int k, int nt, const int32_t* dt, const int32_t* ds, const int32_t* dlp,
                   const int32_t* dgt) {
    hipStream_t st = h->st;
    const Dims& d = h->d;
    if (h->K + k > h->Kmax) CK(flush_factors(h));
    const int K = h->K;
    const LRView<Z> w = lrview<Z>(h);
    // the capacitance solve (one workgroup per frequency) runs on the side stream while the
    // strips are formed; the epilogue waits for both
    HC(h, hipEventRecord(h->ev_fork, st));
    HC(h, hipStreamWaitEvent(h->st2, h->ev_fork, 0));
    Timed tp{};
    prof_begin(h, tp, h->st2);
    launch_lr_accept_prep(h->st2, d, w.v, g, dt, ds, dlp, k, w.acc, h->acc_stride, h->d_info, w.lu);
    prof_end(h, tp, PNGF_K_ACC_SMALL, 0, 0, h->st2);
    HC(h, hipEventRecord(h->ev_join, h->st2));
    Timed tm{};
    prof_begin(h, tm);
    launch_lr_gather(st, d, w.tb, dt, k, K, w.Mc, h->sym ? (Z*)nullptr : w.Ar, (int)h->small_stride);
    prof_end(h, tm, PNGF_K_ACC_SMALL, 0, 0);
    const double N = d.N, kd = k, Kd = K, eb = (double)h->esz;
    Timed t1{};
    prof_begin(h, t1);
    launch_lr_colblk(st, d, w.tb, dt, k, K, w.Mc, (int)h->small_stride, h->skv);
    prof_end(h, t1, PNGF_K_COLBLK, d.nf * eb * (N * Kd + 2 * N * kd + Kd * kd), d.nf * 8.0 * N * Kd * kd);
    if (!h->sym) {
        Timed t2{};
        prof_begin(h, t2);
        launch_lr_rowblk(st, d, w.tb, dt, k, K, w.Ar, (int)h->small_stride, w.Bt, h->bt_stride, h->skv);
        prof_end(h, t2, PNGF_K_ROWBLK, d.nf * eb * (N * Kd + 2 * N * kd + Kd * kd), d.nf * 8.0 * N * Kd * kd);
    }
    HC(h, hipStreamWaitEvent(st, h->ev_join, 0));
    Timed t3{};
    prof_begin(h, t3);
    // M_new = diag(s) cap^-1 W[t,:]; W[t,:]^T is the row strip Bt, or, for symmetric W, the
    // column strip just written to A_all[:, K:K+k]
    if (d.gmax <= kLdsGmax) {
        launch_lr_finish(st, d, w.tb, w.v, gtab(h), (const Z*)w.acc, h->acc_stride, (const Z*)w.Bt, h->bt_stride, ds,
                         k, K, nullptr, h->part ? h->d_left : nullptr, h->nleft, h->ngroups_drv);
    } else {
        if (h->sym)
            launch_apply_capinv<Z>(st, d, w.acc, h->acc_stride, nullptr, 0, d.Kmax, k, ds, w.tb.Mall, nullptr, 0, K,
                                   d.ldN, w.tb.Aall, K);
        else
            launch_apply_capinv<Z>(st, d, w.acc, h->acc_stride, w.Bt, h->bt_stride, d.gmax, k, ds, w.tb.Mall,
                                   nullptr, 0, K, d.ldN);
        launch_lr_small(st, d, w.tb, w.v, (const Z*)w.acc, h->acc_stride, k, K);
        launch_lr_blocks(st, d, w.tb, w.v, gtab(h), k, K);
    }
    launch_mask_update(st, h->d_P, h->d_tile, dt, ds, k, dgt, nt);
    prof_end(h, t3, PNGF_K_ACC_SMALL, 0, 0);
    h->K += k;
    return PNGF_OK;
}

int accept_group(pngf_handle* h, int g) {
    hipStream_t st = h->st;
    const Dims& d = h->d;
    int* t = h->h_pin_i;
    int* s = t + 96;
    int* lp = s + 96;
    int k = host_change_set(h, g, t, s, lp);
    int* gt = lp + 96;
    const int nt = h->grp.ntiles[g];
    for (int i = 0; i < 4; i++) gt[i] = h->grp.tiles[4 * g + i];
    for (int i = 0; i < (int)h->L.port.size(); i++) gt[4 + i] = h->L.port[i];
    HC(h, hipMemcpyAsync(h->d_small, t, 300 * sizeof(int32_t), hipMemcpyHostToDevice, st));
    const int32_t* dt = h->d_small;
    const int32_t* ds = h->d_small + 96;
    const int32_t* dlp = h->d_small + 192;
    const int32_t* dgt = h->d_small + 288;
    auto t0 = std::chrono::high_resolution_clock::now();
    if (k > 0 && h->cfg.method == PNGF_METHOD_LOWRANK) {
        if (h->c64) CK(accept_lowrank<zf>(h, g, k, nt, dt, ds, dlp, dgt));
        else CK(accept_lowrank<zd>(h, g, k, nt, dt, ds, dlp, dgt));
    } else if (k > 0) {
        // PLAIN: re-score the flip keeping its factors, then pngfopt.hpp:256-367 (V C^-1 and the
        // rank-k update on the library's FP64-MFMA kernels)
        h->h_pin_i[kPinBatch] = g;
        HC(h, hipMemcpyAsync(h->d_batch, h->h_pin_i + kPinBatch, sizeof(int32_t), hipMemcpyHostToDevice, st));
        const int nsplit = std::max(1, std::min(64, h->nsplit_items / d.nf));
        Timed tm{};
        prof_begin(h, tm);
        launch_plain_eval(st, d, h->tabs, h->vecs, gtab(h), h->d_tile, h->d_P, h->d_owners, h->d_batch, 1, nsplit,
                          h->d_partial, h->d_eobj_out, h->d_info, h->d_acc, h->acc_stride, h->d_lu);
        launch_build_V(st, d, h->tabs, dt, k, h->d_V, h->v_stride);
        prof_end(h, tm, PNGF_K_ACC_SMALL, 0, 0);
        const double N = d.N, kd = k;
        Timed t1{};
        prof_begin(h, t1);
        launch_plain_vcinv(st, d, h->tabs, k, h->d_V, h->v_stride, h->d_VC, h->v_stride);
        prof_end(h, t1, PNGF_K_VCINV, d.nf * 16.0 * (N * N + 2 * N * kd), d.nf * 8.0 * kd * N * N);
        Timed t2{};
        prof_begin(h, t2);
        launch_apply_capinv<zd>(st, d, h->d_acc, h->acc_stride, h->d_VC, h->v_stride, k, k, nullptr, nullptr, h->d_M,
                            h->v_stride, 0, d.ldN);
        launch_snapshot_cinvu(st, d, h->tabs, dt, ds, k, h->d_CU, h->v_stride);
        prof_end(h, t2, PNGF_K_ACC_SMALL, 0, 0);
        Timed t3{};
        prof_begin(h, t3);
        launch_plain_rankk(st, d, h->tabs, h->d_CU, h->v_stride, h->d_M, h->v_stride, k);
        prof_end(h, t3, PNGF_K_RANKK, d.nf * (32.0 * N * N + 32.0 * N * kd), d.nf * 8.0 * kd * N * N);
        Timed t4{};
        prof_begin(h, t4);
        launch_plain_small(st, d, h->vecs, h->d_acc, h->acc_stride, h->d_CU, h->v_stride, h->d_M, h->v_stride, dt, ds,
                           k, h->d_rut);
        launch_mask_update(st, h->d_P, h->d_tile, dt, ds, k, dgt, nt);
        prof_end(h, t4, PNGF_K_ACC_SMALL, 0, 0);
    } else {
        launch_mask_update(st, h->d_P, h->d_tile, dt, ds, 0, dgt, nt);
    }
    HC(h, hipGetLastError());
    for (int c = 0; c < k; c++) h->P[t[c]] = (int8_t)(h->P[t[c]] + s[c]);
    for (int i = 0; i < nt; i++) h->tile[h->grp.tiles[4 * g + i]] ^= 1;
    if (h->prof) {
        HC(h, hipStreamSynchronize(st));
        h->acc_seconds += std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
    }
    return PNGF_OK;
}

int sync_sel(pngf_handle* h, int nb, int* sel, double* F_sel, bool check_info) {
    hipStream_t st = h->st;
    HC(h, hipMemcpyAsync(h->h_pin_i + kPinSel, h->d_sel, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HC(h, hipMemcpyAsync(h->h_pin_d, h->d_Fsel, sizeof(double), hipMemcpyDeviceToHost, st));
    if (check_info) HC(h, hipMemcpyAsync(h->h_pin_i + kPinInfo, h->d_info, sizeof(int32_t) * nb * h->d.nf, hipMemcpyDeviceToHost, st));
    HC(h, hipStreamSynchronize(st));
    prof_collect(h);
    *sel = h->h_pin_i[kPinSel];
    *F_sel = h->h_pin_d[0];
    if (check_info)
        for (int i = 0; i < nb * h->d.nf; i++)
            if (h->h_pin_i[kPinInfo + i] != 0) return fail(h, PNGF_ERR_SINGULAR, "capacitance matrix is singular");
    return PNGF_OK;
}

double host_objective(pngf_handle* h, const zd* eobj) { return objective_of(h->op, h->fc.data(), eobj); }

int ranker_din(const pngf_handle* h) { return 51 + 8 * h->d.nf; }

int ensure_feat(pngf_handle* h) {
    if (h->cfg.method != PNGF_METHOD_LOWRANK || h->c64)
        return fail(h, PNGF_ERR_STATE, "ranker inputs need the LOWRANK method in complex128");
    if (!h->d_feat) {
        CK(dalloc_t(h, &h->d_feat, (size_t)h->d.ngroups * ranker_din(h)));
        CK(dalloc_t(h, &h->d_lab, (size_t)h->d.ngroups * h->d.nf));
    }
    return PNGF_OK;
}

// Ranker inputs of the groups already uploaded to d_batch (nb of them) into host rows.
int features_batch(pngf_handle* h, int nb, float* host) {
    CK(ensure_feat(h));
    launch_features(h->st, h->d, h->vecs, gtab(h), h->d_tile, h->d_P, h->d_owners, h->d_batch, nb, h->L.Tx, h->L.Ty,
                    h->op, h->d_feat);
    HC(h, hipGetLastError());
    HC(h, hipMemcpyAsync(host, h->d_feat, sizeof(float) * (size_t)nb * ranker_din(h), hipMemcpyDeviceToHost, h->st));
    HC(h, hipStreamSynchronize(h->st));
    return PNGF_OK;
}

// Appends (inputs, exact |S11_f|^2 changes) of the last scored batch to the training log.
int log_batch(pngf_handle* h, const pngf_dbs_params* p, int nb, int64_t* logged) {
    if (!p->log_x || !p->log_y || *logged >= p->log_cap) return PNGF_OK;
    const int n = (int)std::min<int64_t>(nb, p->log_cap - *logged);
    const int din = ranker_din(h), nf = h->d.nf;
    CK(ensure_feat(h));
    launch_labels(h->st, h->d, h->vecs, h->op, h->d_eobj_out, n, h->d_lab);
    HC(h, hipMemcpyAsync(p->log_y + (size_t)*logged * nf, h->d_lab, sizeof(float) * (size_t)n * nf,
                         hipMemcpyDeviceToHost, h->st));
    CK(features_batch(h, n, p->log_x + (size_t)*logged * din));
    *logged += n;
    return PNGF_OK;
}

// One record of the reference Logger (logger.hpp:64-73; main.cpp:304-328, 401-426): tile,
// decision, F, F_best before the evaluation, then per frequency Re S11, Im S11 and D.
void hist_record(pngf_handle* h, const pngf_dbs_params* p, int64_t* cnt, int tile, int dec, double F,
                 double F_best, const zd* e) {
    if (*cnt >= p->hist_cap) return;
    const int nf = h->d.nf, ne = h->d.ne;
    double* pr = p->hist_perf + (size_t)*cnt * (2 + 3 * nf);
    pr[0] = F;
    pr[1] = F_best;
    for (int f = 0; f < nf; f++) {
        double D;
        const zd s = s11_of(h->op, h->fc.data(), e + (size_t)f * ne, f, &D);
        pr[2 + 3 * f] = s.x;
        pr[3 + 3 * f] = s.y;
        pr[4 + 3 * f] = D;
    }
    p->hist_tile[*cnt] = tile;
    p->hist_dec[*cnt] = dec;
    (*cnt)++;
}

bool hist_on(const pngf_dbs_params* p) { return p->hist_tile && p->hist_dec && p->hist_perf && p->hist_cap > 0; }

// Records for the first n candidates of the batch just scored (the ones the sequential loop
// evaluates: up to and including the accepted one), all against the same F_best.
int hist_batch(pngf_handle* h, const pngf_dbs_params* p, const int* groups, int n, double F_best, int64_t* cnt) {
    if (!hist_on(p) || *cnt >= p->hist_cap) return PNGF_OK;
    n = (int)std::min<int64_t>(n, p->hist_cap - *cnt);
    const size_t per = (size_t)h->d.nf * h->d.ne;
    std::vector<zd> e((size_t)n * per);
    std::vector<double> F(n);
    HC(h, hipMemcpyAsync(e.data(), h->d_eobj_out, e.size() * sizeof(zd), hipMemcpyDeviceToHost, h->st));
    HC(h, hipMemcpyAsync(F.data(), h->d_F, F.size() * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HC(h, hipStreamSynchronize(h->st));
    for (int b = 0; b < n; b++)
        hist_record(h, p, cnt, h->grp.prim[groups[b]], F[b] < F_best ? 1 : 0, F[b], F_best, e.data() + b * per);
    return PNGF_OK;
}

int download_eobj(pngf_handle* h, std::vector<zd>& e) {
    e.resize((size_t)h->d.nf * h->d.ne);
    HC(h, hipMemcpyAsync(e.data(), h->vecs.Eobj, e.size() * sizeof(zd), hipMemcpyDeviceToHost, h->st));
    HC(h, hipStreamSynchronize(h->st));
    return PNGF_OK;
}

// ---------------------------------------------------------- driver position ---
// Words of pngf_dbs_params.driver_state: [0] magic, [1] driver, [2] raster position in the
// visit cycle, [3] raster non-improving count, [4] reference driver's last tile tried, [5] hat
// size, [6] draws left in the hat, [7] W = words of the generator state, then W words of the
// std::mt19937 state (its standard text form: 624 state words and the position) and the hat's
// index array.
constexpr uint32_t kDrvMagic = 0x50474E44u;
constexpr int kDrvHead = 8;

int drv_save(pngf_handle* h, const pngf_dbs_params* p, int pos, int64_t nonimp, const Hat* hat, int last_tried) {
    if (!p->driver_state) return PNGF_OK;
    uint32_t* w = p->driver_state;
    std::memset(w, 0, PNGF_DRIVER_STATE_WORDS * sizeof(uint32_t));
    w[0] = kDrvMagic;
    w[1] = (uint32_t)h->cfg.driver;
    w[2] = (uint32_t)pos;
    w[3] = (uint32_t)nonimp;
    w[4] = (uint32_t)last_tried;
    if (hat) {
        std::ostringstream os;
        os << hat->gen;
        std::istringstream is(os.str());
        std::vector<uint32_t> gw;
        unsigned long long x;
        while (is >> x) gw.push_back((uint32_t)x);
        if (kDrvHead + gw.size() + (size_t)hat->total > (size_t)PNGF_DRIVER_STATE_WORDS)
            return fail(h, PNGF_ERR_ARG, "driver position does not fit PNGF_DRIVER_STATE_WORDS");
        w[5] = (uint32_t)hat->total;
        w[6] = (uint32_t)hat->left;
        w[7] = (uint32_t)gw.size();
        for (size_t i = 0; i < gw.size(); i++) w[kDrvHead + i] = gw[i];
        for (int i = 0; i < hat->total; i++) w[kDrvHead + gw.size() + i] = (uint32_t)hat->idx[i];
    }
    return PNGF_OK;
}

// Reads a saved position (resume = 1); the checks keep every index in range.
int drv_load(pngf_handle* h, const pngf_dbs_params* p, int ncl, int* pos, int64_t* nonimp, Hat* hat,
             int* last_tried) {
    const uint32_t* w = p->driver_state;
    if (!w || w[0] != kDrvMagic || (int)w[1] != h->cfg.driver)
        return fail(h, PNGF_ERR_ARG, "resume: no driver position of this driver in driver_state");
    if (h->cfg.driver == PNGF_DRIVER_RASTER) {
        if ((int)w[2] < 0 || (int)w[2] >= std::max(ncl, 1) || (int64_t)w[3] > ncl)
            return fail(h, PNGF_ERR_ARG, "resume: raster position out of range");
        *pos = (int)w[2];
        *nonimp = (int64_t)w[3];
        return PNGF_OK;
    }
    const int total = (int)w[5], left = (int)w[6], nw = (int)w[7];
    if (total != hat->total || left < 0 || left > total || nw < 1 ||
        kDrvHead + (size_t)nw + (size_t)total > (size_t)PNGF_DRIVER_STATE_WORDS)
        return fail(h, PNGF_ERR_ARG, "resume: hat state does not match this problem");
    std::ostringstream os;
    for (int i = 0; i < nw; i++) os << w[kDrvHead + i] << ' ';
    std::istringstream is(os.str());
    is >> hat->gen;
    if (is.fail()) return fail(h, PNGF_ERR_ARG, "resume: unreadable generator state");
    for (int i = 0; i < total; i++) {
        const int v = (int)w[kDrvHead + nw + i];
        if (v < 0 || v >= total) return fail(h, PNGF_ERR_ARG, "resume: hat index out of range");
        hat->idx[i] = v;
    }
    hat->left = left;
    const int lt = (int)w[4];
    if (lt < -1 || lt >= total) return fail(h, PNGF_ERR_ARG, "resume: last tile tried out of range");
    *last_tried = lt;
    return PNGF_OK;
}

}  // namespace

// ===================================================================== C ABI
extern "C" {

int pngf_layout_size(int32_t layout, int32_t tx, int32_t ty, int32_t cx, int32_t cy) {
    if ((layout != 0 && layout != 1) || tx < 1 || ty < 1 || cx < 1 || cy < 1) return PNGF_ERR_ARG;
    Layout L;
    L.setup(layout, tx, ty, cx, cy);
    return L.N;
}

const char* pngf_last_error(const pngf_handle* h) { return h ? h->err.c_str() : "null handle"; }

int pngf_create(const pngf_config* cfg, pngf_handle** out) {
    if (!cfg || !out) return PNGF_ERR_ARG;
    *out = nullptr;
    const pngf_config& c = *cfg;
    if ((c.layout != 0 && c.layout != 1) || c.tiles_x < 1 || c.tiles_y < 1 || c.cells_x < 1 || c.cells_y < 2 ||
        c.num_eobj < 1 || c.num_eobj > 3 || c.num_freq < 1 || c.max_flips < 1 || c.max_flips > 4 ||
        (c.method != 0 && c.method != 1) || (c.objective != 0 && c.objective != 1) || (c.driver != 0 && c.driver != 1))
        return PNGF_ERR_ARG;
    if (c.objective == PNGF_OBJ_REF && c.num_eobj != 3) return PNGF_ERR_ARG;
    if (c.driver == PNGF_DRIVER_REF && (c.max_flips < 4 || c.tiles_x % 2 == 0 || c.tiles_y % 2 == 0)) return PNGF_ERR_ARG;
    pngf_handle* h = new pngf_handle();
    h->cfg = c;
    if (h->cfg.z0 == 0.0) h->cfg.z0 = 50.0;
    if (h->cfg.dx == 0.0) { h->cfg.dx = 0.01 / 60.0; h->cfg.dy = 0.01 / 60.0; h->cfg.dz = 0.01 / 180.0; }
    if (!(h->cfg.synth_soft > 0.0)) h->cfg.synth_soft = kSoftDefault;
    h->L.setup(c.layout, c.tiles_x, c.tiles_y, c.cells_x, c.cells_y);
    if (h->L.per_tile * c.max_flips > 96 || c.tile_variant < 0 || c.tile_variant > 6 ||
        (c.symmetry != PNGF_SYM_AUTO && c.symmetry != PNGF_SYM_OFF && c.symmetry != PNGF_SYM_PROJECT) ||
        (c.symmetry == PNGF_SYM_PROJECT && c.method != PNGF_METHOD_LOWRANK) ||
        c.factors < PNGF_FACTORS_AUTO || c.factors > PNGF_FACTORS_REBUILT ||
        (c.precision != PNGF_PREC_C128 && c.precision != PNGF_PREC_C64) ||
        (c.precision == PNGF_PREC_C64 && c.method != PNGF_METHOD_LOWRANK)) {
        delete h;
        return PNGF_ERR_ARG;
    }
    h->c64 = c.precision == PNGF_PREC_C64;
    h->esz = h->c64 ? sizeof(zf) : sizeof(zd);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= c.device) {
        (void)hipGetLastError();
        h->err = "no HIP device";
        *out = h;
        return PNGF_ERR_HIP;
    }
    *out = h;
    HC(h, hipSetDevice(c.device));
    HC(h, hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
    HC(h, hipStreamCreateWithFlags(&h->st2, hipStreamNonBlocking));
    HC(h, hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
    HC(h, hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming));
    RB(h, rocblas_create_handle(&h->rb));
    RB(h, rocblas_set_stream(h->rb, h->st));
    RB(h, rocblas_set_pointer_mode(h->rb, rocblas_pointer_mode_host));
    const int N = h->L.N, nf = c.num_freq, ne = c.num_eobj;
    h->d.N = N;
    h->d.ldN = (int)round_up(N, 8);
    h->d.nf = nf;
    h->d.ne = ne;
    h->freqs.assign(nf, kFref);
    h->g_loaded.assign(nf, 0);
    h->fc.assign(nf, FreqConst{});
    h->gobj_extra.assign(nf, std::vector<zd>());
    h->Gcm.assign(nf, nullptr); h->Grow.assign(nf, nullptr); h->Cinv.assign(nf, nullptr);
    h->W0.assign(nf, nullptr); h->Aall.assign(nf, nullptr); h->Mall.assign(nf, nullptr);
    h->tile.assign(h->L.Tx * h->L.Ty, 0);
    h->P.assign(N, 0);
    h->owners.resize(2 * (size_t)N);
    for (int i = 0; i < N; i++) h->L.owners(i, h->owners[2 * i], h->owners[2 * i + 1]);
    h->op.mode = c.objective; h->op.nf = nf; h->op.ne = ne;
    h->op.zfac = h->cfg.dx * 0.5 / (2.0 * h->cfg.dy * h->cfg.dz);
    h->op.pfac = h->cfg.dx * 0.5 * (2.0 * h->cfg.dy * h->cfg.dz) * 0.5;
    h->op.z0 = h->cfg.z0;
    synth_positions(h->L, h->cfg.dx, h->cfg.dy, h->pos);
    synth_norm(h->L, h->pos, h->cfg.dx, h->cfg.dy, h->cfg.dz, h->cfg.synth_soft, h->synth_scale, h->synth_beta);
    // small device state
    const size_t ld = h->d.ldN;
    CK(dalloc_t(h, &h->d_tile, h->tile.size()));
    CK(dalloc_t(h, &h->d_P, N));
    CK(dalloc_t(h, &h->d_owners, 2 * (size_t)N));
    CK(dalloc_t(h, &h->d_small, 512));
    CK(dalloc_t(h, &h->d_sel, 4));
    CK(dalloc_t(h, &h->d_info, 64));
    CK(dalloc_t(h, &h->d_pos, 3 * (size_t)N));
    CK(dalloc_t(h, &h->d_flag, 4));

    CK(dalloc_t(h, &h->d_Fbest, 4));
    CK(dalloc_t(h, &h->d_Fsel, 4));
    CK(zalloc(h, &h->d_pe, ld));
    CK(dalloc_t(h, &h->d_tabs, 6 * (size_t)nf));
    CK(zalloc(h, &h->vecs.q, nf * ld));
    CK(zalloc(h, &h->vecs.S, nf * ld));
    CK(zalloc(h, &h->vecs.R, nf * ne * ld));
    CK(zalloc(h, &h->vecs.Gobj, nf * ne * ld));
    CK(zalloc(h, &h->vecs.Einc, nf * ld));
    CK(dalloc_t(h, &h->vecs.Eobj, (size_t)nf * ne));
    CK(dalloc_t(h, &h->vecs.fc, nf));
    HC(h, hipMemset(h->vecs.Gobj, 0, nf * ne * ld * h->esz));
    HC(h, hipMemcpy(h->d_owners, h->owners.data(), h->owners.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HC(h, hipMemcpy(h->d_pos, h->pos.data(), h->pos.size() * sizeof(double), hipMemcpyHostToDevice));
    HC(h, hipHostMalloc((void**)&h->h_pin_i, kPinInts * sizeof(int32_t), hipHostMallocDefault));
    HC(h, hipHostMalloc((void**)&h->h_pin_d, 64 * sizeof(double), hipHostMallocDefault));
    return PNGF_OK;
}

int pngf_destroy(pngf_handle* h) {
    if (!h) return PNGF_ERR_ARG;
    if (h->st) (void)hipStreamSynchronize(h->st);
    for (auto& a : h->allocs) (void)hipFree(a.first);
    h->allocs.clear();
    for (auto e : h->pool) (void)hipEventDestroy(e);
    for (auto& t : h->pending) { (void)hipEventDestroy(t.a); (void)hipEventDestroy(t.b); }
    if (h->h_pin_i) (void)hipHostFree(h->h_pin_i);
    if (h->h_pin_d) (void)hipHostFree(h->h_pin_d);
    if (h->h_loop) (void)hipHostFree(h->h_loop);
    if (h->h_iterlog) (void)hipHostFree(h->h_iterlog);
    for (auto e : h->lev) if (e) (void)hipEventDestroy(e);
    for (auto e : h->levr) if (e) (void)hipEventDestroy(e);
    if (h->loop_exec) (void)hipGraphExecDestroy(h->loop_exec);
    if (h->loop_graph) (void)hipGraphDestroy(h->loop_graph);
    if (h->rb) rocblas_destroy_handle(h->rb);
    if (h->st2) (void)hipStreamSynchronize(h->st2);
    if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
    if (h->ev_join) (void)hipEventDestroy(h->ev_join);
    if (h->st2) (void)hipStreamDestroy(h->st2);
    if (h->st) (void)hipStreamDestroy(h->st);
    delete h;
    return PNGF_OK;
}

int pngf_set_frequency(pngf_handle* h, int32_t f, double hz) {
    if (!h || f < 0 || f >= h->d.nf || !(hz > 0)) return fail(h, PNGF_ERR_ARG, "bad frequency slot or value");
    h->freqs[f] = hz;
    return PNGF_OK;
}

int pngf_set_g(pngf_handle* h, int32_t f, const pngf_c128* file_buf) {
    if (!h || !file_buf || f < 0 || f >= h->d.nf) return fail(h, PNGF_ERR_ARG, "bad G slot or buffer");
    if (h->inited) return fail(h, PNGF_ERR_STATE, "G cannot change after pngf_init_state");
    const int N = h->d.N, ne = h->d.ne;
    const size_t M = (size_t)N + ne - 1;
    CK(alloc_gcm(h, f));
    HC(h, hipMemcpy2D(h->Gcm[f], h->d.ldN * sizeof(zd), file_buf, M * sizeof(zd), N * sizeof(zd), N,
                      hipMemcpyHostToDevice));
    h->gobj_extra[f].assign((size_t)(ne - 1) * N, zmk(0.0, 0.0));
    const zd* fb = (const zd*)file_buf;
    for (int r = 0; r < ne - 1; r++)
        for (int n = 0; n < N; n++) h->gobj_extra[f][(size_t)r * N + n] = fb[(N + r) + M * n];
    h->g_loaded[f] = 1;
    return PNGF_OK;
}

int pngf_load_g(pngf_handle* h, int32_t f, const char* path) {
    if (!h || !path || f < 0 || f >= h->d.nf) return fail(h, PNGF_ERR_ARG, "bad G slot or path");
    if (h->inited) return fail(h, PNGF_ERR_STATE, "G cannot change after pngf_init_state");
    const int N = h->d.N, ne = h->d.ne;
    const size_t M = (size_t)N + ne - 1;
    struct stat sb;
    if (stat(path, &sb) != 0) return fail(h, PNGF_ERR_IO, std::string("G file not found: ") + path);
    if ((size_t)sb.st_size != M * M * sizeof(zd))
        return fail(h, PNGF_ERR_SIZE, std::string("G file ") + path + " has " + std::to_string(sb.st_size) +
                                          " bytes, expected " + std::to_string(M * M * sizeof(zd)));
    CK(alloc_gcm(h, f));
    // Two pinned staging buffers of whole file columns (~64 MB each): the read of one chunk
    // overlaps the upload of the previous one (hipMemcpy2DAsync on the handle's stream, which
    // also drops the Gobj rows N.. of each column and pads the columns to ldN).
    const size_t cols = std::max<size_t>(1, (64u << 20) / (M * sizeof(zd)));
    zd* stage[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    FILE* fp = nullptr;
    int rc = PNGF_OK;
    for (int i = 0; i < 2 && rc == PNGF_OK; i++) {
        if (hipHostMalloc((void**)&stage[i], cols * M * sizeof(zd), hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            stage[i] = nullptr;
            rc = fail(h, PNGF_ERR_NOMEM, "pinned staging buffer");
        } else if (hipEventCreateWithFlags(&done[i], hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            done[i] = nullptr;
            rc = fail(h, PNGF_ERR_HIP, "staging event");
        }
    }
    if (rc == PNGF_OK) {
        fp = fopen(path, "rb");
        if (!fp) rc = fail(h, PNGF_ERR_IO, std::string("cannot open ") + path);
    }
    if (rc == PNGF_OK) h->gobj_extra[f].assign((size_t)(ne - 1) * N, zmk(0.0, 0.0));
    int b = 0;
    for (size_t c0 = 0, chunk = 0; c0 < M && rc == PNGF_OK; c0 += cols, chunk++, b ^= 1) {
        const size_t nc = std::min(cols, M - c0);
        // buffer b was last handed to the upload two chunks ago
        if (chunk >= 2) {
            hipError_t e = hipEventSynchronize(done[b]);
            if (e != hipSuccess) { rc = fail(h, PNGF_ERR_HIP, std::string("hipEventSynchronize: ") + hipGetErrorString(e)); break; }
        }
        if (fread(stage[b], sizeof(zd), nc * M, fp) != nc * M) { rc = fail(h, PNGF_ERR_IO, "short read of G file"); break; }
        for (size_t c = 0; c < nc; c++) {
            const size_t col = c0 + c;
            if (col >= (size_t)N) continue;
            for (int r = 0; r < ne - 1; r++) h->gobj_extra[f][(size_t)r * N + col] = stage[b][c * M + N + r];
        }
        const size_t ncol = std::min(nc, (size_t)N > c0 ? (size_t)N - c0 : 0);
        if (ncol) {
            hipError_t e = hipMemcpy2DAsync(h->Gcm[f] + c0 * h->d.ldN, h->d.ldN * sizeof(zd), stage[b],
                                            M * sizeof(zd), N * sizeof(zd), ncol, hipMemcpyHostToDevice, h->st);
            if (e == hipSuccess) e = hipEventRecord(done[b], h->st);
            if (e != hipSuccess) rc = fail(h, PNGF_ERR_HIP, std::string("hipMemcpy2DAsync: ") + hipGetErrorString(e));
        }
    }
    // the uploads must finish before the staging buffers go
    hipError_t es = h->st ? hipStreamSynchronize(h->st) : hipSuccess;
    if (es != hipSuccess && rc == PNGF_OK) rc = fail(h, PNGF_ERR_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(es));
    if (fp) fclose(fp);
    for (int i = 0; i < 2; i++) {
        if (stage[i]) (void)hipHostFree(stage[i]);
        if (done[i]) (void)hipEventDestroy(done[i]);
    }
    if (rc == PNGF_OK) h->g_loaded[f] = 1;
    return rc;
}

int pngf_synth_g(pngf_handle* h, int32_t f) {
    if (!h || f < 0 || f >= h->d.nf) return fail(h, PNGF_ERR_ARG, "bad G slot");
    if (h->inited) return fail(h, PNGF_ERR_STATE, "G cannot change after pngf_init_state");
    const int N = h->d.N, ne = h->d.ne;
    CK(alloc_gcm(h, f));
    const double hz = h->freqs[f];
    const zd scale = zscale(h->synth_scale, hz / kFref);
    launch_synth_g(h->st, N, h->d.ldN, h->d_pos, kr_of(hz), scale, h->cfg.synth_soft * h->cfg.dx, 0, h->Gcm[f]);
    HC(h, hipGetLastError());
    HC(h, hipStreamSynchronize(h->st));
    h->gobj_extra[f].assign((size_t)(ne - 1) * N, zmk(0.0, 0.0));
    const double b = h->synth_beta * hz / kFref;
    for (int r = 0; r < ne - 1; r++)
        for (int n = 0; n < N; n++) {
            const int o = (int)h->pos[3 * n + 2];
            h->gobj_extra[f][(size_t)r * N + n] = zmk((r == 0 ? (o == 0) : (o == 1)) ? b : 0.0, 0.0);
        }
    h->g_loaded[f] = 1;
    return PNGF_OK;
}

int pngf_perturb_g(pngf_handle* h, int32_t f, double rel, int32_t seed) {
    if (!h || f < 0 || f >= h->d.nf || !(rel >= 0.0) || seed < 0) return fail(h, PNGF_ERR_ARG, "bad perturbation");
    if (h->inited) return fail(h, PNGF_ERR_STATE, "G cannot change after pngf_init_state");
    if (!h->g_loaded[f]) return fail(h, PNGF_ERR_STATE, "G of this slot is not set");
    const int N = h->d.N;
    // amp = rel * max |G_ij| (|z| = |Re z| + |Im z|), the largest entry from the asymmetry kernel
    unsigned long long* m = nullptr;
    CK(dalloc_t(h, &m, 2));
    HC(h, hipMemsetAsync(m, 0, 2 * sizeof(unsigned long long), h->st));
    launch_asym(h->st, N, h->d.ldN, h->Gcm[f], m);
    unsigned long long hm[2];
    HC(h, hipMemcpyAsync(hm, m, sizeof(hm), hipMemcpyDeviceToHost, h->st));
    HC(h, hipStreamSynchronize(h->st));
    dfree(h, m);
    double amax;
    std::memcpy(&amax, &hm[1], sizeof(double));
    launch_perturb(h->st, N, h->d.ldN, h->Gcm[f], rel * amax, (unsigned long long)seed);
    HC(h, hipGetLastError());
    HC(h, hipStreamSynchronize(h->st));
    return PNGF_OK;
}

int pngf_synth_g_host(int32_t layout, int32_t tx, int32_t ty, int32_t cx, int32_t cy, int32_t ne, double hz,
                      double soft, pngf_c128* file_buf) {
    if ((layout != 0 && layout != 1) || tx < 1 || ty < 1 || cx < 1 || cy < 2 || ne < 1 || ne > 3 || !(hz > 0) ||
        !file_buf)
        return PNGF_ERR_ARG;
    Layout L;
    L.setup(layout, tx, ty, cx, cy);
    const double dx = 0.01 / 60.0, dy = dx, dz = 0.01 / 180.0;
    std::vector<double> pos;
    synth_positions(L, dx, dy, pos);
    if (!(soft > 0.0)) soft = kSoftDefault;
    zd scale; double beta;
    synth_norm(L, pos, dx, dy, dz, soft, scale, beta);
    scale = zscale(scale, hz / kFref);
    const double kr = kr_of(hz), soft2 = (soft * dx) * (soft * dx);
    const size_t N = L.N, M = N + ne - 1;
    zd* out = (zd*)file_buf;
    for (size_t n = 0; n < M; n++)
        for (size_t m = 0; m < M; m++) {
            zd v = zmk(0.0, 0.0);
            if (n < N && m < N) v = zmul(synth_entry(&pos[3 * m], &pos[3 * n], kr, soft2), scale);
            else if (n < N && m >= N) {
                const int o = (int)pos[3 * n + 2];
                const int r = (int)(m - N);
                v = zmk((r == 0 ? (o == 0) : (o == 1)) ? beta * hz / kFref : 0.0, 0.0);
            }
            out[m + M * n] = v;
        }
    return PNGF_OK;
}

int pngf_set_tile_map(pngf_handle* h, const int32_t* tm) {
    if (!h || !tm) return fail(h, PNGF_ERR_ARG, "null tile map");
    if (h->inited) return fail(h, PNGF_ERR_STATE, "tile map is fixed after pngf_init_state (use pngf_accept)");
    const int T2 = h->L.Tx * h->L.Ty;
    for (int i = 0; i < T2; i++) h->tile[i] = tm[i] ? 1 : 0;
    h->L.mask(h->tile.data(), h->P);
    HC(h, hipMemcpy(h->d_tile, h->tile.data(), T2 * sizeof(int32_t), hipMemcpyHostToDevice));
    HC(h, hipMemcpy(h->d_P, h->P.data(), h->P.size(), hipMemcpyHostToDevice));
    h->tile_set = true;
    return PNGF_OK;
}

int pngf_set_source_port(pngf_handle* h, int32_t tx, int32_t ty, int32_t yoff, int32_t ysz) {
    if (!h) return PNGF_ERR_ARG;
    if (h->inited) return fail(h, PNGF_ERR_STATE, "port is fixed after pngf_init_state");
    if (!h->L.port_ok(yoff, ysz) || tx < 0 || ty < 0 || tx >= h->L.Tx || ty >= h->L.Ty)
        return fail(h, PNGF_ERR_ARG, "set_source_port: invalid port y position and height parameters");
    if (h->L.kind == PNGF_LAYOUT_DISJOINT && ysz != 2)
        return fail(h, PNGF_ERR_ARG, "the disjoint layout has exactly 2 port rows");
    h->L.set_port(tx, ty, yoff, ysz);
    if (h->L.kind == PNGF_LAYOUT_DISJOINT) {
        synth_positions(h->L, h->cfg.dx, h->cfg.dy, h->pos);
        HC(h, hipMemcpy(h->d_pos, h->pos.data(), h->pos.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    synth_norm(h->L, h->pos, h->cfg.dx, h->cfg.dy, h->cfg.dz, h->cfg.synth_soft, h->synth_scale, h->synth_beta);
    return PNGF_OK;
}

int pngf_init_state(pngf_handle* h) {
    if (!h) return PNGF_ERR_ARG;
    if (h->inited) return fail(h, PNGF_ERR_STATE, "already initialised");
    if (!h->tile_set) return fail(h, PNGF_ERR_STATE, "set_tile_map first");
    for (int f = 0; f < h->d.nf; f++)
        if (!h->g_loaded[f]) return fail(h, PNGF_ERR_STATE, "G of frequency slot " + std::to_string(f) + " not set");
    const int nf = h->d.nf, ne = h->d.ne, N = h->d.N;
    const size_t ld = h->d.ldN;
    build_groups(h);
    const Groups& G = h->grp;
    h->d.ngroups = G.n;
    h->d.gmax = G.gmax;
    {
        std::vector<int> owner(h->d.N, -1);
        bool disjoint = true;
        for (int g = 0; g < h->ngroups_drv && disjoint; g++)
            for (int a = 0; a < G.ncomp[g]; a++) {
                int& o = owner[G.comps[(size_t)g * G.gmax + a]];
                if (o >= 0) { disjoint = false; break; }
                o = g;
            }
        std::vector<int32_t> left;
        for (int i = 0; i < h->d.N && disjoint; i++) if (owner[i] < 0) left.push_back(i);
        h->part = disjoint;
        h->nleft = (int)left.size();
        if (disjoint) {
            CK(dalloc_t(h, &h->d_left, std::max<size_t>(left.size(), 1)));
            HC(h, hipMemcpy(h->d_left, left.data(), left.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        }
    }
    CK(dalloc_t(h, &h->d_gtiles, G.tiles.size()));
    CK(dalloc_t(h, &h->d_gnt, G.ntiles.size()));
    CK(dalloc_t(h, &h->d_gcomps, G.comps.size()));
    CK(dalloc_t(h, &h->d_gnc, G.ncomp.size()));
    CK(dalloc_t(h, &h->d_batch, G.n + 8));
    HC(h, hipMemcpy(h->d_gtiles, G.tiles.data(), G.tiles.size() * 4, hipMemcpyHostToDevice));
    HC(h, hipMemcpy(h->d_gnt, G.ntiles.data(), G.ntiles.size() * 4, hipMemcpyHostToDevice));
    HC(h, hipMemcpy(h->d_gcomps, G.comps.data(), G.comps.size() * 4, hipMemcpyHostToDevice));
    HC(h, hipMemcpy(h->d_gnc, G.ncomp.data(), G.ncomp.size() * 4, hipMemcpyHostToDevice));
    std::vector<int32_t> portv(h->L.port.begin(), h->L.port.end());
    HC(h, hipMemcpy(h->d_small + 400, portv.data(), portv.size() * 4, hipMemcpyHostToDevice));
    const int gmax = G.gmax;
    CK(dalloc_t(h, &h->d_eobj_out, (size_t)G.n * nf * ne));
    CK(dalloc_t(h, &h->d_F, (size_t)G.n));
    dfree(h, h->d_info);
    CK(dalloc_t(h, &h->d_info, (size_t)G.n * nf + 64));
    if ((size_t)G.n * nf + 64 > (size_t)(kPinInts - kPinInfo) || G.n > kPinInfo - kPinBatch)
        return fail(h, PNGF_ERR_ARG, "design too large for the staging buffers");
    h->acc_stride = 2 * gmax * gmax + 2 * gmax + 4 + ne * gmax;
    if (gmax > kLdsGmax)
        CK(zalloc(h, &h->d_lu, (size_t)std::max(G.n, 1) * nf * gmax * (2 * gmax + 1)));
    CK(zalloc(h, &h->d_acc, (size_t)nf * h->acc_stride));
    if (h->cfg.method == PNGF_METHOD_LOWRANK)
        CK(zalloc(h, &h->vecs.Dblk, (size_t)nf * G.n * gmax * gmax));
    HC(h, hipMemcpy(h->vecs.fc, h->fc.data(), nf * sizeof(FreqConst), hipMemcpyHostToDevice));
    h->asym = 0.0;
    h->sym = false;
    h->d.sym = 0;
    if (h->cfg.method == PNGF_METHOD_LOWRANK && h->cfg.symmetry != PNGF_SYM_OFF && !h->d_asym)
        CK(dalloc_t(h, &h->d_asym, 2));
    h->sym_proj_ok = true;
    for (int f = 0; f < nf; f++) CK(init_freq(h, f));
    h->sym = h->cfg.method == PNGF_METHOD_LOWRANK &&
             ((h->cfg.symmetry == PNGF_SYM_AUTO && h->asym <= kSymTol) ||
              (h->cfg.symmetry == PNGF_SYM_PROJECT && h->sym_proj_ok));
    h->d.sym = h->sym ? 1 : 0;
    HC(h, hipMemcpy(h->vecs.fc, h->fc.data(), nf * sizeof(FreqConst), hipMemcpyHostToDevice));
    if (h->cfg.method == PNGF_METHOD_LOWRANK) {
        size_t freeb = 0, totb = 0;
        HC(h, hipMemGetInfo(&freeb, &totb));
        // keep room for the scratch strips and for one checkpoint copy of W0
        const size_t scratch = (size_t)nf * N * gmax * h->esz * 2 + (size_t)nf * ld * N * h->esz +
                               (size_t)kMChunk * ld * h->esz + (64u << 20);
        size_t avail = freeb > scratch ? freeb - scratch : 0;
        if (h->cfg.mem_limit > 0) {
            size_t lim = (size_t)h->cfg.mem_limit > h->mem_used + scratch ? (size_t)h->cfg.mem_limit - h->mem_used - scratch : 0;
            avail = std::min(avail, lim);
        }
        // Folding K deferred columns into W0 costs 8 N^2 K flop; not folding costs 8 N K k more per
        // later accept, so folding pays only when more than N / k accepts remain: by default the
        // factors hold N / gmax accepts (memory permitting) and a typical run never folds.
        int maxdef = h->cfg.max_deferred > 0 ? h->cfg.max_deferred : (N + gmax - 1) / gmax;
        // A column + M row.  When memory caps that below the wanted capacity (or the caller asks)
        // and W is symmetric with groups held in LDS, M rows are instead rebuilt from A and the
        // k x k Lam of their accept: a factor column then costs one A column (twice the capacity;
        // the gather that rebuilds M[:, t] costs ~15 us more per accept than reading stored M).
        const size_t want = (size_t)maxdef * gmax;
        const size_t per_col_m = (size_t)nf * 2 * ld * h->esz;
        const bool capped = (size_t)(0.85 * avail) / per_col_m < want;
        h->lam = h->sym && gmax <= kLdsGmax && h->cfg.factors != PNGF_FACTORS_STORED &&
                 (capped || h->cfg.factors == PNGF_FACTORS_REBUILT);
        h->d.lam = h->lam ? 1 : 0;
        const size_t per_col = h->lam ? (size_t)nf * (ld + gmax) * h->esz : per_col_m;
        size_t cols = std::min<size_t>(want, (size_t)(0.85 * avail) / per_col);
        cols = cols / gmax * gmax;
        if (cols < (size_t)gmax) return fail(h, PNGF_ERR_NOMEM, "no device memory for the deferred factors");
        h->Kmax = (int)cols;
        h->d.Kmax = h->Kmax;
        if (h->lam) h->Lam.assign(nf, nullptr);
        for (int f = 0; f < nf; f++) {
            CK(zalloc(h, &h->Aall[f], (size_t)N * h->Kmax));
            if (h->lam) CK(zalloc(h, &h->Lam[f], (size_t)h->Kmax * gmax));
            else CK(zalloc(h, &h->Mall[f], (size_t)h->Kmax * ld));
        }
        if (h->lam) {
            CK(dalloc_t(h, &h->d_fstart, h->Kmax));
            CK(dalloc_t(h, &h->d_flen, h->Kmax));
            CK(zalloc(h, &h->d_mchunk, (size_t)kMChunk * ld));
            HC(h, hipMemset(h->d_fstart, 0, sizeof(int32_t) * h->Kmax));
            HC(h, hipMemset(h->d_flen, 0, sizeof(int32_t) * h->Kmax));
        }
        h->small_stride = (size_t)h->Kmax * gmax;
        CK(zalloc(h, &h->d_Mc, nf * h->small_stride));
        CK(zalloc(h, &h->d_Ar, nf * h->small_stride));
        h->bt_stride = (size_t)N * gmax;
        CK(zalloc(h, &h->d_Bt, nf * h->bt_stride));
        for (int f = 0; f < nf; f++) {
            HC(h, hipMemset(h->Aall[f], 0, h->esz * (size_t)N * h->Kmax));
            if (h->lam) HC(h, hipMemset(h->Lam[f], 0, h->esz * (size_t)h->Kmax * gmax));
            else HC(h, hipMemset(h->Mall[f], 0, h->esz * (size_t)h->Kmax * ld));
        }
        HC(h, hipMemset(h->d_Mc, 0, h->esz * nf * h->small_stride));
        HC(h, hipMemset(h->d_Ar, 0, h->esz * nf * h->small_stride));
    } else {
        h->d.Kmax = 0;
        h->v_stride = (size_t)gmax * ld;
        CK(dalloc_t(h, &h->d_V, nf * h->v_stride));
        CK(dalloc_t(h, &h->d_VC, nf * h->v_stride));
        CK(dalloc_t(h, &h->d_CU, nf * h->v_stride));
        CK(dalloc_t(h, &h->d_M, nf * h->v_stride));
        CK(dalloc_t(h, &h->d_rut, (size_t)nf * ne * 96));
        const size_t items = std::max<size_t>(h->nsplit_items, (size_t)G.n * nf);
        h->partial_elems = items * (gmax + ne) * (gmax + 1);
        CK(dalloc_t(h, &h->d_partial, h->partial_elems));
    }
    CK(upload_tabs(h));
    if (h->cfg.method == PNGF_METHOD_LOWRANK) {
        if (h->cfg.tile_variant > 0) h->skv = h->cfg.tile_variant - 1;
        else CK(autotune_skinny(h));
    }
    std::vector<zd> e;
    CK(download_eobj(h, e));
    for (const zd& v : e)
        if (!std::isfinite(v.x) || !std::isfinite(v.y))
            return fail(h, PNGF_ERR_NONFINITE, "initial Eobj is not finite (C^-1 overflowed)");
    h->inited = true;
    return PNGF_OK;
}

int pngf_eval(pngf_handle* h, const int32_t* tiles, int32_t nt, pngf_c128* new_eobj, double* F) {
    if (!h) return PNGF_ERR_ARG;
    CK(check_ready(h));
    int g;
    CK(group_for(h, tiles, nt, &g));
    CK(eval_launch(h, &g, 1));
    int sel; double fs;
    CK(sync_sel(h, 1, &sel, &fs, true));
    const size_t n = (size_t)h->d.nf * h->d.ne;
    std::vector<zd> e(n);
    HC(h, hipMemcpy(e.data(), h->d_eobj_out, n * sizeof(zd), hipMemcpyDeviceToHost));
    if (new_eobj) std::memcpy(new_eobj, e.data(), n * sizeof(zd));
    if (F) HC(h, hipMemcpy(F, h->d_F, sizeof(double), hipMemcpyDeviceToHost));
    return PNGF_OK;
}

int pngf_eval_batch(pngf_handle* h, const int32_t* tiles, const int32_t* ntiles, int32_t ncand, double* F_out,
                    pngf_c128* eobj_out) {
    if (!h || !tiles || !ntiles || ncand < 0) return fail(h, PNGF_ERR_ARG, "bad batch");
    CK(check_ready(h));
    const int mf = h->cfg.max_flips;
    const size_t per = (size_t)h->d.nf * h->d.ne;
    std::vector<int> gs;
    std::vector<int> idx;
    auto flush_batch = [&]() -> int {
        if (gs.empty()) return PNGF_OK;
        CK(eval_launch(h, gs.data(), (int)gs.size()));
        int sel; double fs;
        CK(sync_sel(h, (int)gs.size(), &sel, &fs, true));
        std::vector<double> F(gs.size());
        std::vector<zd> e(gs.size() * per);
        HC(h, hipMemcpy(F.data(), h->d_F, F.size() * sizeof(double), hipMemcpyDeviceToHost));
        HC(h, hipMemcpy(e.data(), h->d_eobj_out, e.size() * sizeof(zd), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < gs.size(); i++) {
            if (F_out) F_out[idx[i]] = F[i];
            if (eobj_out) std::memcpy(eobj_out + (size_t)idx[i] * per, &e[i * per], per * sizeof(zd));
        }
        gs.clear(); idx.clear();
        return PNGF_OK;
    };
    for (int c = 0; c < ncand; c++) {
        int g;
        CK(group_for(h, tiles + (size_t)c * mf, ntiles[c], &g));
        if (g == h->adhoc) {
            CK(flush_batch());
            gs.push_back(g); idx.push_back(c);
            CK(flush_batch());
        } else {
            gs.push_back(g); idx.push_back(c);
            if ((int)gs.size() >= h->ngroups_drv) CK(flush_batch());
        }
    }
    return flush_batch();
}

int pngf_sweep(pngf_handle* h, const int32_t* tiles, const int32_t* ntiles, int32_t ncand, double F_best,
               int32_t* accepted_idx, double* F_new, int64_t* evaluated) {
    if (!h || !tiles || !ntiles || ncand < 0 || !accepted_idx) return fail(h, PNGF_ERR_ARG, "bad sweep");
    CK(check_ready(h));
    const int mf = h->cfg.max_flips;
    *accepted_idx = -1;
    if (F_new) *F_new = F_best;
    if (evaluated) *evaluated = ncand;
    h->h_pin_d[8] = F_best;
    HC(h, hipMemcpyAsync(h->d_Fbest, h->h_pin_d + 8, sizeof(double), hipMemcpyHostToDevice, h->st));
    std::vector<int> gs, idx;
    // scores the pending candidates as one batch; true when one of them improves
    auto flush = [&](bool* found) -> int {
        *found = false;
        if (gs.empty()) return PNGF_OK;
        CK(eval_launch(h, gs.data(), (int)gs.size()));
        int sel;
        double fs;
        CK(sync_sel(h, (int)gs.size(), &sel, &fs, true));
        if (sel >= 0) {
            *found = true;
            *accepted_idx = idx[sel];
            if (F_new) *F_new = fs;
            if (evaluated) *evaluated = (int64_t)idx[sel] + 1;
        }
        gs.clear();
        idx.clear();
        return PNGF_OK;
    };
    bool found = false;
    for (int c = 0; c < ncand && !found; c++) {
        int g;
        CK(group_for(h, tiles + (size_t)c * mf, ntiles[c], &g));
        if (g == h->adhoc) {
            // the ad-hoc slot holds one flip at a time: score what is pending, then this one alone
            CK(flush(&found));
            if (found) break;
            gs.push_back(g);
            idx.push_back(c);
            CK(flush(&found));
        } else {
            gs.push_back(g);
            idx.push_back(c);
            if ((int)gs.size() >= h->ngroups_drv) CK(flush(&found));
        }
    }
    if (!found) CK(flush(&found));
    return PNGF_OK;
}

int pngf_accept(pngf_handle* h, const int32_t* tiles, int32_t nt) {
    if (!h) return PNGF_ERR_ARG;
    CK(check_ready(h));
    int g;
    CK(group_for(h, tiles, nt, &g));
    CK(accept_group(h, g));
    HC(h, hipMemcpyAsync(h->h_pin_i + kPinInfo, h->d_info, sizeof(int32_t) * h->d.nf, hipMemcpyDeviceToHost, h->st));
    HC(h, hipStreamSynchronize(h->st));
    if (h->cfg.method == PNGF_METHOD_LOWRANK)
        for (int f = 0; f < h->d.nf; f++)
            if (h->h_pin_i[kPinInfo + f] != 0) return fail(h, PNGF_ERR_SINGULAR, "capacitance matrix is singular");
    HC(h, hipStreamSynchronize(h->st));
    prof_collect(h);
    return PNGF_OK;
}

int pngf_get_state(pngf_handle* h, pngf_c128* eobj, pngf_c128* vinc, pngf_c128* ln, pngf_c128* s11, double* F) {
    if (!h) return PNGF_ERR_ARG;
    CK(check_ready(h));
    std::vector<zd> e;
    CK(download_eobj(h, e));
    const int nf = h->d.nf, ne = h->d.ne;
    if (eobj) std::memcpy(eobj, e.data(), e.size() * sizeof(zd));
    for (int f = 0; f < nf; f++) {
        if (vinc) { vinc[f].re = h->fc[f].vinc.x; vinc[f].im = h->fc[f].vinc.y; }
        if (ln) { ln[2 * f] = {h->fc[f].ln1.x, h->fc[f].ln1.y}; ln[2 * f + 1] = {h->fc[f].ln2.x, h->fc[f].ln2.y}; }
        if (s11) {
            zd u = zadd(e[(size_t)f * ne], h->fc[f].vinc);
            zd zin = zmk(-u.x * h->op.zfac, -u.y * h->op.zfac);
            zd s = zdiv(zmk(zin.x - h->op.z0, zin.y), zmk(zin.x + h->op.z0, zin.y));
            s11[f] = {s.x, s.y};
        }
    }
    if (F) *F = host_objective(h, e.data());
    return PNGF_OK;
}

int pngf_get_tile_map(pngf_handle* h, int32_t* tm) {
    if (!h || !tm) return fail(h, PNGF_ERR_ARG, "null tile map");
    for (size_t i = 0; i < h->tile.size(); i++) tm[i] = h->tile[i];
    return PNGF_OK;
}

int pngf_get_cinv(pngf_handle* h, int32_t f, pngf_c128* out) {
    if (!h || !out || f < 0 || f >= h->d.nf) return fail(h, PNGF_ERR_ARG, "bad slot");
    CK(check_ready(h));
    if (!h->Cinv[f]) return fail(h, PNGF_ERR_STATE, "C^-1 is not kept by this method");
    HC(h, hipMemcpy2D(out, h->d.N * sizeof(zd), h->Cinv[f], h->d.ldN * sizeof(zd), h->d.N * sizeof(zd), h->d.N,
                      hipMemcpyDeviceToHost));
    return PNGF_OK;
}

int pngf_get_w(pngf_handle* h, int32_t f, pngf_c128* out) {
    if (!h || !out || f < 0 || f >= h->d.nf) return fail(h, PNGF_ERR_ARG, "bad slot");
    CK(check_ready(h));
    if (!h->W0[f]) return fail(h, PNGF_ERR_STATE, "W is kept by the LOWRANK method only");
    const int N = h->d.N;
    const size_t ld = h->d.ldN;
    zd* tmp = nullptr;
    CK(zalloc(h, &tmp, ld * N));
    HC(h, hipMemcpyAsync(tmp, h->W0[f], ld * N * h->esz, hipMemcpyDeviceToDevice, h->st));
    if (h->K > 0) CK(fold_into(h, f, tmp));
    HC(h, hipStreamSynchronize(h->st));
    if (h->c64) {
        // complex64 state, returned widened to complex128
        std::vector<zf> w((size_t)N * N);
        HC(h, hipMemcpy2D(w.data(), N * sizeof(zf), tmp, ld * sizeof(zf), N * sizeof(zf), N, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < w.size(); i++) { out[i].re = w[i].x; out[i].im = w[i].y; }
    } else {
        HC(h, hipMemcpy2D(out, N * sizeof(zd), tmp, ld * sizeof(zd), N * sizeof(zd), N, hipMemcpyDeviceToHost));
    }
    dfree(h, tmp);
    return PNGF_OK;
}

int pngf_get_g(pngf_handle* h, int32_t f, pngf_c128* out) {
    if (!h || !out || f < 0 || f >= h->d.nf) return fail(h, PNGF_ERR_ARG, "bad slot");
    CK(check_ready(h));
    if (!h->Grow[f]) return fail(h, PNGF_ERR_STATE, "G is kept by the PLAIN method only");
    // Grow is row-major G = column-major G^T; return column-major G via a device transpose
    const int N = h->d.N;
    const size_t ld = h->d.ldN;
    zd* tmp = nullptr;
    CK(dalloc_t(h, &tmp, ld * N));
    launch_transpose(h->st, N, (int)ld, h->Grow[f], tmp);
    HC(h, hipStreamSynchronize(h->st));
    HC(h, hipMemcpy2D(out, N * sizeof(zd), tmp, ld * sizeof(zd), N * sizeof(zd), N, hipMemcpyDeviceToHost));
    dfree(h, tmp);
    return PNGF_OK;
}

// Device-side snapshot of the whole mutable state (LOWRANK: W0 with the pending
// factors folded in, blocks, q, R, Eobj; PLAIN: C^-1, S, R, Eobj; both: mask and
// tile map).  pngf_restore returns to it; used to repeat a DBS run from the same
// start and as the resume point of a long run.
int pngf_checkpoint(pngf_handle* h) {
    if (!h) return PNGF_ERR_ARG;
    CK(check_ready(h));
    const Dims& d = h->d;
    const size_t ld = d.ldN, big = ld * d.N;
    const bool lr = h->cfg.method == PNGF_METHOD_LOWRANK;
    if (lr) CK(flush_factors(h));
    // the snapshot's vector part: q (or S), R in the state's precision, then Eobj in complex128
    const size_t qb = (size_t)d.nf * ld * h->esz, rb = (size_t)d.nf * d.ne * ld * h->esz;
    if (!h->ckpt) {
        h->ck_big.assign(d.nf, nullptr);
        for (int f = 0; f < d.nf; f++) CK(zalloc(h, &h->ck_big[f], big));
        CK(dalloc(h, (void**)&h->ck_vec, qb + rb + (size_t)d.nf * d.ne * sizeof(zd)));
        if (lr) CK(zalloc(h, &h->ck_D, (size_t)d.nf * d.ngroups * d.gmax * d.gmax));
        CK(dalloc_t(h, &h->ck_P, d.N));
    }
    hipStream_t st = h->st;
    for (int f = 0; f < d.nf; f++)
        HC(h, hipMemcpyAsync(h->ck_big[f], lr ? h->W0[f] : h->Cinv[f], big * h->esz, hipMemcpyDeviceToDevice, st));
    char* v = (char*)h->ck_vec;
    HC(h, hipMemcpyAsync(v, lr ? h->vecs.q : h->vecs.S, qb, hipMemcpyDeviceToDevice, st));
    HC(h, hipMemcpyAsync(v + qb, h->vecs.R, rb, hipMemcpyDeviceToDevice, st));
    HC(h, hipMemcpyAsync(v + qb + rb, h->vecs.Eobj, d.nf * d.ne * sizeof(zd), hipMemcpyDeviceToDevice, st));
    if (lr) HC(h, hipMemcpyAsync(h->ck_D, h->vecs.Dblk, (size_t)d.nf * d.ngroups * d.gmax * d.gmax * h->esz, hipMemcpyDeviceToDevice, st));
    HC(h, hipMemcpyAsync(h->ck_P, h->d_P, d.N, hipMemcpyDeviceToDevice, st));
    HC(h, hipStreamSynchronize(st));
    h->ck_tile = h->tile;
    h->ck_Ph = h->P;
    h->ckpt = true;
    return PNGF_OK;
}

int pngf_restore(pngf_handle* h) {
    if (!h) return PNGF_ERR_ARG;
    CK(check_ready(h));
    if (!h->ckpt) return fail(h, PNGF_ERR_STATE, "no checkpoint");
    const Dims& d = h->d;
    const size_t ld = d.ldN, big = ld * d.N;
    const bool lr = h->cfg.method == PNGF_METHOD_LOWRANK;
    hipStream_t st = h->st;
    const size_t qb = (size_t)d.nf * ld * h->esz, rb = (size_t)d.nf * d.ne * ld * h->esz;
    for (int f = 0; f < d.nf; f++)
        HC(h, hipMemcpyAsync(lr ? h->W0[f] : h->Cinv[f], h->ck_big[f], big * h->esz, hipMemcpyDeviceToDevice, st));
    const char* v = (const char*)h->ck_vec;
    HC(h, hipMemcpyAsync(lr ? h->vecs.q : h->vecs.S, v, qb, hipMemcpyDeviceToDevice, st));
    HC(h, hipMemcpyAsync(h->vecs.R, v + qb, rb, hipMemcpyDeviceToDevice, st));
    HC(h, hipMemcpyAsync(h->vecs.Eobj, v + qb + rb, d.nf * d.ne * sizeof(zd), hipMemcpyDeviceToDevice, st));
    if (lr) HC(h, hipMemcpyAsync(h->vecs.Dblk, h->ck_D, (size_t)d.nf * d.ngroups * d.gmax * d.gmax * h->esz, hipMemcpyDeviceToDevice, st));
    HC(h, hipMemcpyAsync(h->d_P, h->ck_P, d.N, hipMemcpyDeviceToDevice, st));
    h->tile = h->ck_tile;
    h->P = h->ck_Ph;
    HC(h, hipMemcpyAsync(h->d_tile, h->tile.data(), h->tile.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HC(h, hipStreamSynchronize(st));
    h->K = 0;
    return PNGF_OK;
}

int pngf_profile(pngf_handle* h, int32_t enable) {
    if (!h) return PNGF_ERR_ARG;
    h->prof = enable != 0;
    for (auto& k : h->kst) k = pngf_kstat{};
    h->acc_seconds = 0.0;
    return PNGF_OK;
}

int pngf_kernel_stats(pngf_handle* h, int32_t which, pngf_kstat* out) {
    if (!h || !out || which < 0 || which >= PNGF_K_COUNT) return fail(h, PNGF_ERR_ARG, "bad kernel id");
    *out = h->kst[which];
    return PNGF_OK;
}

// ------------------------------------------------ device-driven raster loop ---
// The raster first-improvement iteration of pngf_dbs_run with batch construction, selection
// and the accept bookkeeping in kernels (LoopState in device memory; the accept kernels read
// k, K and whether an accept is pending from it).  The host enqueues up to kLoopChunk
// iterations, each of which accepts at most one flip, and synchronises once per chunk; the
// factor capacity is checked (and folded) between chunks.  Same batches, same selection and
// the same kernels as the host loop, so the accepted sequence is identical.
constexpr int kLoopChunk = kLoopChunkMax;
// host threads of pngf_dbs_run_many
constexpr int kDefaultWorkers = 4;

bool device_loop_ok(const pngf_handle* h, const pngf_dbs_params* p) {
    return p->loop == PNGF_LOOP_AUTO && h->cfg.driver == PNGF_DRIVER_RASTER && p->order == PNGF_ORDER_FIRST &&
           h->cfg.method == PNGF_METHOD_LOWRANK && !p->log_x && !hist_on(p) && h->d.gmax <= kLdsGmax;
}

int device_loop_prepare(pngf_handle* h, const pngf_dbs_params* p, double F_best, int bmin, int bmax, int pos = 0,
                        int nonimp = 0) {
    if (!h->d_loop) {
        CK(dalloc_t(h, &h->d_loop, 1));
        HC(h, hipHostMalloc((void**)&h->h_loop, sizeof(LoopState), hipHostMallocDefault));
        std::vector<int32_t> cl;
        for (int g = 0; g < h->ngroups_drv; g++) if (h->grp.ntiles[g] > 0) cl.push_back(g);
        CK(dalloc_t(h, &h->d_cl, std::max<size_t>(cl.size(), 1)));
        HC(h, hipMemcpy(h->d_cl, cl.data(), cl.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    const int cap = std::max(1, p->max_accepts);
    if (cap > h->acc_alloc) {
        // the captured iterations hold the old record arrays: capture again
        if (h->loop_exec) (void)hipGraphExecDestroy(h->loop_exec);
        if (h->loop_graph) (void)hipGraphDestroy(h->loop_graph);
        h->loop_exec = nullptr;
        h->loop_graph = nullptr;
        dfree(h, h->d_accg);
        dfree(h, h->d_acce);
        dfree(h, h->d_accF);
        h->d_accg = nullptr;
        h->d_acce = nullptr;
        h->d_accF = nullptr;
        CK(dalloc_t(h, &h->d_accg, cap));
        CK(dalloc_t(h, &h->d_acce, cap));
        CK(dalloc_t(h, &h->d_accF, cap));
        h->acc_alloc = cap;
    }
    if (h->prof && !h->d_iterlog) {
        CK(dalloc_t(h, &h->d_iterlog, 3 * kLoopChunk));
        HC(h, hipHostMalloc((void**)&h->h_iterlog, 3 * kLoopChunk * sizeof(int32_t), hipHostMallocDefault));
        for (auto& e : h->lev) HC(h, hipEventCreate(&e));
        for (auto& e : h->levr) HC(h, hipEventCreate(&e));
    }
    LoopState& ls = *h->h_loop;
    std::memset(&ls, 0, sizeof(ls));
    ls.max_evals = p->max_evals;
    ls.F_best = F_best;
    ls.F_sel = F_best;
    ls.bs = bmin;
    ls.bmin = bmin;
    ls.bmax = bmax;
    ls.acc_cap = p->max_accepts;
    int ncl = 0;
    for (int g = 0; g < h->ngroups_drv; g++) ncl += h->grp.ntiles[g] > 0 ? 1 : 0;
    ls.ncl = ncl;
    ls.sel = -1;
    ls.pos = pos;
    ls.nonimp = nonimp;
    ls.dyn[kDynKK] = h->K;
    HC(h, hipMemcpyAsync(h->d_loop, &ls, sizeof(LoopState), hipMemcpyHostToDevice, h->st));
    launch_loop_batch(h->st, h->d_loop, h->d_cl, h->d_batch);
    return PNGF_OK;
}

// Iteration i of a chunk.  With profiling on, events bracket the column strip and the control
// kernel records (pending, k, K) for device_loop_collect.
extern "C++" template <class Z>
void device_loop_iteration_t(pngf_handle* h, int i) {
    const LRView<Z> w = lrview<Z>(h);
    hipStream_t st = h->st;
    const bool tm = h->prof && h->d_iterlog;
    const Dims& d = h->d;
    LoopState* ls = h->d_loop;
    const int32_t* dyn = ls->dyn;
    int32_t* dt = h->d_small;
    int32_t* ds = h->d_small + 96;
    int32_t* dlp = h->d_small + 192;
    int32_t* dgt = h->d_small + 288;
    const int ncl = h->h_loop->ncl;
    LoopCtl c{};
    c.ls = ls;
    c.op = h->op;
    c.fc = h->vecs.fc;
    c.F = h->d_F;
    c.cl = h->d_cl;
    c.t = dt;
    c.s = ds;
    c.lp = dlp;
    c.gtiles = dgt;
    c.acc_group = h->d_accg;
    c.acc_eval = h->d_acce;
    c.acc_F = h->d_accF;
    launch_lr_eval(st, d, w.v, gtab(h), h->d_tile, h->d_P, h->d_owners, h->d_batch, ncl, h->d_eobj_out,
                   h->d_info, w.lu, &ls->nb);
    launch_loop_control(st, d, gtab(h), h->d_tile, h->d_P, h->d_owners, h->d_batch, h->d_eobj_out, h->d_info, c,
                        tm ? h->d_iterlog + 3 * i : nullptr);
    (void)hipEventRecord(h->ev_fork, st);
    (void)hipStreamWaitEvent(h->st2, h->ev_fork, 0);
    launch_lr_accept_prep(h->st2, d, w.v, 0, dt, ds, dlp, d.gmax, w.acc, h->acc_stride, h->d_info, w.lu, dyn, &ls->g);
    (void)hipEventRecord(h->ev_join, h->st2);
    // host k = gmax selects the column-tile count; the kernels read the actual k and K
    launch_lr_gather(st, d, w.tb, dt, d.gmax, 0, w.Mc, h->sym ? (Z*)nullptr : w.Ar, (int)h->small_stride, dyn);
    if (tm) (void)hipEventRecord(h->lev[2 * i], st);
    launch_lr_colblk(st, d, w.tb, dt, d.gmax, 0, w.Mc, (int)h->small_stride, h->skv, dyn);
    if (tm) (void)hipEventRecord(h->lev[2 * i + 1], st);
    if (!h->sym) {
        if (tm) (void)hipEventRecord(h->levr[2 * i], st);
        launch_lr_rowblk(st, d, w.tb, dt, d.gmax, 0, w.Ar, (int)h->small_stride, w.Bt, h->bt_stride, h->skv, dyn);
        if (tm) (void)hipEventRecord(h->levr[2 * i + 1], st);
    }
    (void)hipStreamWaitEvent(st, h->ev_join, 0);
    FlipUpdate fu;
    fu.P = h->d_P;
    fu.tile = h->d_tile;
    fu.t = dt;
    fu.gtiles = dgt;
    launch_lr_finish(st, d, w.tb, w.v, gtab(h), (const Z*)w.acc, h->acc_stride, (const Z*)w.Bt, h->bt_stride, ds,
                     d.gmax, 0, dyn, h->part ? h->d_left : nullptr, h->nleft, h->ngroups_drv, fu);
}

void device_loop_iteration(pngf_handle* h, int i) {
    if (h->c64) device_loop_iteration_t<zf>(h, i);
    else device_loop_iteration_t<zd>(h, i);
}

// Captures kLoopChunk iterations into h->loop_exec once (after device_loop_prepare); a capture
// the runtime refuses falls back to direct launches.  Relaxed capture mode: other threads
// driving other handles keep calling the runtime meanwhile.
void device_loop_capture(pngf_handle* h) {
    if (h->loop_exec || h->graph_failed || (h->prof && h->d_iterlog)) return;
    hipStream_t st = h->st;
    bool ok = hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed) == hipSuccess;
    if (ok) {
        for (int i = 0; i < kLoopChunk; i++) device_loop_iteration(h, i);
        ok = hipStreamEndCapture(st, &h->loop_graph) == hipSuccess && h->loop_graph;
        ok = ok && hipGraphInstantiate(&h->loop_exec, h->loop_graph, nullptr, nullptr, 0) == hipSuccess;
    }
    if (!ok) {
        (void)hipGetLastError();
        if (h->loop_exec) (void)hipGraphExecDestroy(h->loop_exec);
        if (h->loop_graph) (void)hipGraphDestroy(h->loop_graph);
        h->loop_exec = nullptr;
        h->loop_graph = nullptr;
        h->graph_failed = true;
    }
}

// Runs the loop to its end; fills the counters and the accepted-flip arrays.
int device_loop_run(pngf_handle* h, const pngf_dbs_params* p, int64_t* flip_count, int64_t* nacc, int64_t* spec,
                    int64_t* batches, double* F_best, int* converged) {
    hipStream_t st = h->st;
    const int gmax = h->d.gmax;
    LoopState& ls = *h->h_loop;
    int Kknown = h->K;
    while (true) {
        int S = (h->Kmax - Kknown) / gmax;
        if (S < 1) {
            // fold the factors (including a pending accept's columns) before the next chunk
            h->K = Kknown;
            CK(flush_factors(h));
            ls.dyn[kDynKK] = 0;
            ls.dyn[kDynActive] = 0;
            HC(h, hipMemcpyAsync(h->d_loop, &ls, sizeof(LoopState), hipMemcpyHostToDevice, st));
            Kknown = 0;
            S = h->Kmax / gmax;
        }
        S = std::min(S, kLoopChunk);
        // with profiling on, the iterations are launched directly: elapsed times of events that a
        // graph replay recorded are not available
        const bool tm = h->prof && h->d_iterlog;
        if (S == kLoopChunk && !tm) device_loop_capture(h);
        if (S == kLoopChunk && !tm && h->loop_exec)
            HC(h, hipGraphLaunch(h->loop_exec, st));
        else
            for (int i = 0; i < S; i++) device_loop_iteration(h, i);
        HC(h, hipGetLastError());
        if (tm)
            HC(h, hipMemcpyAsync(h->h_iterlog, h->d_iterlog, 3 * S * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        HC(h, hipMemcpyAsync(&ls, h->d_loop, sizeof(LoopState), hipMemcpyDeviceToHost, st));
        HC(h, hipStreamSynchronize(st));
        if (tm) {
            // event time of every column strip (and row strip) that ran (iterations with an accept)
            const Dims& d = h->d;
            for (int which = 0; which < (h->sym ? 1 : 2); which++) {
                auto& ks = h->kst[which == 0 ? PNGF_K_COLBLK : PNGF_K_ROWBLK];
                hipEvent_t* ev = which == 0 ? h->lev : h->levr;
                for (int i = 0; i < S; i++) {
                    const int32_t* r = h->h_iterlog + 3 * i;
                    float ms = 0.f;
                    if (!r[0] || hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]) != hipSuccess) continue;
                    const double N = d.N, kd = r[1], Kd = r[2];
                    ks.launches++;
                    ks.total_ms += ms;
                    ks.bytes += d.nf * 16.0 * (N * Kd + 2 * N * kd + Kd * kd);
                    ks.flops += d.nf * 8.0 * N * Kd * kd;
                }
            }
        }
        Kknown = ls.dyn[kDynKK] + (ls.dyn[kDynActive] ? ls.dyn[kDynK] : 0);
        if (ls.done) break;
    }
    h->K = Kknown;
    // host mirrors of the mask and tile map
    HC(h, hipMemcpy(h->tile.data(), h->d_tile, h->tile.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    HC(h, hipMemcpy(h->P.data(), h->d_P, h->P.size(), hipMemcpyDeviceToHost));
    const int na = ls.nacc;
    if (na > 0) {
        std::vector<int32_t> g(na);
        std::vector<int64_t> e(na);
        std::vector<double> F(na);
        HC(h, hipMemcpy(g.data(), h->d_accg, na * sizeof(int32_t), hipMemcpyDeviceToHost));
        HC(h, hipMemcpy(e.data(), h->d_acce, na * sizeof(int64_t), hipMemcpyDeviceToHost));
        HC(h, hipMemcpy(F.data(), h->d_accF, na * sizeof(double), hipMemcpyDeviceToHost));
        for (int i = 0; i < na; i++) {
            if (p->acc_tiles) p->acc_tiles[i] = h->grp.prim[g[i]];
            if (p->acc_eval) p->acc_eval[i] = e[i];
            if (p->acc_F) p->acc_F[i] = F[i];
        }
    }
    *flip_count = ls.flip_count;
    *nacc = na;
    *spec = ls.spec;
    *batches = ls.batches;
    *F_best = ls.F_best;
    *converged = ls.nonimp >= ls.ncl ? 1 : 0;
    if (ls.status == kLoopFull) return fail(h, PNGF_ERR_ARG, "accepted-flip arrays are full");
    if (ls.status == kLoopSingular) return fail(h, PNGF_ERR_SINGULAR, "capacitance matrix is singular");
    return PNGF_OK;
}

int pngf_dbs_run(pngf_handle* h, const pngf_dbs_params* p, pngf_dbs_result* r) {
    if (!h || !p || !r) return fail(h, PNGF_ERR_ARG, "null argument");
    CK(check_ready(h));
    std::memset(r, 0, sizeof(*r));
    if (p->log_x && h->cfg.method != PNGF_METHOD_LOWRANK)
        return fail(h, PNGF_ERR_STATE, "the training log needs the LOWRANK method");
    if (hist_on(p) && p->order == PNGF_ORDER_BEST)
        return fail(h, PNGF_ERR_ARG, "the optimization history follows first-improvement order");
    std::vector<zd> e;
    CK(download_eobj(h, e));
    double F_best = host_objective(h, e.data());
    r->F_start = F_best;
    if (hist_on(p)) hist_record(h, p, &r->hist_count, -1, -1, F_best, F_best, e.data());
    const bool plain = h->cfg.method == PNGF_METHOD_PLAIN;
    // speculative batch after an accept: PLAIN scores are costly (4), LOWRANK ones cheap (32)
    const int bmin = p->batch_min > 0 ? p->batch_min : (plain ? 4 : 32);
    const int bmax_cfg = p->batch_max > 0 ? p->batch_max : (1 << 30);
    int64_t flip_count = 0, nacc = 0, spec = 0, batches = 0;
    const double acc0 = h->acc_seconds;
    h->flushes = 0;
    if (p->resume && p->order != PNGF_ORDER_FIRST)
        return fail(h, PNGF_ERR_ARG, "resume follows first-improvement order");
    // candidate groups of the raster cycle, in visit order
    std::vector<int> cl;
    for (int g = 0; g < h->ngroups_drv; g++) if (h->grp.ntiles[g] > 0) cl.push_back(g);
    const int ncl = (int)cl.size();
    // where the visit order starts: its beginning, or a saved driver position
    int pos0 = 0, last_tried0 = -1;
    int64_t nonimp0 = 0;
    Hat hat;
    if (h->cfg.driver == PNGF_DRIVER_REF) {
        hat.init(p->seed, h->ngroups_drv);
        if (p->rng_discard > 0 && !p->resume) hat.gen.discard(p->rng_discard);
    }
    if (p->resume) CK(drv_load(h, p, ncl, &pos0, &nonimp0, &hat, &last_tried0));
    const bool dev_loop = device_loop_ok(h, p);
    if (dev_loop)
        CK(device_loop_prepare(h, p, F_best, bmin, (int)std::min<int64_t>(bmax_cfg, 1 << 30), pos0, (int)nonimp0));
    HC(h, hipStreamSynchronize(h->st));
    auto t0 = std::chrono::high_resolution_clock::now();
    auto record = [&](int g, double F) -> int {
        if (nacc >= p->max_accepts) return fail(h, PNGF_ERR_ARG, "accepted-flip arrays are full");
        if (p->acc_tiles) p->acc_tiles[nacc] = h->grp.prim[g];
        if (p->acc_eval) p->acc_eval[nacc] = flip_count;
        if (p->acc_F) p->acc_F[nacc] = F;
        return PNGF_OK;
    };
    int rc = PNGF_OK;
    std::vector<int> batch;
    int bs = bmin;
    if (h->cfg.driver == PNGF_DRIVER_RASTER && p->order == PNGF_ORDER_BEST) {
        // steepest descent: every step scores all candidates exactly and commits the best
        std::vector<double> F(ncl);
        while (true) {
            if (p->max_evals > 0 && flip_count + ncl > p->max_evals) break;
            h->h_pin_d[8] = F_best;
            HC(h, hipMemcpyAsync(h->d_Fbest, h->h_pin_d + 8, sizeof(double), hipMemcpyHostToDevice, h->st));
            CK(eval_launch(h, cl.data(), ncl));
            int sel; double Fs;
            CK(sync_sel(h, ncl, &sel, &Fs, true));
            batches++;
            CK(log_batch(h, p, ncl, &r->logged));
            HC(h, hipMemcpy(F.data(), h->d_F, sizeof(double) * ncl, hipMemcpyDeviceToHost));
            flip_count += ncl;
            int best = -1;
            for (int i = 0; i < ncl; i++)
                if (F[i] < F_best && (best < 0 || F[i] < F[best])) best = i;
            if (best < 0) { r->converged = 1; break; }
            flip_count -= 1;
            if ((rc = record(cl[best], F[best])) != PNGF_OK) break;
            flip_count += 1;
            CK(accept_group(h, cl[best]));
            nacc++;
            F_best = F[best];
        }
    } else if (dev_loop) {
        int conv = 0;
        rc = device_loop_run(h, p, &flip_count, &nacc, &spec, &batches, &F_best, &conv);
        r->converged = conv;
        if (rc == PNGF_OK) rc = drv_save(h, p, h->h_loop->pos, h->h_loop->nonimp, nullptr, -1);
    } else if (h->cfg.driver == PNGF_DRIVER_RASTER) {
        int pos = pos0;
        int64_t nonimp = nonimp0;
        while (nonimp < ncl) {
            if (p->max_evals > 0 && flip_count >= p->max_evals) break;
            int64_t B = ncl - nonimp;
            if (p->max_evals > 0) B = std::min<int64_t>(B, p->max_evals - flip_count);
            B = std::min<int64_t>(B, std::min(bs, bmax_cfg));
            batch.resize(B);
            for (int64_t i = 0; i < B; i++) batch[i] = cl[(pos + i) % ncl];
            h->h_pin_d[8] = F_best;
            HC(h, hipMemcpyAsync(h->d_Fbest, h->h_pin_d + 8, sizeof(double), hipMemcpyHostToDevice, h->st));
            CK(eval_launch(h, batch.data(), (int)B));
            int sel; double Fs;
            CK(sync_sel(h, (int)B, &sel, &Fs, true));
            batches++;
            CK(log_batch(h, p, (int)B, &r->logged));
            CK(hist_batch(h, p, batch.data(), sel >= 0 ? sel + 1 : (int)B, F_best, &r->hist_count));
            if (sel >= 0) {
                flip_count += sel;
                if ((rc = record(batch[sel], Fs)) != PNGF_OK) break;
                CK(accept_group(h, batch[sel]));
                flip_count += 1;
                nacc++;
                spec += B - sel - 1;
                F_best = Fs;
                pos = (pos + sel + 1) % ncl;
                nonimp = 0;
                bs = bmin;
            } else {
                flip_count += B;
                nonimp += B;
                pos = (int)((pos + B) % ncl);
                bs = std::min<int64_t>((int64_t)bs * 2, (int64_t)1 << 30);
            }
        }
        r->converged = nonimp >= ncl;
        if (rc == PNGF_OK) rc = drv_save(h, p, pos, nonimp, nullptr, -1);
    } else {
        int last_tried = last_tried0;
        bool done = false;
        while (!done) {
            if (p->max_evals > 0 && flip_count >= p->max_evals) break;
            // the draws that would follow if nothing were accepted (main.cpp:343-392)
            Hat sim = hat;
            int lt = last_tried;
            batch.clear();
            int64_t cap = p->max_evals > 0 ? p->max_evals - flip_count : (int64_t)1 << 40;
            cap = std::min<int64_t>(cap, std::min(bs, bmax_cfg));
            while ((int64_t)batch.size() < cap) {
                int rr = sim.draw();
                if (rr < 0) break;
                if (rr == lt) continue;
                if (h->grp.ntiles[rr] == 0) continue;
                lt = rr;
                batch.push_back(rr);
            }
            if (batch.empty()) { done = true; r->converged = 1; break; }
            h->h_pin_d[8] = F_best;
            HC(h, hipMemcpyAsync(h->d_Fbest, h->h_pin_d + 8, sizeof(double), hipMemcpyHostToDevice, h->st));
            CK(eval_launch(h, batch.data(), (int)batch.size()));
            int sel; double Fs;
            CK(sync_sel(h, (int)batch.size(), &sel, &Fs, true));
            batches++;
            CK(log_batch(h, p, (int)batch.size(), &r->logged));
            const int upto = sel >= 0 ? sel : (int)batch.size() - 1;
            CK(hist_batch(h, p, batch.data(), upto + 1, F_best, &r->hist_count));
            // replay the real hat through the evaluated draws
            int cnt = 0;
            while (true) {
                int rr = hat.draw();
                if (rr == last_tried) continue;
                if (h->grp.ntiles[rr] == 0) continue;
                last_tried = rr;
                if (cnt == upto) break;
                cnt++;
            }
            if (sel >= 0) {
                flip_count += sel;
                if ((rc = record(batch[sel], Fs)) != PNGF_OK) break;
                CK(accept_group(h, batch[sel]));
                flip_count += 1;
                nacc++;
                spec += (int64_t)batch.size() - sel - 1;
                F_best = Fs;
                hat.reset();
                bs = bmin;
            } else {
                flip_count += (int64_t)batch.size();
                bs = std::min<int64_t>((int64_t)bs * 2, (int64_t)1 << 30);
            }
        }
        if (rc == PNGF_OK) rc = drv_save(h, p, 0, 0, &hat, last_tried);
    }
    HC(h, hipStreamSynchronize(h->st));
    prof_collect(h);
    r->seconds = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
    r->flip_count = flip_count;
    r->accepts = nacc;
    r->speculative_evals = spec;
    r->batches = batches;
    r->F_final = F_best;
    r->accept_seconds = h->acc_seconds - acc0;
    r->flushes = h->flushes;
    return rc;
}

int pngf_ranker_dims(pngf_handle* h, int32_t* din, int32_t* dout) {
    if (!h || !din || !dout) return PNGF_ERR_ARG;
    *din = ranker_din(h);
    *dout = h->d.nf;
    return PNGF_OK;
}

int pngf_features(pngf_handle* h, const int32_t* tiles, int32_t ncand, float* out) {
    if (!h || !tiles || !out || ncand < 0) return fail(h, PNGF_ERR_ARG, "bad feature request");
    CK(check_ready(h));
    if (h->cfg.driver != PNGF_DRIVER_RASTER) return fail(h, PNGF_ERR_STATE, "ranker inputs need the raster driver");
    if (ncand > h->ngroups_drv) return fail(h, PNGF_ERR_ARG, "too many candidates");
    for (int i = 0; i < ncand; i++) {
        if (tiles[i] < 0 || tiles[i] >= h->ngroups_drv || h->grp.ntiles[tiles[i]] == 0)
            return fail(h, PNGF_ERR_ARG, "candidate is not a flippable tile");
        h->h_pin_i[kPinBatch + i] = tiles[i];
    }
    if (ncand == 0) return PNGF_OK;
    HC(h, hipMemcpyAsync(h->d_batch, h->h_pin_i + kPinBatch, ncand * sizeof(int32_t), hipMemcpyHostToDevice, h->st));
    return features_batch(h, ncand, out);
}

int pngf_dbs_run_ranked(pngf_handle* h, pngf_mlp* m, const pngf_rank_params* rp, const pngf_dbs_params* p,
                        pngf_dbs_result* r) {
    if (!h || !m || !rp || !p || !r || rp->top_k < 1) return fail(h, PNGF_ERR_ARG, "null argument");
    CK(check_ready(h));
    if (h->cfg.driver != PNGF_DRIVER_RASTER || h->cfg.objective != PNGF_OBJ_S11 ||
        h->cfg.method != PNGF_METHOD_LOWRANK)
        return fail(h, PNGF_ERR_STATE, "ranked DBS needs the LOWRANK method, raster driver and S11 objective");
    std::memset(r, 0, sizeof(*r));
    const int din = ranker_din(h), nf = h->d.nf;
    std::vector<zd> e;
    CK(download_eobj(h, e));
    double F_best = host_objective(h, e.data());
    r->F_start = F_best;
    std::vector<int> cl;
    for (int g = 0; g < h->ngroups_drv; g++) if (h->grp.ntiles[g] > 0) cl.push_back(g);
    const int ncl = (int)cl.size();
    const int K = std::min(rp->top_k, ncl);
    std::vector<float> X((size_t)ncl * din), Y((size_t)ncl * nf);
    std::vector<double> F(ncl), score(ncl);
    std::vector<int> order(ncl), hits;
    int64_t flip_count = 0, nacc = 0, scored = 0, sweeps = 0;
    const int window = std::max(1, rp->window);
    h->flushes = 0;
    HC(h, hipStreamSynchronize(h->st));
    auto t0 = std::chrono::high_resolution_clock::now();
    // exact scores of `groups` -> F (host); returns index of the best improving or -1
    auto exact = [&](const std::vector<int>& groups, double* Fout, int* best) -> int {
        h->h_pin_d[8] = F_best;
        HC(h, hipMemcpyAsync(h->d_Fbest, h->h_pin_d + 8, sizeof(double), hipMemcpyHostToDevice, h->st));
        CK(eval_launch(h, groups.data(), (int)groups.size()));
        int sel; double Fs;
        CK(sync_sel(h, (int)groups.size(), &sel, &Fs, true));
        HC(h, hipMemcpy(Fout, h->d_F, sizeof(double) * groups.size(), hipMemcpyDeviceToHost));
        CK(log_batch(h, p, (int)groups.size(), &r->logged));
        *best = -1;
        for (size_t i = 0; i < groups.size(); i++)
            if (Fout[i] < F_best && (*best < 0 || Fout[i] < Fout[*best])) *best = (int)i;
        flip_count += (int64_t)groups.size();
        return PNGF_OK;
    };
    auto commit = [&](int g, double Fnew) -> int {
        if (nacc >= p->max_accepts) return fail(h, PNGF_ERR_ARG, "accepted-flip arrays are full");
        if (p->acc_tiles) p->acc_tiles[nacc] = h->grp.prim[g];
        if (p->acc_eval) p->acc_eval[nacc] = flip_count - 1;
        if (p->acc_F) p->acc_F[nacc] = Fnew;
        CK(accept_group(h, g));
        nacc++;
        F_best = Fnew;
        return PNGF_OK;
    };
    while (true) {
        if (p->max_evals > 0 && flip_count >= p->max_evals) break;
        int hsum = 0;
        const int hn = std::min<int>((int)hits.size(), window);
        for (int i = 0; i < hn; i++) hsum += hits[hits.size() - 1 - i];
        const bool fallback = hn == window && hsum < rp->min_hit_rate * window;
        if (!fallback) {
            for (int i = 0; i < ncl; i++) h->h_pin_i[kPinBatch + i] = cl[i];
            HC(h, hipMemcpyAsync(h->d_batch, h->h_pin_i + kPinBatch, ncl * sizeof(int32_t), hipMemcpyHostToDevice, h->st));
            CK(features_batch(h, ncl, X.data()));
            for (int i = 0; i < ncl; i++)
                for (int c = 0; c < din; c++) {
                    float& x = X[(size_t)i * din + c];
                    x = (x - rp->x_mean[c]) / rp->x_std[c];
                }
            int rc = pngf_mlp_predict(m, X.data(), ncl, Y.data());
            if (rc != PNGF_OK) return fail(h, rc, std::string("ranker: ") + pngf_mlp_last_error(m));
            for (int i = 0; i < ncl; i++) {
                double s = 0;
                for (int f = 0; f < nf; f++) s += Y[(size_t)i * nf + f] * rp->y_std[f] + rp->y_mean[f];
                score[i] = s / nf;
                order[i] = i;
            }
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return score[a] < score[b]; });
            scored += ncl;
            // exact scores in rank order, top_k at a time, until a chunk holds an improving flip
            // (its best is accepted); a step that runs through every chunk is an exact sweep that
            // found none: a verified local minimum
            int best = -1;
            for (int pos = 0; pos < ncl && best < 0; pos += K) {
                std::vector<int> chunk;
                for (int i = pos; i < std::min(ncl, pos + K); i++) chunk.push_back(cl[order[i]]);
                CK(exact(chunk, F.data(), &best));
                if (pos == 0) {
                    hits.push_back(best >= 0 ? 1 : 0);
                    r->ranked_steps++;
                    r->ranked_hits += best >= 0 ? 1 : 0;
                }
                if (best >= 0) CK(commit(chunk[best], F[best]));
            }
            if (best < 0) { r->converged = 1; break; }
            continue;
        } else {
            hits.push_back(1);  // a fallback step re-arms the window
        }
        int best;
        CK(exact(cl, F.data(), &best));
        sweeps++;
        if (best < 0) { r->converged = 1; break; }
        CK(commit(cl[best], F[best]));
    }
    HC(h, hipStreamSynchronize(h->st));
    prof_collect(h);
    r->seconds = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
    r->flip_count = flip_count;
    r->accepts = nacc;
    r->F_final = F_best;
    r->ranked_scored = scored;
    r->full_sweeps = sweeps;
    r->flushes = h->flushes;
    return PNGF_OK;
}

int pngf_dbs_run_pool(pngf_handle** hs, int32_t n, const pngf_dbs_params* ps, pngf_dbs_result* rs,
                      int32_t workers) {
    if (!hs || !ps || !rs || n < 1 || workers < 1) return PNGF_ERR_ARG;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < i; j++)
            if (!hs[i] || hs[i] == hs[j]) return PNGF_ERR_ARG;  // one thread per handle
    // device-loop buffers and graphs are set up here, one handle after another: allocations,
    // synchronous copies and stream capture are then done before the threads run
    for (int i = 0; i < n; i++) {
        pngf_handle* h = hs[i];
        if (!h->inited || !device_loop_ok(h, &ps[i])) continue;
        (void)hipSetDevice(h->cfg.device);
        CK(device_loop_prepare(h, &ps[i], 0.0, 1, 1));
        HC(h, hipStreamSynchronize(h->st));
        device_loop_capture(h);
    }
    // largest problems first (N x frequencies), handed out to the workers in that order
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return (double)hs[a]->d.N * hs[a]->d.nf > (double)hs[b]->d.N * hs[b]->d.nf;
    });
    std::vector<int> rc(n, PNGF_OK);
    std::atomic<int> next{0};
    auto work = [&]() {
        while (true) {
            const int q = next.fetch_add(1);
            if (q >= n) break;
            const int i = order[q];
            (void)hipSetDevice(hs[i]->cfg.device);
            rc[i] = pngf_dbs_run(hs[i], &ps[i], &rs[i]);
        }
    };
    const int nw = std::min<int>(workers, n);
    std::vector<std::thread> th;
    th.reserve(nw);
    for (int w = 0; w < nw; w++) th.emplace_back(work);
    for (auto& t : th) t.join();
    for (int i = 0; i < n; i++)
        if (rc[i] != PNGF_OK) return rc[i];
    return PNGF_OK;
}

int pngf_dbs_run_many(pngf_handle** hs, int32_t n, const pngf_dbs_params* ps, pngf_dbs_result* rs) {
    return pngf_dbs_run_pool(hs, n, ps, rs, kDefaultWorkers);
}

int pngf_symmetry(pngf_handle* h, int32_t* used, double* rel_asym) {
    if (!h) return PNGF_ERR_ARG;
    if (used) *used = h->sym ? 1 : 0;
    if (rel_asym) *rel_asym = h->asym;
    return PNGF_OK;
}

int pngf_factor_capacity(pngf_handle* h, int32_t* columns, int32_t* rebuilt) {
    if (!h) return PNGF_ERR_ARG;
    if (columns) *columns = h->Kmax;
    if (rebuilt) *rebuilt = h->lam ? 1 : 0;
    return PNGF_OK;
}

int pngf_autotune_result(pngf_handle* h, int32_t* chosen, int32_t* nvariants, double* ms) {
    if (!h || !chosen || !nvariants) return PNGF_ERR_ARG;
    *chosen = h->skv;
    *nvariants = (int32_t)h->tune_ms.size();
    if (ms) for (size_t i = 0; i < h->tune_ms.size(); i++) ms[i] = h->tune_ms[i];
    return PNGF_OK;
}

}  // extern "C"
