// Learned flip ranker: a 3-layer MLP (din -> dh -> dh -> dout, GELU) trained and
// evaluated with this file's kernels on gfx950.  Weights are kept as FP32 master
// copies (Adam state beside them) and enter every product rounded to BF16; products
// run on BF16 MFMA (v_mfma_f32_16x16x16bf16_1k) with FP32 accumulation.
//
// The network predicts, for one candidate flip, the change of |S11_f|^2 at every
// frequency from cheap per-candidate features (tile neighbourhood, position and
// capacitance-diagonal / first-order port-voltage terms, see runtime.hip).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/pngf_gpu.h"

namespace {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ inline short bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// C[i][j] (+)= sum_k A(i,k) B(k,j), A(i,k) = A[i*sai + k*sak], B(k,j) = B[k*sbk + j*sbj],
// operands rounded to BF16, FP32 accumulation.  One wave per 16x16 tile of C: the ranker's
// products are small (training batches of 512 rows, widths up to 256: at most 34 M MACs each) and
// bound by latency, and a tile per wave puts the most waves in flight; the operands come from
// the L2 without LDS staging, each entry read by the 16 (or 4) tiles that share it.
// Fragment layouts (16x16x16 bf16): A[i][k] in lane i + 16(k/4), element k%4;
// B[k][j] in lane j + 16(k/4), element k%4; D[i][j] in lane j + 16(i/4), element i%4.
__global__ void __launch_bounds__(256) gemm_bf16_kernel(int M, int N, int K, const float* __restrict__ A, int sai,
                                                        int sak, const float* __restrict__ B, int sbk, int sbj,
                                                        float* C, int ldc, const float* __restrict__ bias,
                                                        int accumulate) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tiles_n = (N + 15) / 16;
    const int tile = blockIdx.x * 4 + wave;
    const int ti = tile / tiles_n, tj = tile - ti * tiles_n;
    if (ti * 16 >= M) return;
    const int r = lane & 15, g = lane >> 4;
    const int i = ti * 16 + r, j = tj * 16 + r;
    f4v acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        s4v a, b;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int k = k0 + 4 * g + e;
            const float av = (i < M && k < K) ? A[(size_t)i * sai + (size_t)k * sak] : 0.f;
            const float bv = (j < N && k < K) ? B[(size_t)k * sbk + (size_t)j * sbj] : 0.f;
            a[e] = bf16_rne(av);
            b[e] = bf16_rne(bv);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int ii = ti * 16 + 4 * g + v, jj = tj * 16 + r;
        if (ii < M && jj < N) {
            float x = acc[v] + (bias ? bias[jj] : 0.f);
            float* p = &C[(size_t)ii * ldc + jj];
            *p = accumulate ? *p + x : x;
        }
    }
}

__device__ inline float gelu_f(float x) {
    const float c = 0.7978845608028654f;  // sqrt(2/pi)
    return 0.5f * x * (1.f + tanhf(c * (x + 0.044715f * x * x * x)));
}
__device__ inline float gelu_grad(float x) {
    const float c = 0.7978845608028654f;
    const float u = c * (x + 0.044715f * x * x * x);
    const float t = tanhf(u);
    return 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * c * (1.f + 3.f * 0.044715f * x * x);
}

__global__ void gelu_kernel(int n, const float* __restrict__ z, float* h) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) h[i] = gelu_f(z[i]);
}
// dz = dh * gelu'(z)
__global__ void gelu_back_kernel(int n, const float* __restrict__ z, const float* __restrict__ dh, float* dz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dz[i] = dh[i] * gelu_grad(z[i]);
}
// dy = 2 (y - t) / (rows * cols); per-block partial loss sums (fixed order)
__global__ void mse_kernel(int n, int denom, const float* __restrict__ y, const float* __restrict__ t, float* dy,
                           float* partial) {
    __shared__ float red[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < n) {
        const float d = y[i] - t[i];
        dy[i] = 2.f * d / denom;
        l = d * d;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}
// db[j] = sum_i dz[i][j]
__global__ void colsum_kernel(int rows, int cols, const float* __restrict__ dz, float* db) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    float s = 0.f;
    for (int i = 0; i < rows; i++) s += dz[(size_t)i * cols + j];
    db[j] = s;
}
// rows r < n of dst = src rows idx[r]
__global__ void gather_rows_kernel(int n, int w, const int* __restrict__ idx, const float* __restrict__ src, float* dst) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * w) return;
    const int r = e / w, c = e - r * w;
    dst[e] = src[(size_t)idx[r] * w + c];
}
__global__ void adam_kernel(int n, float* w, const float* __restrict__ g, float* m, float* v, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace

// Scratch of pngf_mlp_predict_dev (the entry point that allocates nothing): the two n x dh
// activation arrays of one forward pass, at most kNoAllocFloats floats (4 MiB) in all.  Each GPU
// holds its own copy for the life of the process; host code reaches it only through the address
// hipGetSymbolAddress returns on that device (the name itself is a host-side stand-in).
constexpr size_t kNoAllocFloats = (size_t)1 << 20;
static __device__ __align__(256) float scratch[kNoAllocFloats];

struct pngf_mlp {
    int din = 0, dh = 0, dout = 0, cap = 0;
    int device = 0;
    float* scratch_dev = nullptr;  // this device's `scratch`, from hipGetSymbolAddress
    int nparam = 0;
    float* w = nullptr;   // all parameters: W1 b1 W2 b2 W3 b3 (row-major W[in][out])
    float* g = nullptr;   // gradients, same layout
    float* m = nullptr;
    float* v = nullptr;
    float *X = nullptr, *Y = nullptr, *Z1 = nullptr, *H1 = nullptr, *Z2 = nullptr, *H2 = nullptr, *O = nullptr,
          *dO = nullptr, *dH = nullptr, *dZ = nullptr, *loss = nullptr;
    int step = 0;
    hipStream_t st = nullptr;
    std::string err;
    size_t off_W1, off_b1, off_W2, off_b2, off_W3, off_b3;
};

namespace {

int mfail(pngf_mlp* p, int code, const std::string& m) {
    if (p) p->err = m;
    return code;
}
#define MHC(p, x)                                                                                         \
    do {                                                                                                  \
        hipError_t e_ = (x);                                                                              \
        if (e_ != hipSuccess) return mfail(p, PNGF_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

void gemm_on(hipStream_t st, int M, int N, int K, const float* A, int sai, int sak, const float* B, int sbk,
             int sbj, float* C, int ldc, const float* bias, int acc) {
    const int tiles = ((M + 15) / 16) * ((N + 15) / 16);
    hipLaunchKernelGGL(gemm_bf16_kernel, dim3((tiles + 3) / 4), dim3(256), 0, st, M, N, K, A, sai, sak, B, sbk, sbj,
                       C, ldc, bias, acc);
}

void gemm(pngf_mlp* p, int M, int N, int K, const float* A, int sai, int sak, const float* B, int sbk, int sbj,
          float* C, int ldc, const float* bias, int acc) {
    gemm_on(p->st, M, N, K, A, sai, sak, B, sbk, sbj, C, ldc, bias, acc);
}

int grow(pngf_mlp* p, int n) {
    if (n <= p->cap) return PNGF_OK;
    float** bufs[] = {&p->X, &p->Y, &p->Z1, &p->H1, &p->Z2, &p->H2, &p->O, &p->dO, &p->dH, &p->dZ, &p->loss};
    for (auto b : bufs) { if (*b) (void)hipFree(*b); *b = nullptr; }
    const size_t wid = (size_t)std::max(std::max(p->din, p->dh), p->dout);
    MHC(p, hipMalloc(&p->X, sizeof(float) * n * p->din));
    MHC(p, hipMalloc(&p->Y, sizeof(float) * n * p->dout));
    MHC(p, hipMalloc(&p->Z1, sizeof(float) * n * p->dh));
    MHC(p, hipMalloc(&p->H1, sizeof(float) * n * p->dh));
    MHC(p, hipMalloc(&p->Z2, sizeof(float) * n * p->dh));
    MHC(p, hipMalloc(&p->H2, sizeof(float) * n * p->dh));
    MHC(p, hipMalloc(&p->O, sizeof(float) * n * p->dout));
    MHC(p, hipMalloc(&p->dO, sizeof(float) * n * p->dout));
    MHC(p, hipMalloc(&p->dH, sizeof(float) * n * wid));
    MHC(p, hipMalloc(&p->dZ, sizeof(float) * n * wid));
    MHC(p, hipMalloc(&p->loss, sizeof(float) * ((size_t)n * p->dout / 256 + 2)));
    p->cap = n;
    return PNGF_OK;
}

// Forward pass of n rows X -> O on stream st through the given activation arrays (every one is
// written before it is read: the products overwrite their outputs).
void forward_into(pngf_mlp* p, int n, const float* X, float* Z1, float* H1, float* Z2, float* H2, float* O,
                  hipStream_t st) {
    const float* w = p->w;
    const int e = 256;
    gemm_on(st, n, p->dh, p->din, X, p->din, 1, w + p->off_W1, p->dh, 1, Z1, p->dh, w + p->off_b1, 0);
    hipLaunchKernelGGL(gelu_kernel, dim3((n * p->dh + e - 1) / e), dim3(e), 0, st, n * p->dh, Z1, H1);
    gemm_on(st, n, p->dh, p->dh, H1, p->dh, 1, w + p->off_W2, p->dh, 1, Z2, p->dh, w + p->off_b2, 0);
    hipLaunchKernelGGL(gelu_kernel, dim3((n * p->dh + e - 1) / e), dim3(e), 0, st, n * p->dh, Z2, H2);
    gemm_on(st, n, p->dout, p->dh, H2, p->dh, 1, w + p->off_W3, p->dout, 1, O, p->dout, w + p->off_b3, 0);
}

void forward(pngf_mlp* p, int n) { forward_into(p, n, p->X, p->Z1, p->H1, p->Z2, p->H2, p->O, p->st); }

// This device's address of `scratch`, looked up once per device (with that device current).
int scratch_of_device(pngf_mlp* p, int dev) {
    static std::mutex mu;
    static float* addr[64] = {};
    if (dev < 0 || dev >= 64) return mfail(p, PNGF_ERR_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* a = nullptr;
        MHC(p, hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)));
        addr[dev] = (float*)a;
    }
    p->scratch_dev = addr[dev];
    return PNGF_OK;
}

// Gradients of the mean squared error over the n rows into p->g.
void backward(pngf_mlp* p, int n) {
    float* w = p->w;
    float* g = p->g;
    const int e = 256;
    const int no = n * p->dout;
    hipLaunchKernelGGL(mse_kernel, dim3((no + e - 1) / e), dim3(e), 0, p->st, no, no, p->O, p->Y, p->dO, p->loss);
    // layer 3: dW3 = H2^T dO, db3 = colsum dO, dH2 = dO W3^T
    gemm(p, p->dh, p->dout, n, p->H2, 1, p->dh, p->dO, p->dout, 1, g + p->off_W3, p->dout, nullptr, 0);
    hipLaunchKernelGGL(colsum_kernel, dim3((p->dout + 63) / 64), dim3(64), 0, p->st, n, p->dout, p->dO, g + p->off_b3);
    gemm(p, n, p->dh, p->dout, p->dO, p->dout, 1, w + p->off_W3, 1, p->dout, p->dH, p->dh, nullptr, 0);
    hipLaunchKernelGGL(gelu_back_kernel, dim3((n * p->dh + e - 1) / e), dim3(e), 0, p->st, n * p->dh, p->Z2, p->dH, p->dZ);
    // layer 2
    gemm(p, p->dh, p->dh, n, p->H1, 1, p->dh, p->dZ, p->dh, 1, g + p->off_W2, p->dh, nullptr, 0);
    hipLaunchKernelGGL(colsum_kernel, dim3((p->dh + 63) / 64), dim3(64), 0, p->st, n, p->dh, p->dZ, g + p->off_b2);
    gemm(p, n, p->dh, p->dh, p->dZ, p->dh, 1, w + p->off_W2, 1, p->dh, p->dH, p->dh, nullptr, 0);
    hipLaunchKernelGGL(gelu_back_kernel, dim3((n * p->dh + e - 1) / e), dim3(e), 0, p->st, n * p->dh, p->Z1, p->dH, p->dZ);
    // layer 1
    gemm(p, p->din, p->dh, n, p->X, 1, p->din, p->dZ, p->dh, 1, g + p->off_W1, p->dh, nullptr, 0);
    hipLaunchKernelGGL(colsum_kernel, dim3((p->dh + 63) / 64), dim3(64), 0, p->st, n, p->dh, p->dZ, g + p->off_b1);
}

}  // namespace

extern "C" {

const char* pngf_mlp_last_error(const pngf_mlp* p) { return p ? p->err.c_str() : "null mlp"; }

int pngf_mlp_create(int32_t din, int32_t dh, int32_t dout, int32_t seed, pngf_mlp** out) {
    if (!out || din < 1 || dh < 1 || dout < 1) return PNGF_ERR_ARG;
    pngf_mlp* p = new pngf_mlp();
    *out = p;
    p->din = din; p->dh = dh; p->dout = dout;
    size_t o = 0;
    p->off_W1 = o; o += (size_t)din * dh;
    p->off_b1 = o; o += dh;
    p->off_W2 = o; o += (size_t)dh * dh;
    p->off_b2 = o; o += dh;
    p->off_W3 = o; o += (size_t)dh * dout;
    p->off_b3 = o; o += dout;
    p->nparam = (int)o;
    MHC(p, hipGetDevice(&p->device));
    int rc = scratch_of_device(p, p->device);
    if (rc) return rc;
    MHC(p, hipStreamCreateWithFlags(&p->st, hipStreamNonBlocking));
    MHC(p, hipMalloc(&p->w, sizeof(float) * o));
    MHC(p, hipMalloc(&p->g, sizeof(float) * o));
    MHC(p, hipMalloc(&p->m, sizeof(float) * o));
    MHC(p, hipMalloc(&p->v, sizeof(float) * o));
    // He-style uniform init from std::mt19937(seed); biases zero
    std::vector<float> hw(o, 0.f);
    std::mt19937 gen(seed);
    std::uniform_real_distribution<float> U(-1.f, 1.f);
    auto fill = [&](size_t off, int fan_in, size_t cnt) {
        const float s = std::sqrt(3.f / fan_in);
        for (size_t i = 0; i < cnt; i++) hw[off + i] = s * U(gen);
    };
    fill(p->off_W1, din, (size_t)din * dh);
    fill(p->off_W2, dh, (size_t)dh * dh);
    fill(p->off_W3, dh, (size_t)dh * dout);
    MHC(p, hipMemcpy(p->w, hw.data(), sizeof(float) * o, hipMemcpyHostToDevice));
    MHC(p, hipMemset(p->m, 0, sizeof(float) * o));
    MHC(p, hipMemset(p->v, 0, sizeof(float) * o));
    return PNGF_OK;
}

int pngf_mlp_destroy(pngf_mlp* p) {
    if (!p) return PNGF_ERR_ARG;
    if (p->st) (void)hipStreamSynchronize(p->st);
    float* bufs[] = {p->w, p->g, p->m, p->v, p->X, p->Y, p->Z1, p->H1, p->Z2, p->H2, p->O, p->dO, p->dH, p->dZ, p->loss};
    for (auto b : bufs) if (b) (void)hipFree(b);
    if (p->st) (void)hipStreamDestroy(p->st);
    delete p;
    return PNGF_OK;
}

int pngf_mlp_num_params(const pngf_mlp* p) { return p ? p->nparam : PNGF_ERR_ARG; }

int pngf_mlp_get_params(pngf_mlp* p, float* out) {
    if (!p || !out) return PNGF_ERR_ARG;
    MHC(p, hipMemcpy(out, p->w, sizeof(float) * p->nparam, hipMemcpyDeviceToHost));
    return PNGF_OK;
}

int pngf_mlp_set_params(pngf_mlp* p, const float* in) {
    if (!p || !in) return PNGF_ERR_ARG;
    MHC(p, hipMemcpy(p->w, in, sizeof(float) * p->nparam, hipMemcpyHostToDevice));
    return PNGF_OK;
}

int pngf_mlp_predict(pngf_mlp* p, const float* X, int32_t n, float* out) {
    if (!p || !X || !out || n < 0) return PNGF_ERR_ARG;
    if (n == 0) return PNGF_OK;
    int rc = grow(p, n);
    if (rc) return rc;
    MHC(p, hipMemcpyAsync(p->X, X, sizeof(float) * n * p->din, hipMemcpyHostToDevice, p->st));
    forward(p, n);
    MHC(p, hipGetLastError());
    MHC(p, hipMemcpyAsync(out, p->O, sizeof(float) * n * p->dout, hipMemcpyDeviceToHost, p->st));
    MHC(p, hipStreamSynchronize(p->st));
    return PNGF_OK;
}

// Forward pass of n rows of device-resident inputs into device-resident outputs, enqueued on
// `stream` (null: the network's own stream) without waiting and without allocating: the
// activations live in this device's `scratch` (2 n dh floats, at most kNoAllocFloats).  Calls
// that overlap on one GPU share that scratch and race; nothing in it needs clearing between calls.
int pngf_mlp_predict_dev(pngf_mlp* p, const float* X_dev, int32_t n, float* out_dev, void* stream) {
    if (!p || n < 0 || (n > 0 && (!X_dev || !out_dev))) return mfail(p, PNGF_ERR_ARG, "bad predict_dev arguments");
    if (n == 0) return PNGF_OK;
    const size_t need = 2 * (size_t)n * p->dh;
    if (need > kNoAllocFloats)
        return mfail(p, PNGF_ERR_ARG, "predict_dev: 2 n dh = " + std::to_string(need) + " floats exceed the " +
                                          std::to_string(kNoAllocFloats) + " of its scratch");
    int dev = -1;
    MHC(p, hipGetDevice(&dev));
    if (dev != p->device) return mfail(p, PNGF_ERR_ARG, "predict_dev: the network lives on another device");
    hipStream_t st = stream ? (hipStream_t)stream : p->st;
    float* A = p->scratch_dev;
    float* B = p->scratch_dev + (size_t)n * p->dh;
    forward_into(p, n, X_dev, A, B, A, B, out_dev, st);
    MHC(p, hipGetLastError());
    return PNGF_OK;
}

// Gradient of the mean squared error for one batch (no update); grads: [nparam].
int pngf_mlp_grad(pngf_mlp* p, const float* X, const float* Y, int32_t n, float* grads, float* loss) {
    if (!p || !X || !Y || n < 1) return PNGF_ERR_ARG;
    int rc = grow(p, n);
    if (rc) return rc;
    MHC(p, hipMemcpyAsync(p->X, X, sizeof(float) * n * p->din, hipMemcpyHostToDevice, p->st));
    MHC(p, hipMemcpyAsync(p->Y, Y, sizeof(float) * n * p->dout, hipMemcpyHostToDevice, p->st));
    forward(p, n);
    backward(p, n);
    MHC(p, hipGetLastError());
    const int nb = (n * p->dout + 255) / 256;
    std::vector<float> part(nb);
    MHC(p, hipMemcpyAsync(part.data(), p->loss, sizeof(float) * nb, hipMemcpyDeviceToHost, p->st));
    if (grads) MHC(p, hipMemcpyAsync(grads, p->g, sizeof(float) * p->nparam, hipMemcpyDeviceToHost, p->st));
    MHC(p, hipStreamSynchronize(p->st));
    if (loss) { double s = 0; for (float x : part) s += x; *loss = (float)(s / ((double)n * p->dout)); }
    return PNGF_OK;
}

// Adam on mini-batches of the (n x din, n x dout) training set for `epochs` passes,
// rows shuffled every epoch with std::mt19937(seed).  loss_out: mean loss of the
// last epoch.
int pngf_mlp_train(pngf_mlp* p, const float* X, const float* Y, int32_t n, int32_t epochs, int32_t batch, float lr,
                   int32_t seed, float* loss_out) {
    if (!p || !X || !Y || n < 1 || epochs < 1 || batch < 1) return PNGF_ERR_ARG;
    batch = std::min(batch, n);
    int rc = grow(p, batch);
    if (rc) return rc;
    float *dX = nullptr, *dY = nullptr;
    int* dIdx = nullptr;
    MHC(p, hipMalloc(&dIdx, sizeof(int) * (size_t)n));
    MHC(p, hipMalloc(&dX, sizeof(float) * (size_t)n * p->din));
    MHC(p, hipMalloc(&dY, sizeof(float) * (size_t)n * p->dout));
    MHC(p, hipMemcpy(dX, X, sizeof(float) * (size_t)n * p->din, hipMemcpyHostToDevice));
    MHC(p, hipMemcpy(dY, Y, sizeof(float) * (size_t)n * p->dout, hipMemcpyHostToDevice));
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::mt19937 gen(seed);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    double last = 0.0;
    std::vector<float> part;
    for (int ep = 0; ep < epochs; ep++) {
        std::shuffle(order.begin(), order.end(), gen);
        MHC(p, hipMemcpyAsync(dIdx, order.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, p->st));
        double sum = 0.0;
        int cnt = 0;
        for (int s = 0; s + batch <= n; s += batch) {
            hipLaunchKernelGGL(gather_rows_kernel, dim3((batch * p->din + 255) / 256), dim3(256), 0, p->st, batch,
                               p->din, dIdx + s, dX, p->X);
            hipLaunchKernelGGL(gather_rows_kernel, dim3((batch * p->dout + 255) / 256), dim3(256), 0, p->st, batch,
                               p->dout, dIdx + s, dY, p->Y);
            forward(p, batch);
            backward(p, batch);
            p->step++;
            const float c1 = 1.f - std::pow(b1, (float)p->step), c2 = 1.f - std::pow(b2, (float)p->step);
            hipLaunchKernelGGL(adam_kernel, dim3((p->nparam + 255) / 256), dim3(256), 0, p->st, p->nparam, p->w, p->g,
                               p->m, p->v, lr, b1, b2, eps, c1, c2);
            const int nb = (batch * p->dout + 255) / 256;
            part.resize(nb);
            MHC(p, hipMemcpyAsync(part.data(), p->loss, sizeof(float) * nb, hipMemcpyDeviceToHost, p->st));
            MHC(p, hipStreamSynchronize(p->st));
            double l = 0;
            for (float x : part) l += x;
            sum += l / ((double)batch * p->dout);
            cnt++;
        }
        last = cnt ? sum / cnt : 0.0;
    }
    MHC(p, hipGetLastError());
    (void)hipFree(dX);
    (void)hipFree(dY);
    (void)hipFree(dIdx);
    if (loss_out) *loss_out = (float)last;
    return PNGF_OK;
}

}  // extern "C"
