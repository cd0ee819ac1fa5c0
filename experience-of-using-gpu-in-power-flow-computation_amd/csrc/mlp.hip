// Learned warm start: an MLP (bus P, Q injections and PV setpoints -> |V|,
// theta) trained on the solver's own NR solutions, with hand-written
// training and inference kernels.  BF16 weights and activations, FP32
// accumulation on the matrix cores (v_mfma_f32_16x16x32_bf16), FP32 master
// weights updated by AdamW.  The NR solve that follows judges the
// prediction; flat start remains the fallback (batch.hip).
//
// No counterpart in the reference: it always starts from v_inicial /
// theta_inicial (CPU/estrutura/barra.cpp:20-25).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "pf_internal.hpp"

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace pf {

enum { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_GELU = 2, EPI_DGELU = 3 };

constexpr int BM = 64, BN = 64, BK = 32, LDS_K = BK + 8;   // +8: rows land on different banks

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

// C[M x N] = op(A)[M x K] * op(B)[K x N], bf16 operands, fp32 accumulate.
//   TA = false: A row-major [M x K] (lda);  TA = true: A stored [K x M].
//   TB = false: B row-major [K x N] (ldb);  TB = true: B stored [N x K].
// Every dimension is a multiple of 8 (checked on the host), so 16-byte
// loads along the contiguous dimension never straddle the edge.
// 256 threads: 4 waves, each a 32x32 quarter of the 64x64 tile as 2x2
// MFMA 16x16x32 blocks.
template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, const bf16 *__restrict__ A, int lda,
                                              const bf16 *__restrict__ B, int ldb, float *__restrict__ C,
                                              int ldc, bf16 *__restrict__ Cb, int ldcb,
                                              const float *__restrict__ bias, const float *__restrict__ Zs) {
    __shared__ __align__(16) bf16 As[BM][LDS_K];
    __shared__ __align__(16) bf16 Bs[BN][LDS_K];
    const int t = threadIdx.x;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int w = t >> 6, l = t & 63;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    f32x4 acc[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const bf16x8 zero8 = {};
    for (int k0 = 0; k0 < K; k0 += BK) {
        // ---- stage A tile as As[m][k]
        if (!TA) {
            const int m = t >> 2, kk = (t & 3) * 8;
            bf16x8 v = zero8;
            if (m0 + m < M && k0 + kk < K) v = *reinterpret_cast<const bf16x8 *>(A + (size_t)(m0 + m) * lda + k0 + kk);
            *reinterpret_cast<bf16x8 *>(&As[m][kk]) = v;
        } else {
            const int kk = t >> 3, mm = (t & 7) * 8;
            bf16x8 v = zero8;
            if (k0 + kk < K && m0 + mm < M) v = *reinterpret_cast<const bf16x8 *>(A + (size_t)(k0 + kk) * lda + m0 + mm);
#pragma unroll
            for (int i = 0; i < 8; i++) As[mm + i][kk] = v[i];
        }
        // ---- stage B tile as Bs[n][k]
        if (TB) {
            const int n = t >> 2, kk = (t & 3) * 8;
            bf16x8 v = zero8;
            if (n0 + n < N && k0 + kk < K) v = *reinterpret_cast<const bf16x8 *>(B + (size_t)(n0 + n) * ldb + k0 + kk);
            *reinterpret_cast<bf16x8 *>(&Bs[n][kk]) = v;
        } else {
            const int kk = t >> 3, nn = (t & 7) * 8;
            bf16x8 v = zero8;
            if (k0 + kk < K && n0 + nn < N) v = *reinterpret_cast<const bf16x8 *>(B + (size_t)(k0 + kk) * ldb + n0 + nn);
#pragma unroll
            for (int i = 0; i < 8; i++) Bs[nn + i][kk] = v[i];
        }
        __syncthreads();
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; i++) a[i] = *reinterpret_cast<const bf16x8 *>(&As[wm + 16 * i + (l & 15)][8 * (l >> 4)]);
#pragma unroll
        for (int j = 0; j < 2; j++) b[j] = *reinterpret_cast<const bf16x8 *>(&Bs[wn + 16 * j + (l & 15)][8 * (l >> 4)]);
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    // ---- epilogue: lane l owns rows 4*(l/16)+v, column l%16 of each 16x16 block
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = m0 + wm + 16 * i + 4 * (l >> 4) + v;
                const int col = n0 + wn + 16 * j + (l & 15);
                if (row >= M || col >= N) continue;
                float z = acc[i][j][v];
                const size_t o = (size_t)row * ldc + col;
                if (EPI == EPI_STORE) {
                    C[o] = z;
                } else if (EPI == EPI_BIAS) {
                    C[o] = z + bias[col];
                } else if (EPI == EPI_BIAS_GELU) {
                    z += bias[col];
                    C[o] = z;
                    Cb[(size_t)row * ldcb + col] = (bf16)gelu(z);
                } else {   // EPI_DGELU: z is dH, Zs the saved pre-activation
                    const float d = z * gelu_grad(Zs[o]);
                    C[o] = d;
                    Cb[(size_t)row * ldcb + col] = (bf16)d;
                }
            }
}

__global__ void k_norm_in(const float *X, int B, int n, const float *mean, const float *inv_std, bf16 *Xn) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)B * n) return;
    const int c = (int)(i % n);
    Xn[i] = (bf16)((X[i] - mean[c]) * inv_std[c]);
}

// dY = 2 (Y - (T - mean)/std) / (B n); loss partial sums per block
__global__ void k_out_grad(const float *Y, const float *T, int B, int n, const float *mean, const float *inv_std,
                           float *dY, bf16 *dYb, float *loss_part) {
    __shared__ float red[256];
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    float e2 = 0.f;
    if (i < (size_t)B * n) {
        const int c = (int)(i % n);
        const float d = Y[i] - (T[i] - mean[c]) * inv_std[c];
        const float g = 2.f * d / ((float)B * (float)n);
        dY[i] = g;
        dYb[i] = (bf16)g;
        e2 = d * d;
    }
    red[threadIdx.x] = e2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// column sums (bias gradient): db[c] = sum_r dZ[r, c]
__global__ void k_colsum(const float *dZ, int B, int n, float *db) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    float s = 0.f;
    for (int r = 0; r < B; r++) s += dZ[(size_t)r * n + c];
    db[c] = s;
}

__global__ void k_adamw(float *p, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2,
                        float eps, float wd, float bc1, float bc2) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / bc1, vh = vi / bc2;
    p[i] -= lr * (mh / (sqrtf(vh) + eps) + wd * p[i]);
}

__global__ void k_to_bf16(const float *src, bf16 *dst, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (bf16)src[i];
}

__global__ void k_gather_rows(const float *src, const int *idx, int B, int n, float *dst) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)B * n) return;
    const size_t r = i / n, c = i % n;
    dst[i] = src[(size_t)idx[r] * n + c];
}

__global__ void k_denorm(const float *Y, int B, int n, const float *mean, const float *std, float *out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)B * n) return;
    const int c = (int)(i % n);
    out[i] = Y[i] * std[c] + mean[c];
}

}  // namespace pf

// ======================================================================= host
using namespace pf;

namespace {
#define MHIP(expr)                                                                 \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) {                                                    \
            pf::set_last_error(std::string(#expr) + ": " + hipGetErrorString(e_));       \
            return PF_DEVICE_ERROR;                                                \
        }                                                                          \
    } while (0)
#define MCHK(expr)                  \
    do {                            \
        pf_status s_ = (expr);      \
        if (s_ != PF_OK) return s_; \
    } while (0)

// SplitMix64 for deterministic weight initialisation
inline unsigned long long sm64(unsigned long long &s) {
    s += 0x9E3779B97F4A7C15ull;
    unsigned long long z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
}  // namespace

struct pf_warmstart {
    int device = 0;
    int n_in = 0, n_out = 0, hidden = 0, n_layers = 0;   // n_layers = hidden layers + 1
    std::vector<int> K, N;                               // per layer
    std::vector<size_t> w_off, b_off;                    // into the parameter vector
    size_t n_params = 0;
    int step = 0;
    int Bcap = 0;
    int staged = 0;   // rows of features on the device for pf_warmstart_predict_into
    hipStream_t stream = nullptr;
    std::vector<void *> allocs, ws_allocs;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    std::vector<bf16 *> Wb;
    float *xm = nullptr, *xis = nullptr, *ym = nullptr, *yis = nullptr, *ys = nullptr;
    // workspace
    float *X = nullptr, *T = nullptr, *Yp = nullptr, *Yo = nullptr, *loss_part = nullptr;
    bf16 *Xn = nullptr;
    std::vector<float *> Z;    // per layer pre-activation / dZ fp32
    std::vector<bf16 *> H;     // per hidden layer activation
    float *dZf[2] = {nullptr, nullptr};
    bf16 *dZb[2] = {nullptr, nullptr};

    template <typename T_>
    pf_status alloc(T_ **q, size_t n, std::vector<void *> &lst) {
        void *r = nullptr;
        MHIP(hipMalloc(&r, sizeof(T_) * std::max<size_t>(n, 1)));
        lst.push_back(r);
        *q = (T_ *)r;
        return PF_OK;
    }
    ~pf_warmstart() {
        for (void *q : ws_allocs) (void)hipFree(q);
        for (void *q : allocs) (void)hipFree(q);
        if (stream) (void)hipStreamDestroy(stream);
    }
    pf_status ensure(int B);
    pf_status sync_bf16();
    template <bool TA, bool TB, int EPI>
    pf_status gemm(int M_, int N_, int K_, const bf16 *A, int lda, const bf16 *B, int ldb, float *C, int ldc,
                   bf16 *Cb, int ldcb, const float *bias, const float *Zs);
    pf_status forward(int B);
    pf_status backward(int B, double *loss);
};

pf_status pf_warmstart::ensure(int B) {
    if (B <= Bcap) return PF_OK;
    for (void *q : ws_allocs) MHIP(hipFree(q));
    ws_allocs.clear();
    staged = 0;   // the staged rows went with the old workspace
    Bcap = std::max(B, 64);
    const size_t b = Bcap;
    int maxN = std::max(n_out, hidden);
    MCHK(alloc(&X, b * n_in, ws_allocs));
    MCHK(alloc(&T, b * n_out, ws_allocs));
    MCHK(alloc(&Yp, b * n_out, ws_allocs));
    MCHK(alloc(&Yo, b * n_out, ws_allocs));
    MCHK(alloc(&loss_part, (b * n_out + 255) / 256, ws_allocs));
    MCHK(alloc(&Xn, b * n_in, ws_allocs));
    Z.assign(n_layers, nullptr);
    H.assign(n_layers, nullptr);
    for (int l = 0; l < n_layers; l++) {
        MCHK(alloc(&Z[l], b * N[l], ws_allocs));
        if (l < n_layers - 1) MCHK(alloc(&H[l], b * N[l], ws_allocs));
    }
    for (int i = 0; i < 2; i++) {
        MCHK(alloc(&dZf[i], b * maxN, ws_allocs));
        MCHK(alloc(&dZb[i], b * maxN, ws_allocs));
    }
    return PF_OK;
}

pf_status pf_warmstart::sync_bf16() {
    for (int l = 0; l < n_layers; l++) {
        size_t n = (size_t)K[l] * N[l];
        k_to_bf16<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(p + w_off[l], Wb[l], n);
    }
    MHIP(hipGetLastError());
    return PF_OK;
}

template <bool TA, bool TB, int EPI>
pf_status pf_warmstart::gemm(int M_, int N_, int K_, const bf16 *A, int lda, const bf16 *B, int ldb, float *C,
                             int ldc, bf16 *Cb, int ldcb, const float *bias, const float *Zs) {
    if (M_ % 8 || N_ % 8 || K_ % 8) return PF_INVALID_ARG;
    dim3 grid((N_ + BN - 1) / BN, (M_ + BM - 1) / BM);
    k_gemm<TA, TB, EPI><<<grid, 256, 0, stream>>>(M_, N_, K_, A, lda, B, ldb, C, ldc, Cb, ldcb, bias, Zs);
    MHIP(hipGetLastError());
    return PF_OK;
}

pf_status pf_warmstart::forward(int B) {
    size_t nx = (size_t)B * n_in;
    k_norm_in<<<(unsigned)((nx + 255) / 256), 256, 0, stream>>>(X, B, n_in, xm, xis, Xn);
    MHIP(hipGetLastError());
    const bf16 *in = Xn;
    for (int l = 0; l < n_layers; l++) {
        const float *bias = p + b_off[l];
        if (l < n_layers - 1) {
            MCHK((gemm<false, false, EPI_BIAS_GELU>(B, N[l], K[l], in, K[l], Wb[l], N[l], Z[l], N[l], H[l], N[l],
                                                    bias, nullptr)));
            in = H[l];
        } else {
            MCHK((gemm<false, false, EPI_BIAS>(B, N[l], K[l], in, K[l], Wb[l], N[l], Yp, N[l], nullptr, 0, bias,
                                               nullptr)));
        }
    }
    return PF_OK;
}

pf_status pf_warmstart::backward(int B, double *loss) {
    size_t ny = (size_t)B * n_out;
    unsigned nb = (unsigned)((ny + 255) / 256);
    int cur = 0;
    k_out_grad<<<nb, 256, 0, stream>>>(Yp, T, B, n_out, ym, yis, dZf[cur], dZb[cur], loss_part);
    MHIP(hipGetLastError());
    for (int l = n_layers - 1; l >= 0; l--) {
        const bf16 *hin = (l == 0) ? Xn : H[l - 1];
        // dW_l = H_{l-1}^T dZ_l : [K_l x N_l], reduction over the batch
        MCHK((gemm<true, false, EPI_STORE>(K[l], N[l], B, hin, K[l], dZb[cur], N[l], g + w_off[l], N[l], nullptr,
                                           0, nullptr, nullptr)));
        k_colsum<<<(N[l] + 255) / 256, 256, 0, stream>>>(dZf[cur], B, N[l], g + b_off[l]);
        MHIP(hipGetLastError());
        if (l > 0) {
            // dZ_{l-1} = (dZ_l W_l^T) * gelu'(Z_{l-1})
            MCHK((gemm<false, true, EPI_DGELU>(B, K[l], N[l], dZb[cur], N[l], Wb[l], N[l], dZf[cur ^ 1], K[l],
                                               dZb[cur ^ 1], K[l], nullptr, Z[l - 1])));
            cur ^= 1;
        }
    }
    if (loss) {
        std::vector<float> parts(nb);
        MHIP(hipMemcpyAsync(parts.data(), loss_part, sizeof(float) * nb, hipMemcpyDeviceToHost, stream));
        MHIP(hipStreamSynchronize(stream));
        double s = 0;
        for (float x : parts) s += x;
        *loss = s / ((double)B * n_out);
    }
    return PF_OK;
}

extern "C" {

pf_status pf_warmstart_create(int n_in, int n_out, int hidden, int n_hidden_layers, unsigned long long seed,
                              int device, pf_warmstart **out) {
    if (!out || n_in <= 0 || n_out <= 0 || hidden <= 0 || n_hidden_layers < 0 || n_in % 8 || n_out % 8 ||
        hidden % 8)
        return PF_INVALID_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return PF_DEVICE_ERROR;
    auto w = std::make_unique<pf_warmstart>();
    w->device = device;
    MHIP(hipSetDevice(device));
    MHIP(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
    w->n_in = n_in;
    w->n_out = n_out;
    w->hidden = hidden;
    w->n_layers = n_hidden_layers + 1;
    size_t off = 0;
    for (int l = 0; l < w->n_layers; l++) {
        int k = (l == 0) ? n_in : hidden;
        int n = (l == w->n_layers - 1) ? n_out : hidden;
        w->K.push_back(k);
        w->N.push_back(n);
        w->w_off.push_back(off);
        off += (size_t)k * n;
        w->b_off.push_back(off);
        off += n;
    }
    w->n_params = off;
    // He-uniform initialisation (deterministic), zero biases
    std::vector<float> h(off, 0.f);
    unsigned long long s = seed;
    for (int l = 0; l < w->n_layers; l++) {
        const float a = std::sqrt(6.0f / (float)w->K[l]) * (l == w->n_layers - 1 ? 0.1f : 1.0f);
        for (size_t i = 0; i < (size_t)w->K[l] * w->N[l]; i++) {
            double u = (double)(sm64(s) >> 11) * (1.0 / 9007199254740992.0);
            h[w->w_off[l] + i] = (float)((2.0 * u - 1.0) * a);
        }
    }
    MCHK(w->alloc(&w->p, off, w->allocs));
    MCHK(w->alloc(&w->g, off, w->allocs));
    MCHK(w->alloc(&w->m, off, w->allocs));
    MCHK(w->alloc(&w->v, off, w->allocs));
    MHIP(hipMemcpy(w->p, h.data(), sizeof(float) * off, hipMemcpyHostToDevice));
    MHIP(hipMemset(w->m, 0, sizeof(float) * off));
    MHIP(hipMemset(w->v, 0, sizeof(float) * off));
    MHIP(hipMemset(w->g, 0, sizeof(float) * off));
    w->Wb.resize(w->n_layers);
    for (int l = 0; l < w->n_layers; l++) MCHK(w->alloc(&w->Wb[l], (size_t)w->K[l] * w->N[l], w->allocs));
    MCHK(w->alloc(&w->xm, n_in, w->allocs));
    MCHK(w->alloc(&w->xis, n_in, w->allocs));
    MCHK(w->alloc(&w->ym, n_out, w->allocs));
    MCHK(w->alloc(&w->yis, n_out, w->allocs));
    MCHK(w->alloc(&w->ys, n_out, w->allocs));
    std::vector<float> zeros(std::max(n_in, n_out), 0.f), ones(std::max(n_in, n_out), 1.f);
    MHIP(hipMemcpy(w->xm, zeros.data(), 4 * n_in, hipMemcpyHostToDevice));
    MHIP(hipMemcpy(w->xis, ones.data(), 4 * n_in, hipMemcpyHostToDevice));
    MHIP(hipMemcpy(w->ym, zeros.data(), 4 * n_out, hipMemcpyHostToDevice));
    MHIP(hipMemcpy(w->yis, ones.data(), 4 * n_out, hipMemcpyHostToDevice));
    MHIP(hipMemcpy(w->ys, ones.data(), 4 * n_out, hipMemcpyHostToDevice));
    MCHK(w->sync_bf16());
    MHIP(hipStreamSynchronize(w->stream));
    *out = w.release();
    return PF_OK;
}

void pf_warmstart_destroy(pf_warmstart *w) {
    if (!w) return;
    (void)hipSetDevice(w->device);
    delete w;
}

pf_status pf_warmstart_info(pf_warmstart *w, size_t *n_params, int *n_layers) {
    if (!w) return PF_INVALID_ARG;
    if (n_params) *n_params = w->n_params;
    if (n_layers) *n_layers = w->n_layers;
    return PF_OK;
}

pf_status pf_warmstart_set_norm(pf_warmstart *w, const float *x_mean, const float *x_std, const float *y_mean,
                                const float *y_std) {
    if (!w || !x_mean || !x_std || !y_mean || !y_std) return PF_INVALID_ARG;
    MHIP(hipSetDevice(w->device));
    std::vector<float> xi(w->n_in), yi(w->n_out);
    for (int i = 0; i < w->n_in; i++) xi[i] = x_std[i] > 0 ? 1.0f / x_std[i] : 1.0f;
    for (int i = 0; i < w->n_out; i++) yi[i] = y_std[i] > 0 ? 1.0f / y_std[i] : 1.0f;
    std::vector<float> ysd(y_std, y_std + w->n_out);
    for (auto &x : ysd)
        if (!(x > 0)) x = 1.0f;
    MHIP(hipMemcpy(w->xm, x_mean, 4 * w->n_in, hipMemcpyHostToDevice));
    MHIP(hipMemcpy(w->xis, xi.data(), 4 * w->n_in, hipMemcpyHostToDevice));
    MHIP(hipMemcpy(w->ym, y_mean, 4 * w->n_out, hipMemcpyHostToDevice));
    MHIP(hipMemcpy(w->yis, yi.data(), 4 * w->n_out, hipMemcpyHostToDevice));
    MHIP(hipMemcpy(w->ys, ysd.data(), 4 * w->n_out, hipMemcpyHostToDevice));
    return PF_OK;
}

pf_status pf_warmstart_get_params(pf_warmstart *w, float *out) {
    if (!w || !out) return PF_INVALID_ARG;
    MHIP(hipSetDevice(w->device));
    MHIP(hipMemcpy(out, w->p, sizeof(float) * w->n_params, hipMemcpyDeviceToHost));
    return PF_OK;
}

pf_status pf_warmstart_set_params(pf_warmstart *w, const float *in) {
    if (!w || !in) return PF_INVALID_ARG;
    MHIP(hipSetDevice(w->device));
    MHIP(hipMemcpy(w->p, in, sizeof(float) * w->n_params, hipMemcpyHostToDevice));
    MCHK(w->sync_bf16());
    MHIP(hipStreamSynchronize(w->stream));
    return PF_OK;
}

// forward + backward on one minibatch (raw inputs X [B x n_in], raw targets
// Y [B x n_out]); gradients (parameter layout) copied to grads if not NULL
pf_status pf_warmstart_grad(pf_warmstart *w, const float *X, const float *Y, int B, float *grads, double *loss) {
    if (!w || !X || !Y || B <= 0 || B % 8) return PF_INVALID_ARG;
    w->staged = 0;   // its rows replace the staged ones
    MHIP(hipSetDevice(w->device));
    MCHK(w->ensure(B));
    MHIP(hipMemcpyAsync(w->X, X, sizeof(float) * (size_t)B * w->n_in, hipMemcpyHostToDevice, w->stream));
    MHIP(hipMemcpyAsync(w->T, Y, sizeof(float) * (size_t)B * w->n_out, hipMemcpyHostToDevice, w->stream));
    MCHK(w->forward(B));
    MCHK(w->backward(B, loss));
    if (grads) MHIP(hipMemcpyAsync(grads, w->g, sizeof(float) * w->n_params, hipMemcpyDeviceToHost, w->stream));
    MHIP(hipStreamSynchronize(w->stream));
    return PF_OK;
}

// AdamW training over a resident data set: X [n x n_in], Y [n x n_out] raw;
// `epochs` passes in minibatches of B rows (a permutation per epoch from
// `seed`); per-epoch mean loss written to epoch_loss (may be NULL)
pf_status pf_warmstart_fit(pf_warmstart *w, const float *X, const float *Y, int n, int epochs, int B, float lr,
                           float weight_decay, unsigned long long seed, double *epoch_loss) {
    if (!w || !X || !Y || n < B || B <= 0 || B % 8 || epochs <= 0) return PF_INVALID_ARG;
    w->staged = 0;   // the minibatches replace the staged rows
    MHIP(hipSetDevice(w->device));
    MCHK(w->ensure(B));
    float *dX = nullptr, *dY = nullptr;
    int *dperm = nullptr;
    std::vector<void *> tmp;
    MCHK(w->alloc(&dX, (size_t)n * w->n_in, tmp));
    MCHK(w->alloc(&dY, (size_t)n * w->n_out, tmp));
    MCHK(w->alloc(&dperm, (size_t)n, tmp));
    pf_status st = PF_OK;
    MHIP(hipMemcpy(dX, X, sizeof(float) * (size_t)n * w->n_in, hipMemcpyHostToDevice));
    MHIP(hipMemcpy(dY, Y, sizeof(float) * (size_t)n * w->n_out, hipMemcpyHostToDevice));
    std::vector<int> perm(n);
    for (int i = 0; i < n; i++) perm[i] = i;
    unsigned long long s = seed;
    const int nbat = n / B;
    for (int ep = 0; ep < epochs && st == PF_OK; ep++) {
        for (int i = n - 1; i > 0; i--) {
            int j = (int)(sm64(s) % (unsigned long long)(i + 1));
            std::swap(perm[i], perm[j]);
        }
        if (hipMemcpyAsync(dperm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, w->stream) != hipSuccess ||
            hipStreamSynchronize(w->stream) != hipSuccess) {
            st = PF_DEVICE_ERROR;
            break;
        }
        double acc = 0;
        for (int bi = 0; bi < nbat && st == PF_OK; bi++) {
            const size_t nx = (size_t)B * w->n_in, ny = (size_t)B * w->n_out;
            k_gather_rows<<<(unsigned)((nx + 255) / 256), 256, 0, w->stream>>>(dX, dperm + (size_t)bi * B, B, w->n_in, w->X);
            k_gather_rows<<<(unsigned)((ny + 255) / 256), 256, 0, w->stream>>>(dY, dperm + (size_t)bi * B, B, w->n_out, w->T);
            double loss = 0;
            st = w->forward(B);
            if (st == PF_OK) st = w->backward(B, epoch_loss ? &loss : nullptr);
            if (st != PF_OK) break;
            w->step++;
            const float b1 = 0.9f, b2 = 0.999f;
            const float bc1 = 1.0f - std::pow(b1, (float)w->step), bc2 = 1.0f - std::pow(b2, (float)w->step);
            k_adamw<<<(unsigned)((w->n_params + 255) / 256), 256, 0, w->stream>>>(w->p, w->g, w->m, w->v, w->n_params,
                                                                                 lr, b1, b2, 1e-8f, weight_decay,
                                                                                 bc1, bc2);
            st = w->sync_bf16();
            acc += loss;
        }
        if (epoch_loss) epoch_loss[ep] = acc / nbat;
    }
    (void)hipStreamSynchronize(w->stream);
    for (void *q : tmp) (void)hipFree(q);
    return st;
}

// inference: X [B x n_in] raw -> Y [B x n_out] in target units
pf_status pf_warmstart_predict(pf_warmstart *w, const float *X, int B, float *Y) {
    if (!w || !X || !Y || B <= 0 || B % 8) return PF_INVALID_ARG;
    w->staged = 0;   // its rows replace the staged ones
    MHIP(hipSetDevice(w->device));
    MCHK(w->ensure(B));
    MHIP(hipMemcpyAsync(w->X, X, sizeof(float) * (size_t)B * w->n_in, hipMemcpyHostToDevice, w->stream));
    MCHK(w->forward(B));
    size_t ny = (size_t)B * w->n_out;
    k_denorm<<<(unsigned)((ny + 255) / 256), 256, 0, w->stream>>>(w->Yp, B, w->n_out, w->ym, w->ys, w->Yo);
    MHIP(hipGetLastError());
    MHIP(hipMemcpyAsync(Y, w->Yo, sizeof(float) * ny, hipMemcpyDeviceToHost, w->stream));
    MHIP(hipStreamSynchronize(w->stream));
    return PF_OK;
}

pf_status pf_warmstart_stage(pf_warmstart *w, const float *X, int B) {
    if (!w || !X || B <= 0 || B % 8) return PF_INVALID_ARG;
    MHIP(hipSetDevice(w->device));
    MCHK(w->ensure(B));
    MHIP(hipMemcpyAsync(w->X, X, sizeof(float) * (size_t)B * w->n_in, hipMemcpyHostToDevice, w->stream));
    MHIP(hipStreamSynchronize(w->stream));
    w->staged = B;
    return PF_OK;
}

pf_status pf_warmstart_predict_into(pf_warmstart *w, pf_batch *b, int v_col, int th_col) {
    if (!w || !b || w->staged <= 0) return PF_INVALID_ARG;
    MHIP(hipSetDevice(w->device));
    MCHK(w->forward(w->staged));
    return batch_start_from_rows(b, w->device, w->staged, w->Yp, w->n_out, v_col, th_col, w->ym, w->ys,
                                 (void *)w->stream);
}

// one GEMM through the training kernel (tests): C = op(A) op(B), bf16 in, fp32 out
pf_status pf_warmstart_gemm_probe(pf_warmstart *w, int ta, int tb, int M, int N, int K, const float *A,
                                  const float *B, float *C) {
    if (!w || !A || !B || !C || M % 8 || N % 8 || K % 8) return PF_INVALID_ARG;
    MHIP(hipSetDevice(w->device));
    std::vector<void *> tmp;
    bf16 *dA, *dB;
    float *dC, *fA, *fB;
    MCHK(w->alloc(&fA, (size_t)M * K, tmp));
    MCHK(w->alloc(&fB, (size_t)K * N, tmp));
    MCHK(w->alloc(&dA, (size_t)M * K, tmp));
    MCHK(w->alloc(&dB, (size_t)K * N, tmp));
    MCHK(w->alloc(&dC, (size_t)M * N, tmp));
    pf_status st = PF_OK;
    (void)hipMemcpy(fA, A, 4 * (size_t)M * K, hipMemcpyHostToDevice);
    (void)hipMemcpy(fB, B, 4 * (size_t)K * N, hipMemcpyHostToDevice);
    k_to_bf16<<<(unsigned)(((size_t)M * K + 255) / 256), 256, 0, w->stream>>>(fA, dA, (size_t)M * K);
    k_to_bf16<<<(unsigned)(((size_t)K * N + 255) / 256), 256, 0, w->stream>>>(fB, dB, (size_t)K * N);
    const int lda = ta ? M : K, ldb = tb ? K : N;
    if (!ta && !tb) st = w->gemm<false, false, EPI_STORE>(M, N, K, dA, lda, dB, ldb, dC, N, nullptr, 0, nullptr, nullptr);
    else if (ta && !tb) st = w->gemm<true, false, EPI_STORE>(M, N, K, dA, lda, dB, ldb, dC, N, nullptr, 0, nullptr, nullptr);
    else if (!ta && tb) st = w->gemm<false, true, EPI_STORE>(M, N, K, dA, lda, dB, ldb, dC, N, nullptr, 0, nullptr, nullptr);
    else st = PF_INVALID_ARG;
    if (st == PF_OK && hipMemcpyAsync(C, dC, 4 * (size_t)M * N, hipMemcpyDeviceToHost, w->stream) != hipSuccess)
        st = PF_DEVICE_ERROR;
    (void)hipStreamSynchronize(w->stream);
    for (void *q : tmp) (void)hipFree(q);
    return st;
}

}  // extern "C"
