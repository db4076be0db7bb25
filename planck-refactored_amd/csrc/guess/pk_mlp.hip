// pk_mlp.hip — the atom-pair-block network of the learned initial guess
// (include/pk_mlp.h): FP32 tiled GEMMs with fused bias / tanh, their backward
// passes, the mean-squared-error loss and Adam, all hand-written for gfx950.
//
// GEMM: C[m][n] = sum_k A(m,k) B(k,n), operands addressed through (row, col)
// strides so one kernel serves X W, X^T dY and dY W^T.  A workgroup of 256
// lanes computes a 64 x 64 tile, each lane a 4 x 4 block, with 64 x 16 slabs of
// A and B staged through LDS (padded rows: no bank conflicts on the column
// reads).  Epilogues: + bias, tanh, or multiply by (1 - H^2) (tanh backward).
#include "pk_fock.h"
#include "pk_mlp.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define MLP_HIP(expr)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return fail(PK_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

enum { EPI_NONE = 0, EPI_BIAS = 1, EPI_BIAS_TANH = 2, EPI_DTANH = 3 };

constexpr int TM = 64, TN = 64, TK = 16;

struct GemmArgs {
    const float *A;
    long long sam, sak;  // A(m, k) = A[m * sam + k * sak]
    const float *B;
    long long sbk, sbn;  // B(k, n) = B[k * sbk + n * sbn]
    float *C;            // C[m * ldc + n]
    long long ldc;
    const float *bias;   // EPI_BIAS*: bias[n]
    const float *H;      // EPI_DTANH: C *= 1 - H[m * ldc + n]^2
    int M, N, K;
};

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
    __shared__ float As[TK][TM + 4];
    __shared__ float Bs[TK][TN + 4];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    float acc[4][4] = {};
    for (int k0 = 0; k0 < g.K; k0 += TK) {
        // stage A (64 x 16) and B (16 x 64): 1024 values each, 4 per lane
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int e = threadIdx.x + 256 * r;
            const int am = e / TK, ak = e % TK;  // A slab: row-major in k
            const int gm = m0 + am, gk = k0 + ak;
            As[ak][am] = (gm < g.M && gk < g.K) ? g.A[gm * g.sam + gk * g.sak] : 0.0f;
            const int bk = e / TN, bn = e % TN;
            const int hk = k0 + bk, hn = n0 + bn;
            Bs[bk][bn] = (hk < g.K && hn < g.N) ? g.B[hk * g.sbk + hn * g.sbn] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int m = m0 + ty + 16 * i;
        if (m >= g.M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int n = n0 + tx + 16 * j;
            if (n >= g.N) continue;
            float v = acc[i][j];
            if (EPI == EPI_BIAS || EPI == EPI_BIAS_TANH) v += g.bias[n];
            if (EPI == EPI_BIAS_TANH) v = tanhf(v);
            if (EPI == EPI_DTANH) {
                const float h = g.H[m * g.ldc + n];
                v *= 1.0f - h * h;
            }
            g.C[m * g.ldc + n] = v;
        }
    }
}

// column sums of D[rows][cols] -> out[cols] (bias gradients); one lane per column
__global__ void k_colsum(const float *__restrict__ D, int rows, int cols, float *__restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.0f;
    for (int r = 0; r < rows; ++r) s += D[(size_t)r * cols + c];
    out[c] = s;
}

// dY = 2 (Y - T) / (rows cols); loss partial sums of (Y - T)^2 / (rows cols)
__global__ void k_mse(const float *__restrict__ Y, const float *__restrict__ T, size_t n, float scale,
                      float *__restrict__ dY, double *__restrict__ loss) {
    __shared__ double part[256];
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float d = Y[i] - T[i];
        dY[i] = 2.0f * d * scale;
        s += (double)d * d;
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, part[0] * (double)scale);
}

__global__ void k_zero_d(double *x) { *x = 0.0; }

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, size_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

}  // namespace

struct pk_mlp {
    int nin = 0, hid = 0, nout = 0, maxb = 0;
    size_t np = 0;
    size_t oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0;
    float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
    float *h1 = nullptr, *h2 = nullptr, *y = nullptr, *dy = nullptr, *d1 = nullptr, *d2 = nullptr;
    double *loss = nullptr;
    long long step = 0;
    hipStream_t s = nullptr;
    bool own = false;
};

namespace {

template <int EPI>
int gemm(pk_mlp *m, const GemmArgs &a) {
    const dim3 grid((a.N + TN - 1) / TN, (a.M + TM - 1) / TM);
    hipLaunchKernelGGL(k_gemm<EPI>, grid, dim3(256), 0, m->s, a);
    MLP_HIP(hipGetLastError());
    return PK_OK;
}

// row-major X[rows][k] times W[k][n] (+ bias, act) -> C[rows][n]
template <int EPI>
int dense(pk_mlp *m, const float *X, int rows, int k, const float *W, int n, const float *bias, float *C) {
    GemmArgs a{X, k, 1, W, n, 1, C, n, bias, nullptr, rows, n, k};
    return gemm<EPI>(m, a);
}

int forward(pk_mlp *m, const float *X, int b, float *Y) {
    int rc;
    if ((rc = dense<EPI_BIAS_TANH>(m, X, b, m->nin, m->p + m->oW1, m->hid, m->p + m->ob1, m->h1)) != PK_OK) return rc;
    if ((rc = dense<EPI_BIAS_TANH>(m, m->h1, b, m->hid, m->p + m->oW2, m->hid, m->p + m->ob2, m->h2)) != PK_OK)
        return rc;
    return dense<EPI_BIAS>(m, m->h2, b, m->hid, m->p + m->oW3, m->nout, m->p + m->ob3, Y);
}

// dW[k][n] = In^T dZ  (In [rows][k], dZ [rows][n]); db = colsum dZ
int weight_grad(pk_mlp *m, const float *In, int rows, int k, const float *dZ, int n, float *dW, float *db) {
    GemmArgs a{In, 1, k, dZ, n, 1, dW, n, nullptr, nullptr, k, n, rows};
    int rc;
    if ((rc = gemm<EPI_NONE>(m, a)) != PK_OK) return rc;
    hipLaunchKernelGGL(k_colsum, dim3((n + 255) / 256), dim3(256), 0, m->s, dZ, rows, n, db);
    MLP_HIP(hipGetLastError());
    return PK_OK;
}

// dIn[rows][k] = (dZ W^T) * (1 - H^2)   (W [k][n], H = tanh output of the layer below)
int input_grad(pk_mlp *m, const float *dZ, int rows, int n, const float *W, int k, const float *H, float *dIn) {
    GemmArgs a{dZ, n, 1, W, 1, n, dIn, k, nullptr, H, rows, k, n};
    return gemm<EPI_DTANH>(m, a);
}

}  // namespace

extern "C" {

const char *pk_mlp_last_error(void) { return g_err.c_str(); }

int pk_mlp_create(int n_in, int hidden, int n_out, int max_batch, uint64_t seed, pk_mlp **out) {
    if (!out || n_in <= 0 || hidden <= 0 || n_out <= 0 || max_batch <= 0)
        return fail(PK_ERR_ARG, "pk_mlp_create: bad sizes");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(PK_ERR_HIP, "pk_mlp_create: no HIP device visible");
    pk_mlp *m = new pk_mlp;
    m->nin = n_in;
    m->hid = hidden;
    m->nout = n_out;
    m->maxb = max_batch;
    size_t o = 0;
    m->oW1 = o; o += (size_t)n_in * hidden;
    m->ob1 = o; o += hidden;
    m->oW2 = o; o += (size_t)hidden * hidden;
    m->ob2 = o; o += hidden;
    m->oW3 = o; o += (size_t)hidden * n_out;
    m->ob3 = o; o += n_out;
    m->np = o;
    std::vector<float> init(o, 0.0f);
    std::mt19937_64 rng(seed);
    auto glorot = [&](size_t off, int fi, int fo) {
        const double a = std::sqrt(6.0 / (fi + fo));
        std::uniform_real_distribution<double> u(-a, a);
        for (size_t i = 0; i < (size_t)fi * fo; ++i) init[off + i] = (float)u(rng);
    };
    glorot(m->oW1, n_in, hidden);
    glorot(m->oW2, hidden, hidden);
    glorot(m->oW3, hidden, n_out);
    const size_t bh = (size_t)max_batch * hidden, bo = (size_t)max_batch * n_out;
    bool ok = hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking) == hipSuccess;
    m->own = ok;
    for (float **b : {&m->p, &m->g, &m->m1, &m->m2})
        ok = ok && hipMalloc((void **)b, o * sizeof(float)) == hipSuccess;
    for (float **b : {&m->h1, &m->h2, &m->d1, &m->d2}) ok = ok && hipMalloc((void **)b, bh * sizeof(float)) == hipSuccess;
    for (float **b : {&m->y, &m->dy}) ok = ok && hipMalloc((void **)b, bo * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc((void **)&m->loss, sizeof(double)) == hipSuccess;
    ok = ok && hipMemcpy(m->p, init.data(), o * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemset(m->m1, 0, o * sizeof(float)) == hipSuccess && hipMemset(m->m2, 0, o * sizeof(float)) == hipSuccess &&
         hipMemset(m->g, 0, o * sizeof(float)) == hipSuccess;
    if (!ok) {
        pk_mlp_destroy(m);
        return fail(PK_ERR_HIP, "pk_mlp_create: device allocation failed");
    }
    *out = m;
    return PK_OK;
}

int pk_mlp_destroy(pk_mlp *m) {
    if (!m) return PK_OK;
    if (m->s) (void)hipStreamSynchronize(m->s);
    for (void *b : {(void *)m->p, (void *)m->g, (void *)m->m1, (void *)m->m2, (void *)m->h1, (void *)m->h2,
                    (void *)m->y, (void *)m->dy, (void *)m->d1, (void *)m->d2, (void *)m->loss})
        if (b) (void)hipFree(b);
    if (m->own && m->s) (void)hipStreamDestroy(m->s);
    delete m;
    return PK_OK;
}

int pk_mlp_set_stream(pk_mlp *m, void *stream) {
    if (!m) return fail(PK_ERR_ARG, "pk_mlp_set_stream: null network");
    MLP_HIP(hipStreamSynchronize(m->s));
    if (m->own) (void)hipStreamDestroy(m->s);
    if (stream) {
        m->s = (hipStream_t)stream;
        m->own = false;
    } else {
        MLP_HIP(hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking));
        m->own = true;
    }
    return PK_OK;
}

size_t pk_mlp_nparams(const pk_mlp *m) { return m ? m->np : 0; }

int pk_mlp_get_params(pk_mlp *m, float *host) {
    if (!m || !host) return fail(PK_ERR_ARG, "pk_mlp_get_params: null argument");
    MLP_HIP(hipMemcpyAsync(host, m->p, m->np * sizeof(float), hipMemcpyDeviceToHost, m->s));
    MLP_HIP(hipStreamSynchronize(m->s));
    return PK_OK;
}

int pk_mlp_set_params(pk_mlp *m, const float *host) {
    if (!m || !host) return fail(PK_ERR_ARG, "pk_mlp_set_params: null argument");
    MLP_HIP(hipMemcpyAsync(m->p, host, m->np * sizeof(float), hipMemcpyHostToDevice, m->s));
    MLP_HIP(hipStreamSynchronize(m->s));
    return PK_OK;
}

int pk_mlp_get_grads(pk_mlp *m, float *host) {
    if (!m || !host) return fail(PK_ERR_ARG, "pk_mlp_get_grads: null argument");
    MLP_HIP(hipMemcpyAsync(host, m->g, m->np * sizeof(float), hipMemcpyDeviceToHost, m->s));
    MLP_HIP(hipStreamSynchronize(m->s));
    return PK_OK;
}

int pk_mlp_forward(pk_mlp *m, const float *X, int batch, float *Y) {
    if (!m || !X || !Y) return fail(PK_ERR_ARG, "pk_mlp_forward: null argument");
    if (batch <= 0 || batch > m->maxb) return fail(PK_ERR_ARG, "pk_mlp_forward: batch outside 1..max_batch");
    return forward(m, X, batch, Y);
}

int pk_mlp_loss_grad(pk_mlp *m, const float *X, const float *T, int batch, float *loss) {
    if (!m || !X || !T) return fail(PK_ERR_ARG, "pk_mlp_loss_grad: null argument");
    if (batch <= 0 || batch > m->maxb) return fail(PK_ERR_ARG, "pk_mlp_loss_grad: batch outside 1..max_batch");
    int rc;
    if ((rc = forward(m, X, batch, m->y)) != PK_OK) return rc;
    const size_t ny = (size_t)batch * m->nout;
    hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(1), 0, m->s, m->loss);
    hipLaunchKernelGGL(k_mse, dim3((unsigned)std::min<size_t>((ny + 255) / 256, 1024)), dim3(256), 0, m->s, m->y, T, ny,
                       1.0f / (float)ny, m->dy, m->loss);
    MLP_HIP(hipGetLastError());
    float *g = m->g;
    // layer 3 (linear): dW3 = h2^T dy, db3 = colsum dy; d2 = (dy W3^T) (1 - h2^2)
    if ((rc = weight_grad(m, m->h2, batch, m->hid, m->dy, m->nout, g + m->oW3, g + m->ob3)) != PK_OK) return rc;
    if ((rc = input_grad(m, m->dy, batch, m->nout, m->p + m->oW3, m->hid, m->h2, m->d2)) != PK_OK) return rc;
    // layer 2: dW2 = h1^T d2, db2; d1 = (d2 W2^T) (1 - h1^2)
    if ((rc = weight_grad(m, m->h1, batch, m->hid, m->d2, m->hid, g + m->oW2, g + m->ob2)) != PK_OK) return rc;
    if ((rc = input_grad(m, m->d2, batch, m->hid, m->p + m->oW2, m->hid, m->h1, m->d1)) != PK_OK) return rc;
    // layer 1: dW1 = X^T d1, db1
    if ((rc = weight_grad(m, X, batch, m->nin, m->d1, m->hid, g + m->oW1, g + m->ob1)) != PK_OK) return rc;
    if (loss) {
        double l = 0.0;
        MLP_HIP(hipMemcpyAsync(&l, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->s));
        MLP_HIP(hipStreamSynchronize(m->s));
        *loss = (float)l;
    }
    return PK_OK;
}

int pk_mlp_adam_step(pk_mlp *m, float lr, float beta1, float beta2, float eps) {
    if (!m) return fail(PK_ERR_ARG, "pk_mlp_adam_step: null network");
    m->step += 1;
    const float c1 = 1.0f - std::pow(beta1, (float)m->step), c2 = 1.0f - std::pow(beta2, (float)m->step);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((m->np + 255) / 256)), dim3(256), 0, m->s, m->p, m->g, m->m1, m->m2, m->np,
                       lr, beta1, beta2, eps, c1, c2);
    MLP_HIP(hipGetLastError());
    return PK_OK;
}

}  // extern "C"
