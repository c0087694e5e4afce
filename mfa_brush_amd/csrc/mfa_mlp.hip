// mfa_mlp.hip -- the profile surrogate: a small MLP trained and evaluated with
// hand-written BF16 MFMA kernels (FP32 accumulation, FP32 master weights).
//
// Network: x (n_in) -> [Linear -> SiLU] x (L-1) -> Linear -> y (n_out).
// Parameters are one flat FP32 buffer: W_0 (n_0 x n_1, row-major), b_0 (n_1),
// W_1, b_1, ...  Activations are FP32 in memory and rounded to BF16 (nearest
// even) when a GEMM loads them, so every product runs on the matrix cores.
//
// GEMM tile: one 64-lane wave computes a 32x32 block of C with
// v_mfma_f32_32x32x8_bf16 (K = 8 per instruction).  Lane l holds
//   A: rows i = l % 32, columns k = 4 (l / 32) + j, j = 0..3
//   B: rows k = 4 (l / 32) + j, columns n = l % 32
//   C: column n = l % 32, rows i = 8 (r / 4) + 4 (l / 32) + r % 4, r = 0..15
// (checked on the device by tests/test_gpu_mlp.py against exact integer
// products).  Strides give the transposes the backward pass needs.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/mfa_mlp.h"

namespace mlp {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float16_t __attribute__((ext_vector_type(16)));

// FP32 -> BF16, round to nearest even (NaN kept quiet)
__device__ __forceinline__ short bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (short)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// C[i, n] (+)= sum_k A(i, k) B(k, n);  A(i, k) = A[i*sai + k*sak], B(k, n) = B[k*sbk + n*sbn]
__global__ __launch_bounds__(64) void k_gemm(const float *__restrict__ A, long sai, long sak,
                                            const float *__restrict__ Bm, long sbk, long sbn,
                                            float *__restrict__ C, long ldc, int M, int N, int K, int accumulate) {
    const int l = threadIdx.x;
    const int i0 = blockIdx.x * 32, n0 = blockIdx.y * 32;
    const int ia = i0 + (l & 31), nb = n0 + (l & 31), kq = 4 * (l >> 5);
    float16_t acc = {};
    for (int k0 = 0; k0 < K; k0 += 8) {
        short4_t a, b;
        for (int j = 0; j < 4; ++j) {
            const int k = k0 + kq + j;
            a[j] = (ia < M && k < K) ? bf16_rne(A[(long)ia * sai + (long)k * sak]) : (short)0;
            b[j] = (nb < N && k < K) ? bf16_rne(Bm[(long)k * sbk + (long)nb * sbn]) : (short)0;
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    for (int r = 0; r < 16; ++r) {
        const int i = i0 + 8 * (r >> 2) + kq + (r & 3);
        const int n = n0 + (l & 31);
        if (i < M && n < N) {
            float *c = C + (long)i * ldc + n;
            *c = accumulate ? *c + acc[r] : acc[r];
        }
    }
}

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// z += b (row-wise bias); h = SiLU(z) for hidden layers, h = z for the output
__global__ __launch_bounds__(256) void k_bias_act(float *__restrict__ z, float *__restrict__ h,
                                                 const float *__restrict__ b, int M, int N, int act) {
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long)M * N) return;
    const float v = z[t] + b[t % N];
    z[t] = v;
    h[t] = act ? silu(v) : v;
}

// dz = dh * SiLU'(z)
__global__ __launch_bounds__(256) void k_act_back(float *__restrict__ dz, const float *__restrict__ z, long n) {
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t < n) dz[t] *= dsilu(z[t]);
}

// MSE: dy = 2 (y - t) / (M N); per-block partial sums of (y - t)^2
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ y, const float *__restrict__ tg,
                                            float *__restrict__ dy, float *__restrict__ part, long n, float scale) {
    __shared__ float sh[256];
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    float e = 0.f;
    if (t < n) {
        const float d = y[t] - tg[t];
        dy[t] = 2.0f * d * scale;
        e = d * d;
    }
    sh[threadIdx.x] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// db[n] = sum_i dz[i, n] (fixed order per column)
__global__ __launch_bounds__(256) void k_colsum(const float *__restrict__ dz, float *__restrict__ db, int M, int N) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (int i = 0; i < M; ++i) s += dz[(long)i * N + n];
    db[n] = s;
}

// sum of the per-block partials (one block)
__global__ __launch_bounds__(256) void k_sum(const float *__restrict__ part, int n, float *__restrict__ out) {
    __shared__ float sh[256];
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += 256) s += part[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m1,
                                             float *__restrict__ m2, long n, float lr, float b1, float b2, float eps,
                                             float c1, float c2) {
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const float gt = g[t];
    const float a = b1 * m1[t] + (1.0f - b1) * gt;
    const float b = b2 * m2[t] + (1.0f - b2) * gt * gt;
    m1[t] = a; m2[t] = b;
    p[t] -= lr * (a * c1) / (sqrtf(b * c2) + eps);
}

// gather rows idx[] of x (n_in) and t (n_out) into the batch buffers
__global__ __launch_bounds__(256) void k_gather(const float *__restrict__ x, const float *__restrict__ t,
                                               const int *__restrict__ idx, int m, int n_in, int n_out,
                                               float *__restrict__ xb, float *__restrict__ tb) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= m) return;
    const long s = idx[r];
    for (int q = 0; q < n_in; ++q) xb[(long)r * n_in + q] = x[s * n_in + q];
    for (int q = 0; q < n_out; ++q) tb[(long)r * n_out + q] = t[s * n_out + q];
}

}  // namespace mlp

using namespace mlp;

static thread_local std::string g_err;
static int fail(int code, const std::string &m) { g_err = m; return code; }
#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return fail(MFA_MLP_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); } while (0)

struct mfa_mlp {
    std::vector<int> w;                 // layer widths n_0 .. n_L
    std::vector<long> woff, boff;       // parameter offsets of W_l, b_l
    long n_par = 0;
    hipStream_t st = nullptr;
    float *d_p = nullptr, *d_g = nullptr, *d_m1 = nullptr, *d_m2 = nullptr;
    long adam_t = 0;
    // activation buffers for a batch of up to cap rows
    int cap = 0;
    std::vector<float *> d_h, d_z;      // h_l (input of layer l, l = 0..L) and z_l (pre-activation, l = 0..L-1)
    float *d_dy = nullptr, *d_da = nullptr, *d_db = nullptr, *d_t = nullptr, *d_part = nullptr, *d_loss = nullptr;
};

static void free_batch(mfa_mlp *m) {
    for (float *p : m->d_h) if (p) (void)hipFree(p);
    for (float *p : m->d_z) if (p) (void)hipFree(p);
    m->d_h.clear(); m->d_z.clear();
    float *bufs[] = {m->d_dy, m->d_da, m->d_db, m->d_t, m->d_part, m->d_loss};
    for (float *p : bufs) if (p) (void)hipFree(p);
    m->d_dy = m->d_da = m->d_db = m->d_t = m->d_part = m->d_loss = nullptr;
    m->cap = 0;
}

static int ensure_batch(mfa_mlp *m, int rows) {
    if (rows <= m->cap) return 0;
    free_batch(m);
    const int L = (int)m->w.size() - 1;
    int wmax = 0;
    for (int v : m->w) wmax = std::max(wmax, v);
    m->d_h.assign(L + 1, nullptr);
    m->d_z.assign(L, nullptr);
    for (int l = 0; l <= L; ++l) HIPCHK(hipMalloc((void **)&m->d_h[l], sizeof(float) * (size_t)rows * m->w[l]));
    for (int l = 0; l < L; ++l) HIPCHK(hipMalloc((void **)&m->d_z[l], sizeof(float) * (size_t)rows * m->w[l + 1]));
    HIPCHK(hipMalloc((void **)&m->d_dy, sizeof(float) * (size_t)rows * wmax));
    HIPCHK(hipMalloc((void **)&m->d_da, sizeof(float) * (size_t)rows * wmax));
    HIPCHK(hipMalloc((void **)&m->d_db, sizeof(float) * (size_t)rows * wmax));
    HIPCHK(hipMalloc((void **)&m->d_t, sizeof(float) * (size_t)rows * m->w[L]));
    const long nblk = ((long)rows * m->w[L] + 255) / 256;
    HIPCHK(hipMalloc((void **)&m->d_part, sizeof(float) * (size_t)std::max(1L, nblk)));
    HIPCHK(hipMalloc((void **)&m->d_loss, sizeof(float)));
    m->cap = rows;
    return 0;
}

static void gemm(mfa_mlp *m, const float *A, long sai, long sak, const float *B, long sbk, long sbn, float *C, long ldc,
                 int M, int N, int K, int acc) {
    dim3 g((M + 31) / 32, (N + 31) / 32);
    hipLaunchKernelGGL(k_gemm, g, dim3(64), 0, m->st, A, sai, sak, B, sbk, sbn, C, ldc, M, N, K, acc);
}

static int blocks(long n) { return (int)((n + 255) / 256); }

// forward of rows already in d_h[0]; output in d_h[L]
static void forward(mfa_mlp *m, int rows) {
    const int L = (int)m->w.size() - 1;
    for (int l = 0; l < L; ++l) {
        const int ni = m->w[l], no = m->w[l + 1];
        gemm(m, m->d_h[l], ni, 1, m->d_p + m->woff[l], no, 1, m->d_z[l], no, rows, no, ni, 0);
        hipLaunchKernelGGL(k_bias_act, dim3(blocks((long)rows * no)), dim3(256), 0, m->st, m->d_z[l], m->d_h[l + 1],
                           m->d_p + m->boff[l], rows, no, l < L - 1 ? 1 : 0);
    }
}

// loss and gradient (into d_g) of the rows in d_h[0] against targets d_t
static void backward(mfa_mlp *m, int rows) {
    const int L = (int)m->w.size() - 1;
    const long nout = (long)rows * m->w[L];
    hipLaunchKernelGGL(k_mse, dim3(blocks(nout)), dim3(256), 0, m->st, m->d_h[L], m->d_t, m->d_dy, m->d_part, nout,
                       1.0f / (float)nout);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, m->st, m->d_part, blocks(nout), m->d_loss);
    float *dz = m->d_dy;            // dL/dz of the current layer (rows x no)
    for (int l = L - 1; l >= 0; --l) {
        const int ni = m->w[l], no = m->w[l + 1];
        // dW_l = h_l^T dz  (ni x no), K = rows
        gemm(m, m->d_h[l], 1, ni, dz, no, 1, m->d_g + m->woff[l], no, ni, no, rows, 0);
        hipLaunchKernelGGL(k_colsum, dim3(blocks(no)), dim3(256), 0, m->st, dz, m->d_g + m->boff[l], rows, no);
        if (l == 0) break;
        // dh_l = dz W_l^T  (rows x ni), then through SiLU of layer l-1
        float *dh = dz == m->d_da ? m->d_db : m->d_da;
        gemm(m, dz, no, 1, m->d_p + m->woff[l], 1, no, dh, ni, rows, ni, no, 0);
        hipLaunchKernelGGL(k_act_back, dim3(blocks((long)rows * ni)), dim3(256), 0, m->st, dh, m->d_z[l - 1],
                           (long)rows * ni);
        dz = dh;
    }
}

extern "C" {

const char *mfa_mlp_last_error(void) { return g_err.c_str(); }

int mfa_mlp_create(const int32_t *widths, int32_t n_widths, uint64_t seed, mfa_mlp **out) {
    if (!widths || !out || n_widths < 2 || n_widths > 16) return fail(MFA_MLP_E_ARG, "create: need 2..16 layer widths");
    for (int i = 0; i < n_widths; ++i)
        if (widths[i] < 1 || widths[i] > 4096) return fail(MFA_MLP_E_ARG, "create: widths must be 1..4096");
    mfa_mlp *m = new mfa_mlp;
    m->w.assign(widths, widths + n_widths);
    const int L = n_widths - 1;
    for (int l = 0; l < L; ++l) {
        m->woff.push_back(m->n_par); m->n_par += (long)m->w[l] * m->w[l + 1];
        m->boff.push_back(m->n_par); m->n_par += m->w[l + 1];
    }
    // He-uniform init from a splitmix64 stream of the seed (deterministic)
    std::vector<float> p(m->n_par, 0.0f);
    uint64_t s = seed;
    auto next = [&s]() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    for (int l = 0; l < L; ++l) {
        const double a = std::sqrt(6.0 / m->w[l]);
        for (long q = 0; q < (long)m->w[l] * m->w[l + 1]; ++q)
            p[m->woff[l] + q] = (float)(a * (2.0 * ((next() >> 11) * (1.0 / 9007199254740992.0)) - 1.0));
    }
    if (hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking) != hipSuccess) { delete m; return fail(MFA_MLP_E_HIP, "stream"); }
    const size_t b = sizeof(float) * (size_t)m->n_par;
    if (hipMalloc((void **)&m->d_p, b) != hipSuccess || hipMalloc((void **)&m->d_g, b) != hipSuccess ||
        hipMalloc((void **)&m->d_m1, b) != hipSuccess || hipMalloc((void **)&m->d_m2, b) != hipSuccess) {
        mfa_mlp_destroy(m); return fail(MFA_MLP_E_HIP, "parameter allocation");
    }
    if (hipMemcpy(m->d_p, p.data(), b, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(m->d_m1, 0, b) != hipSuccess || hipMemset(m->d_m2, 0, b) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {
        mfa_mlp_destroy(m); return fail(MFA_MLP_E_HIP, "parameter upload");
    }
    *out = m;
    return 0;
}

int mfa_mlp_destroy(mfa_mlp *m) {
    if (!m) return 0;
    if (m->st) (void)hipStreamSynchronize(m->st);
    free_batch(m);
    float *bufs[] = {m->d_p, m->d_g, m->d_m1, m->d_m2};
    for (float *p : bufs) if (p) (void)hipFree(p);
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
    return 0;
}

int mfa_mlp_num_params(mfa_mlp *m, int64_t *n) {
    if (!m || !n) return fail(MFA_MLP_E_ARG, "num_params: null");
    *n = m->n_par;
    return 0;
}

int mfa_mlp_get_params(mfa_mlp *m, float *p, int64_t n) {
    if (!m || !p || n != m->n_par) return fail(MFA_MLP_E_ARG, "get_params: bad argument");
    HIPCHK(hipStreamSynchronize(m->st));
    HIPCHK(hipMemcpy(p, m->d_p, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost));
    return 0;
}

int mfa_mlp_set_params(mfa_mlp *m, const float *p, int64_t n) {
    if (!m || !p || n != m->n_par) return fail(MFA_MLP_E_ARG, "set_params: bad argument");
    HIPCHK(hipStreamSynchronize(m->st));
    HIPCHK(hipMemcpy(m->d_p, p, sizeof(float) * (size_t)n, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(m->d_m1, 0, sizeof(float) * (size_t)n));
    HIPCHK(hipMemset(m->d_m2, 0, sizeof(float) * (size_t)n));
    HIPCHK(hipDeviceSynchronize());
    m->adam_t = 0;
    return 0;
}

int mfa_mlp_forward(mfa_mlp *m, const float *x, int32_t rows, float *y) {
    if (!m || !x || !y || rows < 1) return fail(MFA_MLP_E_ARG, "forward: bad argument");
    if (int rc = ensure_batch(m, rows)) return rc;
    const int L = (int)m->w.size() - 1;
    HIPCHK(hipStreamSynchronize(m->st));
    HIPCHK(hipMemcpy(m->d_h[0], x, sizeof(float) * (size_t)rows * m->w[0], hipMemcpyHostToDevice));
    forward(m, rows);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(m->st));
    HIPCHK(hipMemcpy(y, m->d_h[L], sizeof(float) * (size_t)rows * m->w[L], hipMemcpyDeviceToHost));
    return 0;
}

int mfa_mlp_gradient(mfa_mlp *m, const float *x, const float *t, int32_t rows, float *loss, float *grad) {
    if (!m || !x || !t || rows < 1) return fail(MFA_MLP_E_ARG, "gradient: bad argument");
    if (int rc = ensure_batch(m, rows)) return rc;
    const int L = (int)m->w.size() - 1;
    HIPCHK(hipStreamSynchronize(m->st));
    HIPCHK(hipMemcpy(m->d_h[0], x, sizeof(float) * (size_t)rows * m->w[0], hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m->d_t, t, sizeof(float) * (size_t)rows * m->w[L], hipMemcpyHostToDevice));
    forward(m, rows);
    backward(m, rows);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(m->st));
    float ls = 0.f;
    HIPCHK(hipMemcpy(&ls, m->d_loss, sizeof(float), hipMemcpyDeviceToHost));
    if (loss) *loss = ls / (float)((long)rows * m->w[L]);
    if (grad) HIPCHK(hipMemcpy(grad, m->d_g, sizeof(float) * (size_t)m->n_par, hipMemcpyDeviceToHost));
    return 0;
}

int mfa_mlp_train(mfa_mlp *m, const float *x, const float *t, int64_t n_rows, int32_t steps, int32_t batch,
                  float lr, uint64_t seed, float *loss_history) {
    if (!m || !x || !t || n_rows < 1 || n_rows > (1 << 30) || steps < 0 || batch < 1 || !(lr > 0.0f))
        return fail(MFA_MLP_E_ARG, "train: bad argument");
    if (steps == 0) return 0;
    const int L = (int)m->w.size() - 1;
    const int bt = (int)std::min<int64_t>(batch, n_rows);
    if (int rc = ensure_batch(m, bt)) return rc;
    // minibatches: a fresh permutation of the rows each epoch (splitmix64 of
    // the seed), all drawn up front and uploaded once
    std::vector<int> perm(n_rows), idx((size_t)steps * bt);
    for (int64_t i = 0; i < n_rows; ++i) perm[i] = (int)i;
    uint64_t s = seed;
    auto next = [&s]() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    int64_t pos = n_rows;
    for (size_t r = 0; r < idx.size(); ++r) {
        if (pos >= n_rows) {
            for (int64_t i = n_rows - 1; i > 0; --i) std::swap(perm[i], perm[next() % (uint64_t)(i + 1)]);
            pos = 0;
        }
        idx[r] = perm[pos++];
    }
    HIPCHK(hipStreamSynchronize(m->st));
    float *d_x = nullptr, *d_t = nullptr, *d_lh = nullptr;
    int *d_ix = nullptr;
    HIPCHK(hipMalloc((void **)&d_x, sizeof(float) * (size_t)n_rows * m->w[0]));
    HIPCHK(hipMalloc((void **)&d_t, sizeof(float) * (size_t)n_rows * m->w[L]));
    HIPCHK(hipMalloc((void **)&d_ix, sizeof(int) * idx.size()));
    HIPCHK(hipMalloc((void **)&d_lh, sizeof(float) * (size_t)steps));
    HIPCHK(hipMemcpy(d_x, x, sizeof(float) * (size_t)n_rows * m->w[0], hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_t, t, sizeof(float) * (size_t)n_rows * m->w[L], hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_ix, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const long nout = (long)bt * m->w[L];
    for (int it = 0; it < steps; ++it) {
        hipLaunchKernelGGL(k_gather, dim3(blocks(bt)), dim3(256), 0, m->st, (const float *)d_x, (const float *)d_t,
                           (const int *)(d_ix + (size_t)it * bt), bt, m->w[0], m->w[L], m->d_h[0], m->d_t);
        forward(m, bt);
        backward(m, bt);
        hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, m->st, m->d_part, blocks(nout), d_lh + it);
        m->adam_t += 1;
        const float c1 = 1.0f / (1.0f - std::pow(b1, (float)m->adam_t));
        const float c2 = 1.0f / (1.0f - std::pow(b2, (float)m->adam_t));
        hipLaunchKernelGGL(k_adam, dim3(blocks(m->n_par)), dim3(256), 0, m->st, m->d_p, m->d_g, m->d_m1, m->d_m2,
                           m->n_par, lr, b1, b2, eps, c1, c2);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(m->st);
    std::vector<float> lh(steps);
    if (e == hipSuccess) e = hipMemcpy(lh.data(), d_lh, sizeof(float) * (size_t)steps, hipMemcpyDeviceToHost);
    (void)hipFree(d_x); (void)hipFree(d_t); (void)hipFree(d_ix); (void)hipFree(d_lh);
    if (e != hipSuccess) return fail(MFA_MLP_E_HIP, std::string("train: ") + hipGetErrorString(e));
    if (loss_history)
        for (int it = 0; it < steps; ++it) loss_history[it] = lh[it] / (float)nout;
    return 0;
}

int mfa_mlp_gemm_check(const float *a, const float *b, int32_t M, int32_t N, int32_t K, float *c) {
    // C = A B with A (M x K) and B (K x N) row-major, through the MFMA tile
    if (!a || !b || !c || M < 1 || N < 1 || K < 1) return fail(MFA_MLP_E_ARG, "gemm_check: bad argument");
    float *da = nullptr, *db = nullptr, *dc = nullptr;
    HIPCHK(hipMalloc((void **)&da, sizeof(float) * (size_t)M * K));
    HIPCHK(hipMalloc((void **)&db, sizeof(float) * (size_t)K * N));
    HIPCHK(hipMalloc((void **)&dc, sizeof(float) * (size_t)M * N));
    HIPCHK(hipMemcpy(da, a, sizeof(float) * (size_t)M * K, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(db, b, sizeof(float) * (size_t)K * N, hipMemcpyHostToDevice));
    dim3 g((M + 31) / 32, (N + 31) / 32);
    hipLaunchKernelGGL(k_gemm, g, dim3(64), 0, 0, (const float *)da, (long)K, 1L, (const float *)db, (long)N, 1L, dc,
                       (long)N, M, N, K, 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(c, dc, sizeof(float) * (size_t)M * N, hipMemcpyDeviceToHost));
    (void)hipFree(da); (void)hipFree(db); (void)hipFree(dc);
    return 0;
}

}  // extern "C"
