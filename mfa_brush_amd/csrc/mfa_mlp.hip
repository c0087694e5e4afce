// mfa_mlp.hip -- the profile surrogate: a small MLP trained and evaluated with
// hand-written BF16 MFMA kernels (FP32 accumulation, FP32 master weights).
//
// Network: x (n_in) -> [Linear -> SiLU] x (L-1) -> Linear -> y (n_out).
// Parameters are one flat FP32 buffer: W_0 (n_0 x n_1, row-major), b_0 (n_1),
// W_1, b_1, ...  Activations are FP32 in memory and rounded to BF16 (nearest
// even) when a GEMM loads them, so every product runs on the matrix cores.
//
// GEMM: a block of four 64-lane waves computes a 64x64 tile of C; the
// operands are staged per K chunk of 32 in LDS as BF16 (rounded once, nearest
// even), A row-major and B transposed, so each wave reads its fragments with
// 16-byte LDS loads and issues v_mfma_f32_32x32x16_bf16 (gfx950, K = 16 per
// instruction) on its 32x32 quarter.  Lane l of a wave holds
//   A: row i = l % 32, columns k = 8 (l / 32) + j, j = 0..7
//   B: column n = l % 32, rows k = 8 (l / 32) + j
//   C: column n = l % 32, rows i = 8 (r / 4) + 4 (l / 32) + r % 4, r = 0..15
// (checked on the device by tests/test_gpu_mlp.py against exact integer
// products).  Strides give the transposes the backward pass needs; the
// staging loop maps consecutive threads along whichever operand dimension
// has unit stride, so the global loads coalesce either way.  The forward
// pass adds the bias and the SiLU in the GEMM's epilogue.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/mfa_mlp.h"

namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float float16_t __attribute__((ext_vector_type(16)));

// FP32 -> BF16 bits, round to nearest even (NaN kept quiet)
__device__ __forceinline__ unsigned short bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float silu(float z);

constexpr int GT = 64;          // block tile of C (rows and columns)
constexpr int KC = 32;          // K chunk staged in LDS
constexpr int KP = KC + 8;      // LDS row length in BF16 elements (80 bytes: 16-byte aligned rows)

// Epilogue of the GEMM: plain store, accumulate into C, or z = acc + b stored
// in C with h = z (EPI_BIAS) or h = SiLU(z) (EPI_SILU) stored in H.
enum Epi { EPI_STORE = 0, EPI_ACC, EPI_BIAS, EPI_SILU };

__device__ __forceinline__ float apply_epi(float v, float *c, int epi, const float *bias, float *h, int n) {
    if (epi == EPI_ACC) v += *c;
    if (epi >= EPI_BIAS) {
        v += bias[n];
        *h = epi == EPI_SILU ? silu(v) : v;
    }
    return v;
}

// C[i, n] (op)= sum_k A(i, k) B(k, n);  A(i, k) = A[i*sai + k*sak], B(k, n) = B[k*sbk + n*sbn].
// Split K: block z of gridDim.z takes k in [z kslice, (z + 1) kslice) and, when
// gridDim.z > 1, writes its raw partial tile to ws + z M N (row-major, ld N)
// for k_splitk_reduce.  The next chunk's operands are loaded into registers
// while the current chunk's MFMAs run.
__global__ __launch_bounds__(256) void k_gemm(const float *__restrict__ A, long sai, long sak,
                                             const float *__restrict__ Bm, long sbk, long sbn,
                                             float *__restrict__ C, long ldc, int M, int N, int K, int epi,
                                             const float *__restrict__ bias, float *__restrict__ H,
                                             int kslice, float *__restrict__ ws) {
    __shared__ __attribute__((aligned(16))) unsigned short As[GT * KP];     // As[i][k]
    __shared__ __attribute__((aligned(16))) unsigned short Bs[GT * KP];     // Bs[n][k]
    constexpr int PER = GT * KC / 256;                                       // staged elements per thread
    const int t = threadIdx.x, l = t & 63, w = t >> 6;
    const int i0 = blockIdx.x * GT, n0 = blockIdx.y * GT;
    const int kb = blockIdx.z * kslice, ke = min(K, kb + kslice);
    const int wi = 32 * (w & 1), wn = 32 * (w >> 1);
    const bool a_k_fast = sak == 1, b_n_fast = sbn == 1;
    float ra[PER], rb[PER];
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = t + 256 * q;
            const int r = a_k_fast ? e / KC : e % GT, ka = a_k_fast ? e % KC : e / GT;
            const int gi = i0 + r, gka = k0 + ka;
            ra[q] = (gi < M && gka < ke) ? A[(long)gi * sai + (long)gka * sak] : 0.0f;
            const int n = b_n_fast ? e % GT : e / KC, kbb = b_n_fast ? e / GT : e % KC;
            const int gn = n0 + n, gkb = k0 + kbb;
            rb[q] = (gn < N && gkb < ke) ? Bm[(long)gkb * sbk + (long)gn * sbn] : 0.0f;
        }
    };
    float16_t acc = {};
    if (kb < ke) load(kb);
    for (int k0 = kb; k0 < ke; k0 += KC) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = t + 256 * q;
            const int r = a_k_fast ? e / KC : e % GT, ka = a_k_fast ? e % KC : e / GT;
            As[r * KP + ka] = bf16_rne(ra[q]);
            const int n = b_n_fast ? e % GT : e / KC, kbb = b_n_fast ? e / GT : e % KC;
            Bs[n * KP + kbb] = bf16_rne(rb[q]);
        }
        __syncthreads();
        if (k0 + KC < ke) load(k0 + KC);
#pragma unroll
        for (int kk = 0; kk < KC; kk += 16) {
            const bf16x8 a = *(const bf16x8 *)&As[(wi + (l & 31)) * KP + kk + 8 * (l >> 5)];
            const bf16x8 b = *(const bf16x8 *)&Bs[(wn + (l & 31)) * KP + kk + 8 * (l >> 5)];
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const int n = n0 + wn + (l & 31);
    for (int r = 0; r < 16; ++r) {
        const int i = i0 + wi + 8 * (r >> 2) + 4 * (l >> 5) + (r & 3);
        if (i < M && n < N) {
            if (gridDim.z > 1) {
                ws[(long)blockIdx.z * M * N + (long)i * N + n] = acc[r];
            } else {
                float *c = C + (long)i * ldc + n;
                *c = apply_epi(acc[r], c, epi, bias, H + (long)i * ldc + n, n);
            }
        }
    }
}

// C (op)= sum over z = 0 .. S-1, in order, of the split-K partials
__global__ __launch_bounds__(256) void k_splitk_reduce(const float *__restrict__ ws, int S, int M, int N,
                                                      float *__restrict__ C, long ldc, int epi,
                                                      const float *__restrict__ bias, float *__restrict__ H) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)M * N) return;
    const int i = (int)(e / N), n = (int)(e % N);
    float v = 0.0f;
    for (int z = 0; z < S; ++z) v += ws[(long)z * M * N + e];
    float *c = C + (long)i * ldc + n;
    *c = apply_epi(v, c, epi, bias, H + (long)i * ldc + n, n);
}

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
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

// db[n] = sum_i dz[i, n]: a block takes 32 columns, its 8 row groups sum
// rows g, g + 8, ... (four running sums, so four loads are in flight) and the
// groups' sums are added in order (deterministic)
__global__ __launch_bounds__(256) void k_colsum(const float *__restrict__ dz, float *__restrict__ db, int M, int N) {
    __shared__ float sh[8][32];
    const int c = threadIdx.x & 31, g = threadIdx.x >> 5;
    const int n = blockIdx.x * 32 + c;
    float s = 0.f;
    if (n < N) {
        float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
        int i = g;
        for (; i + 24 < M; i += 32) {
            s0 += dz[(long)i * N + n];
            s1 += dz[(long)(i + 8) * N + n];
            s2 += dz[(long)(i + 16) * N + n];
            s3 += dz[(long)(i + 24) * N + n];
        }
        for (; i < M; i += 8) s0 += dz[(long)i * N + n];
        s = (s0 + s1) + (s2 + s3);
    }
    sh[g][c] = s;
    __syncthreads();
    if (g == 0 && n < N) {
        float v = sh[0][c];
        for (int q = 1; q < 8; ++q) v += sh[q][c];
        db[n] = v;
    }
}

// sum of the per-block partials (one block), into out[*slot] (out[0] when
// slot is null)
__global__ __launch_bounds__(256) void k_sum(const float *__restrict__ part, int n, float *__restrict__ out,
                                            const int *__restrict__ slot) {
    __shared__ float sh[256];
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += 256) s += part[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[slot ? *slot : 0] = sh[0];
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters, at Adam step
// t0 + *step + 1 (the bias corrections computed on the device, so a captured
// training step replays unchanged).  With lr_final < 1 the learning rate
// follows a cosine from lr at the call's first step to lr * lr_final at its
// last (step *step of n_steps); lr_final == 1 keeps it constant.
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m1,
                                             float *__restrict__ m2, long n, float lr, float b1, float b2, float eps,
                                             const int *__restrict__ step, long t0, int n_steps, float lr_final) {
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    if (lr_final != 1.0f) {
        const float u = n_steps > 1 ? (float)*step / (float)(n_steps - 1) : 0.0f;
        lr = lr * (lr_final + (1.0f - lr_final) * 0.5f * (1.0f + cosf(3.14159265f * u)));
    }
    const float ts = (float)(t0 + *step + 1);
    const float c1 = 1.0f / (1.0f - powf(b1, ts));
    const float c2 = 1.0f / (1.0f - powf(b2, ts));
    const float gt = g[t];
    const float a = b1 * m1[t] + (1.0f - b1) * gt;
    const float b = b2 * m2[t] + (1.0f - b2) * gt * gt;
    m1[t] = a; m2[t] = b;
    p[t] -= lr * (a * c1) / (sqrtf(b * c2) + eps);
}

// the training loop's step counter
__global__ void k_step_inc(int *step) { *step += 1; }

// gather rows idx[*step * m + r] of x (n_in) and t (n_out) into the batch buffers
__global__ __launch_bounds__(256) void k_gather(const float *__restrict__ x, const float *__restrict__ t,
                                               const int *__restrict__ idx, const int *__restrict__ step, int m,
                                               int n_in, int n_out, float *__restrict__ xb, float *__restrict__ tb) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= m) return;
    const long s = idx[(long)*step * m + r];
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
    float *d_ws = nullptr;              // split-K workspace of the batch's GEMMs
    int *d_step = nullptr;              // the training loop's step counter
    bool use_graph = true;              // train by replaying one captured step
    float lr_final = 1.0f;              // learning rate at a train call's last step, relative (1: constant)
};

static int blocks(long n) { return (int)((n + 255) / 256); }

// Split-K plan of a GEMM: with fewer than 64 output tiles and at least 4 K
// chunks, the K range is split over up to 16 blocks per tile (each at least
// two chunks), so a small weight-gradient GEMM with a long K fills more CUs.
// Returns the number of slices S and their length.
static int split_plan(int M, int N, int K, int *kslice) {
    const int tiles = ((M + GT - 1) / GT) * ((N + GT - 1) / GT);
    const int chunks = (K + KC - 1) / KC;
    int S = 1;
    if (tiles < 64 && chunks >= 4) S = std::max(1, std::min({chunks / 2, (128 + tiles - 1) / tiles, 16}));
    const int per = (chunks + S - 1) / S;
    *kslice = per * KC;
    return (chunks + per - 1) / per;
}

static long split_ws(int M, int N, int K) {
    int ks;
    const int S = split_plan(M, N, K, &ks);
    return S > 1 ? (long)S * M * N : 0;
}

static void free_batch(mfa_mlp *m) {
    for (float *p : m->d_h) if (p) (void)hipFree(p);
    for (float *p : m->d_z) if (p) (void)hipFree(p);
    m->d_h.clear(); m->d_z.clear();
    float *bufs[] = {m->d_dy, m->d_da, m->d_db, m->d_t, m->d_part, m->d_loss, m->d_ws};
    for (float *p : bufs) if (p) (void)hipFree(p);
    m->d_dy = m->d_da = m->d_db = m->d_t = m->d_part = m->d_loss = m->d_ws = nullptr;
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
    long ws = 1;
    for (int l = 0; l < L; ++l) {
        const int ni = m->w[l], no = m->w[l + 1];
        ws = std::max({ws, split_ws(rows, no, ni), split_ws(ni, no, rows), split_ws(rows, ni, no)});
    }
    HIPCHK(hipMalloc((void **)&m->d_ws, sizeof(float) * (size_t)ws));
    m->cap = rows;
    return 0;
}

// C (op)= A B; ws: the split-K workspace (at least split_ws floats), or null
// for one pass over K
static void gemm(hipStream_t st, const float *A, long sai, long sak, const float *B, long sbk, long sbn, float *C,
                 long ldc, int M, int N, int K, int epi, float *ws, const float *bias = nullptr, float *H = nullptr) {
    int ks = K;
    const int S = ws ? split_plan(M, N, K, &ks) : 1;
    if (S == 1) ks = K;
    dim3 g((M + GT - 1) / GT, (N + GT - 1) / GT, S);
    hipLaunchKernelGGL(k_gemm, g, dim3(256), 0, st, A, sai, sak, B, sbk, sbn, C, ldc, M, N, K, epi, bias, H, ks, ws);
    if (S > 1)
        hipLaunchKernelGGL(k_splitk_reduce, dim3(blocks((long)M * N)), dim3(256), 0, st, (const float *)ws, S, M, N, C,
                           ldc, epi, bias, H);
}


// forward of rows already in d_h[0]; output in d_h[L]
static void forward(mfa_mlp *m, int rows) {
    const int L = (int)m->w.size() - 1;
    for (int l = 0; l < L; ++l) {
        const int ni = m->w[l], no = m->w[l + 1];
        gemm(m->st, m->d_h[l], ni, 1, m->d_p + m->woff[l], no, 1, m->d_z[l], no, rows, no, ni,
             l < L - 1 ? EPI_SILU : EPI_BIAS, m->d_ws, m->d_p + m->boff[l], m->d_h[l + 1]);
    }
}

// loss and gradient (into d_g) of the rows in d_h[0] against targets d_t
static void backward(mfa_mlp *m, int rows) {
    const int L = (int)m->w.size() - 1;
    const long nout = (long)rows * m->w[L];
    hipLaunchKernelGGL(k_mse, dim3(blocks(nout)), dim3(256), 0, m->st, m->d_h[L], m->d_t, m->d_dy, m->d_part, nout,
                       1.0f / (float)nout);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, m->st, m->d_part, blocks(nout), m->d_loss, (const int *)nullptr);
    float *dz = m->d_dy;            // dL/dz of the current layer (rows x no)
    for (int l = L - 1; l >= 0; --l) {
        const int ni = m->w[l], no = m->w[l + 1];
        // dW_l = h_l^T dz  (ni x no), K = rows
        gemm(m->st, m->d_h[l], 1, ni, dz, no, 1, m->d_g + m->woff[l], no, ni, no, rows, EPI_STORE, m->d_ws);
        hipLaunchKernelGGL(k_colsum, dim3((no + 31) / 32), dim3(256), 0, m->st, dz, m->d_g + m->boff[l], rows, no);
        if (l == 0) break;
        // dh_l = dz W_l^T  (rows x ni), then through SiLU of layer l-1
        float *dh = dz == m->d_da ? m->d_db : m->d_da;
        gemm(m->st, dz, no, 1, m->d_p + m->woff[l], 1, no, dh, ni, rows, ni, no, EPI_STORE, m->d_ws);
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
    if (hipMalloc((void **)&m->d_step, sizeof(int)) != hipSuccess) {
        mfa_mlp_destroy(m); return fail(MFA_MLP_E_HIP, "step counter allocation");
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
    if (m->d_step) (void)hipFree(m->d_step);
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
    const long t0 = m->adam_t;
    const int zero = 0;
    HIPCHK(hipMemcpy(m->d_step, &zero, sizeof(int), hipMemcpyHostToDevice));
    // one training step: minibatch gather, forward, backward, loss into the
    // history at the step counter, Adam, counter + 1; every per-step value is
    // read from the counter on the device, so the step replays as a graph
    auto one_step = [&](hipStream_t st) {
        hipLaunchKernelGGL(k_gather, dim3(blocks(bt)), dim3(256), 0, st, (const float *)d_x, (const float *)d_t,
                           (const int *)d_ix, (const int *)m->d_step, bt, m->w[0], m->w[L], m->d_h[0], m->d_t);
        forward(m, bt);
        backward(m, bt);
        hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, st, m->d_part, blocks(nout), d_lh, (const int *)m->d_step);
        hipLaunchKernelGGL(k_adam, dim3(blocks(m->n_par)), dim3(256), 0, st, m->d_p, m->d_g, m->d_m1, m->d_m2,
                           m->n_par, lr, b1, b2, eps, (const int *)m->d_step, t0, steps, m->lr_final);
        hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(1), 0, st, m->d_step);
    };
    hipError_t e = hipSuccess;
    if (m->use_graph) {
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        e = hipStreamBeginCapture(m->st, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            one_step(m->st);
            e = hipStreamEndCapture(m->st, &g);
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        for (int it = 0; it < steps && e == hipSuccess; ++it) e = hipGraphLaunch(ge, m->st);
        if (e == hipSuccess) e = hipStreamSynchronize(m->st);
        if (ge) (void)hipGraphExecDestroy(ge);
        if (g) (void)hipGraphDestroy(g);
    } else {
        for (int it = 0; it < steps; ++it) one_step(m->st);
    }
    m->adam_t += steps;
    if (e != hipSuccess) {
        (void)hipFree(d_x); (void)hipFree(d_t); (void)hipFree(d_ix); (void)hipFree(d_lh);
        return fail(MFA_MLP_E_HIP, std::string("train: ") + hipGetErrorString(e));
    }
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(m->st);
    std::vector<float> lh(steps);
    if (e == hipSuccess) e = hipMemcpy(lh.data(), d_lh, sizeof(float) * (size_t)steps, hipMemcpyDeviceToHost);
    (void)hipFree(d_x); (void)hipFree(d_t); (void)hipFree(d_ix); (void)hipFree(d_lh);
    if (e != hipSuccess) return fail(MFA_MLP_E_HIP, std::string("train: ") + hipGetErrorString(e));
    if (loss_history)
        for (int it = 0; it < steps; ++it) loss_history[it] = lh[it] / (float)nout;
    return 0;
}

int mfa_mlp_set_lr_schedule(mfa_mlp *m, float lr_final) {
    if (!m || !(lr_final > 0.0f && lr_final <= 1.0f)) return fail(MFA_MLP_E_ARG, "set_lr_schedule: lr_final must be in (0, 1]");
    m->lr_final = lr_final;
    return 0;
}

int mfa_mlp_set_graph(mfa_mlp *m, int32_t on) {
    if (!m) return fail(MFA_MLP_E_ARG, "set_graph: null");
    m->use_graph = on != 0;
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
    float *dw = nullptr;
    const long nws = split_ws(M, N, K);
    if (nws > 0) HIPCHK(hipMalloc((void **)&dw, sizeof(float) * (size_t)nws));
    gemm(0, da, (long)K, 1L, db, (long)N, 1L, dc, (long)N, M, N, K, EPI_STORE, dw);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(c, dc, sizeof(float) * (size_t)M * N, hipMemcpyDeviceToHost));
    (void)hipFree(da); (void)hipFree(db); (void)hipFree(dc);
    if (dw) (void)hipFree(dw);
    return 0;
}

}  // extern "C"
