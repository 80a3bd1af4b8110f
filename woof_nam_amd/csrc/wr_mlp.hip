// wr_mlp.hip -- the gas-optics surrogate: layer-optics extraction (training targets from the plain
// FP64 path) and the BF16 MFMA kernels of the MLP (forward, backward, Adam).
//
// GEMM: C[M x N] = A[M x K] * B[K x N] with arbitrary element strides on A, B and C, operands
// staged through LDS as BF16, v_mfma_f32_16x16x16_bf16 tiles, FP32 accumulation.  A 256-thread
// workgroup computes a 64 x 64 tile of C (4 waves, 2 x 2 MFMA tiles of 16 x 16 each); K advances
// 16 per step.  Split-K writes deterministic partial sums, reduced by k_splitk_reduce.
// Epilogues: bias, bias + SiLU (keeps the pre-activation), multiply by SiLU'(Z) (backward), plain.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wr_internal.h"

namespace {
using namespace wr;

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <typename T> __device__ __forceinline__ float ld(const T *p, size_t i) { return (float)p[i]; }

enum { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_SILU = 2, EPI_DSILU = 3 };

struct GemmArgs {
    int M, N, K;
    const void *A;
    long long lda_m, lda_k;
    const void *B;
    long long ldb_k, ldb_n;
    float *C;                // f32 output (or partials [split][M][N] with ldc)
    long long ldc_m, ldc_n;
    __bf16 *Cb;              // optional BF16 copy of the (activated) output, same strides
    const float *bias;       // [N]
    float *Zout;             // EPI_BIAS_SILU: pre-activation, same strides as C
    const float *Zin;        // EPI_DSILU: pre-activation the derivative is taken at, same strides as C
    int k_per_split;         // K range of one split (grid.z)
    long long split_stride;  // element offset between split partial outputs
};

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z)
{
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

template <typename TA, typename TB, int EPI>
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g)
{
    __shared__ __bf16 As[64][16 + 4];
    __shared__ __bf16 Bs[64][16 + 4];  // transposed: Bs[n][k]
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;  // M on x: up to 2^31 rows
    const int k_begin = blockIdx.z * g.k_per_split;
    const int k_end = min(g.K, k_begin + g.k_per_split);
    const TA *A = (const TA *)g.A;
    const TB *B = (const TB *)g.B;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;  // wave's 32 x 32 sub-tile
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};

    for (int kk = k_begin; kk < k_end; kk += 16) {
        // stage 64 x 16 of A and 16 x 64 of B (4 elements per thread each), zero outside the matrix
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int idx = tid * 4 + e;  // 0..1023
            const int r = idx >> 4, c = idx & 15;
            const int m = m0 + r, k = kk + c;
            float va = 0.f;
            if (m < g.M && k < k_end) va = ld(A, (size_t)m * g.lda_m + (size_t)k * g.lda_k);
            As[r][c] = (__bf16)va;
            const int n = n0 + r;
            float vb = 0.f;
            if (n < g.N && k < k_end) vb = ld(B, (size_t)k * g.ldb_k + (size_t)n * g.ldb_n);
            Bs[r][c] = (__bf16)vb;
        }
        __syncthreads();
        const int row = lane & 15, kq = (lane >> 4) * 4;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            bf16x4 a;
#pragma unroll
            for (int q = 0; q < 4; ++q) a[q] = As[wm + i * 16 + row][kq + q];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                bf16x4 b;
#pragma unroll
                for (int q = 0; q < 4; ++q) b[q] = Bs[wn + j * 16 + row][kq + q];
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // epilogue: lane holds D[4*(lane/16) + q][lane % 16] of each 16 x 16 tile
    float *C = g.C + (size_t)blockIdx.z * g.split_stride;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int m = m0 + wm + i * 16 + (lane >> 4) * 4 + q;
                const int n = n0 + wn + j * 16 + (lane & 15);
                if (m >= g.M || n >= g.N) continue;
                const size_t o = (size_t)m * g.ldc_m + (size_t)n * g.ldc_n;
                float v = acc[i][j][q];
                if (EPI == EPI_BIAS || EPI == EPI_BIAS_SILU) v += g.bias[n];
                if (EPI == EPI_BIAS_SILU) {
                    if (g.Zout) g.Zout[o] = v;
                    v = silu(v);
                }
                if (EPI == EPI_DSILU) v *= dsilu(g.Zin[o]);
                if (g.C) C[o] = v;
                if (g.Cb) g.Cb[o] = (__bf16)v;
            }
}

// sum of split partials: out[i] = sum_s part[s * stride + i]  (fixed order)
__global__ void k_splitk_reduce(const float *__restrict__ part, int splits, long long stride, float *__restrict__ out,
                                long long n)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int k = 0; k < splits; ++k) s += part[k * stride + i];
    out[i] = s;
}

// column sums db[n] = sum_m X[m * ld + n] (fixed order; one thread per column)
__global__ void k_colsum(const float *__restrict__ X, int M, int N, long long ld, float *__restrict__ out)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (int m = 0; m < M; ++m) s += X[(size_t)m * ld + n];
    out[n] = s;
}

// dY = 2 (Y - T) * scale; per-block partial sum of squared error
__global__ void k_mse(const float *__restrict__ Y, const float *__restrict__ T, long long n, float scale,
                      float *__restrict__ dY, float *__restrict__ part)
{
    __shared__ float red[256];
    float s = 0.f;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float d = Y[i] - T[i];
        dY[i] = 2.f * d * scale;
        s += d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Adam (Kingma & Ba 2015) on FP32 master weights, BF16 copy written alongside
__global__ void k_adam(float *__restrict__ w, __bf16 *__restrict__ wb, const float *__restrict__ grad,
                       float *__restrict__ m1, float *__restrict__ m2, long long n, float lr, float b1, float b2, float eps,
                       float bc1, float bc2)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gr = grad[i];
    const float a = b1 * m1[i] + (1.f - b1) * gr;
    const float b = b2 * m2[i] + (1.f - b2) * gr * gr;
    m1[i] = a;
    m2[i] = b;
    const float v = w[i] - lr * (a / bc1) / (sqrtf(b / bc2) + eps);
    w[i] = v;
    if (wb) wb[i] = (__bf16)v;
}

__global__ void k_to_bf16(const float *__restrict__ a, __bf16 *__restrict__ b, long long n)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) b[i] = (__bf16)a[i];
}

// ---- layer optics of the plain FP64 path: surrogate inputs and targets ------------------------
// Per radiation layer (surface-first k of column i), sample s = k * ncol + i:
//   feat[s][8]: ln p (hPa), T (K), ln q (kg/kg), ln o3 (kg/kg), dp (hPa), 0, 0, 0
//   targ[s][504]: ln tau_LW (140), Planck fraction (140), ln tau_SW (112, gas + Rayleigh), SW ssa (112)
constexpr int NOUT = 504;

template <int B> __device__ __forceinline__ void lw_all(const Tabs<double> *T, const LwLay<double> &L, float *y)
{
    if constexpr (B < 16) {
        double tau[LW_NG[B]], fr[LW_NG[B]];
        lw_optics<double, B>(T, L, tau, fr);
        for (int g = 0; g < LW_NG[B]; ++g) {
            y[LW_NS[B] + g] = (float)log_d(fmax(tau[g], 1.0e-10));
            y[140 + LW_NS[B] + g] = (float)fr[g];
        }
        lw_all<B + 1>(T, L, y);
    }
}
template <int B>
__device__ __forceinline__ void sw_all(const Tabs<double> *T, const SwLay<double> &L, int kk, int lt, float *y)
{
    if constexpr (B < 14) {
        double tg[SW_NG[B]], tr[SW_NG[B]];
        sw_optics<double, B>(T, L, kk, lt, tg, tr);
        for (int g = 0; g < SW_NG[B]; ++g) {
            const double t = tg[g] + tr[g];
            y[280 + SW_NS[B] + g] = (float)log_d(fmax(t, 1.0e-10));
            y[392 + SW_NS[B] + g] = (float)(t > 0.0 ? tr[g] / t : 0.0);
        }
        sw_all<B + 1>(T, L, kk, lt, y);
    }
}

__global__ void __launch_bounds__(64) k_optics(const Tabs<double> *__restrict__ T, ColIn in, RunPar par, float *feat,
                                               float *targ)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= in.ncol) return;
    const RadCol rc{in, i, in.nlay + 1};
    const int nrl = rc.nrl;
    int swlt = nrl;
    for (int k = 0; k < nrl; ++k)
        if (log_d(rc.plyr(nrl - 1 - k)) > 4.56) swlt = k + 1;
    for (int k = 0; k < nrl; ++k) {
        const int kin = nrl - 1 - k;
        const size_t s = (size_t)k * in.ncol + i;
        const double la = log_d(rc.plyr(kin));
        float *f = feat + s * 8;
        f[0] = (float)la;
        f[1] = (float)rc.tlyr(kin);
        f[2] = (float)log_d(rc.qlyr(kin));
        f[3] = (float)log_d(rc.olyr(kin));
        f[4] = (float)(rc.plvl(kin + 1) - rc.plvl(kin));
        f[5] = f[6] = f[7] = 0.f;
        if (!targ) continue;
        float *y = targ + s * NOUT;
        LwLay<double> L;
        lw_layer<double>(T, rc, par, kin, la, L);
        lw_all<0>(T, L, y);
        SwLay<double> S;
        sw_layer<double>(T, rc, par, kin, la, S);
        sw_all<0>(T, S, k + 1, swlt, y);
    }
}

template <typename TA, typename TB>
int gemm_dispatch(int epi, const GemmArgs &g, int splits, hipStream_t s)
{
    dim3 grid((g.M + 63) / 64, (g.N + 63) / 64, splits);
    switch (epi) {
    case EPI_STORE: hipLaunchKernelGGL((k_gemm<TA, TB, EPI_STORE>), grid, dim3(256), 0, s, g); break;
    case EPI_BIAS: hipLaunchKernelGGL((k_gemm<TA, TB, EPI_BIAS>), grid, dim3(256), 0, s, g); break;
    case EPI_BIAS_SILU: hipLaunchKernelGGL((k_gemm<TA, TB, EPI_BIAS_SILU>), grid, dim3(256), 0, s, g); break;
    case EPI_DSILU: hipLaunchKernelGGL((k_gemm<TA, TB, EPI_DSILU>), grid, dim3(256), 0, s, g); break;
    default: return WR_ERR_ARG;
    }
    return hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
}

}  // namespace

extern "C" {

int wr_mlp_gemm(int32_t M, int32_t N, int32_t K, const void *A, int32_t a_bf16, int64_t lda_m, int64_t lda_k,
                const void *B, int32_t b_bf16, int64_t ldb_k, int64_t ldb_n, float *C, void *Cb, int64_t ldc_m,
                int64_t ldc_n, const float *bias, float *Zout, const float *Zin, int32_t epilogue, int32_t splits,
                int64_t split_stride, void *stream)
{
    if (M < 0 || N < 0 || K < 0 || splits < 1 || splits > 65535 || N > 65535 * 64 || !A || !B || (!C && !Cb))
        return WR_ERR_ARG;
    if ((epilogue == EPI_BIAS || epilogue == EPI_BIAS_SILU) && !bias) return WR_ERR_ARG;
    if (epilogue == EPI_DSILU && !Zin) return WR_ERR_ARG;
    if (splits > 1 && (epilogue != EPI_STORE || !C || Cb)) return WR_ERR_ARG;
    if (M == 0 || N == 0) return WR_OK;
    GemmArgs g;
    g.M = M, g.N = N, g.K = K;
    g.A = A, g.lda_m = lda_m, g.lda_k = lda_k;
    g.B = B, g.ldb_k = ldb_k, g.ldb_n = ldb_n;
    g.C = C, g.Cb = (__bf16 *)Cb, g.ldc_m = ldc_m, g.ldc_n = ldc_n;
    g.bias = bias, g.Zout = Zout, g.Zin = Zin;
    const int kps = ((K + splits - 1) / splits + 15) / 16 * 16;
    g.k_per_split = kps > 0 ? kps : 16;
    g.split_stride = split_stride;
    hipStream_t s = (hipStream_t)stream;
    if (a_bf16 && b_bf16) return gemm_dispatch<__bf16, __bf16>(epilogue, g, splits, s);
    if (a_bf16) return gemm_dispatch<__bf16, float>(epilogue, g, splits, s);
    if (b_bf16) return gemm_dispatch<float, __bf16>(epilogue, g, splits, s);
    return gemm_dispatch<float, float>(epilogue, g, splits, s);
}

int wr_mlp_splitk_reduce(const float *part, int32_t splits, int64_t stride, float *out, int64_t n, void *stream)
{
    if (!part || !out || splits < 1 || n < 0) return WR_ERR_ARG;
    if (n == 0) return WR_OK;
    hipLaunchKernelGGL(k_splitk_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, part, splits,
                       (long long)stride, out, (long long)n);
    return hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
}

int wr_mlp_colsum(const float *X, int32_t M, int32_t N, int64_t ld, float *out, void *stream)
{
    if (!X || !out || M < 0 || N < 0) return WR_ERR_ARG;
    if (N == 0) return WR_OK;
    hipLaunchKernelGGL(k_colsum, dim3((N + 255) / 256), dim3(256), 0, (hipStream_t)stream, X, M, N, (long long)ld, out);
    return hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
}

int wr_mlp_mse(const float *Y, const float *T, int64_t n, float scale, float *dY, float *partial, int32_t nblocks,
               void *stream)
{
    if (!Y || !T || !dY || !partial || n < 0 || nblocks < 1) return WR_ERR_ARG;
    hipLaunchKernelGGL(k_mse, dim3(nblocks), dim3(256), 0, (hipStream_t)stream, Y, T, (long long)n, scale, dY, partial);
    return hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
}

int wr_mlp_adam(float *w, void *wb, const float *grad, float *m1, float *m2, int64_t n, float lr, float b1, float b2,
                float eps, int32_t step, void *stream)
{
    if (!w || !grad || !m1 || !m2 || n < 0 || step < 1) return WR_ERR_ARG;
    if (n == 0) return WR_OK;
    float bc1 = 1.f, bc2 = 1.f;
    for (int t = 0; t < step && t < 100000; ++t) bc1 *= b1, bc2 *= b2;
    bc1 = 1.f - bc1, bc2 = 1.f - bc2;
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, w, (__bf16 *)wb, grad,
                       m1, m2, (long long)n, lr, b1, b2, eps, bc1, bc2);
    return hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
}

int wr_mlp_to_bf16(const float *a, void *b, int64_t n, void *stream)
{
    if (!a || !b || n < 0) return WR_ERR_ARG;
    if (n == 0) return WR_OK;
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, a, (__bf16 *)b,
                       (long long)n);
    return hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
}

int wr_layer_optics(wr_ctx *c, const wr_columns *in, const wr_params *p, float *feat, float *targ, void *stream)
{
    if (!c || !in || !p || !feat || in->ncol < 0 || in->nlay < 2 || in->nlay + 1 > c->nl_max) return WR_ERR_ARG;
    if (in->ncol == 0) return WR_OK;
    if (!in->p_lev || !in->t_lay || !in->q_lay || !in->o3_lay || !in->cldf || !in->qc || !in->qi || !in->t_sfc)
        return WR_ERR_ARG;
    const ColIn ci = to_colin(in);
    RunPar par;
    par.mcica = p->mcica, par.iovr = p->iovr, par.seed = p->seed;
    par.solcon = p->solcon, par.re_liq = p->re_liq, par.re_ice = p->re_ice;
    for (int g = 0; g < 10; ++g) par.vmr[g] = p->vmr[g];
    int dev_prev = -1;
    (void)hipGetDevice(&dev_prev);
    if (dev_prev != c->device) (void)hipSetDevice(c->device);
    hipLaunchKernelGGL(k_optics, dim3((in->ncol + 63) / 64), dim3(64), 0, (hipStream_t)stream, c->tabs_d, ci, par, feat,
                       targ);
    const int rc = hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
    if (dev_prev >= 0 && dev_prev != c->device) (void)hipSetDevice(dev_prev);
    return rc;
}

}  // extern "C"
