// wr_mlp.hip -- the gas-optics surrogate: layer-optics extraction (training targets from the plain
// FP64 path) and the BF16 MFMA kernels of the MLP (forward, backward, Adam).
//
// GEMM: C[M x N] = A[M x K] * B[K x N] with arbitrary element strides on A, B and C, operands
// staged through LDS as BF16, gfx950's v_mfma_f32_16x16x32_bf16 tiles, FP32 accumulation.  A
// 256-thread workgroup computes a 64 x 64 tile of C (4 waves, 2 x 2 MFMA tiles of 16 x 16 each); K
// advances 32 per step.  Each thread stages a run of 8 elements of A and of B along whichever of
// their dimensions is contiguous in memory (one 16- or 32-byte load when aligned and whole), so the
// staging reads are coalesced for row-major activations, transposed activations and weights alike.
// Split-K writes deterministic partial sums, reduced by k_splitk_reduce.
// Epilogues: bias, bias + SiLU (keeps the pre-activation), multiply by SiLU'(Z) (backward), plain.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wr_internal.h"

namespace {
using namespace wr;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

enum { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_SILU = 2, EPI_DSILU = 3 };
constexpr int GK = 32;      // K per step (one MFMA's depth)
constexpr int GLD = GK + 8; // LDS row stride in BF16 elements: 80 bytes, rows 16-byte aligned

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
    int k_per_split;         // K range of one split (grid.z), a multiple of GK
    long long split_stride;  // element offset between split partial outputs
};

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z)
{
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// cnt (0..8) elements p[j * stride] as BF16, zero beyond; one vector load when the run is whole,
// unit-stride and 16-byte aligned
__device__ __forceinline__ bf16x8 run8(const float *p, long long stride, int cnt)
{
    bf16x8 v;
    if (cnt == 8 && stride == 1 && ((uintptr_t)p & 15) == 0) {
        const float4 a = *(const float4 *)p, b = *(const float4 *)(p + 4);
        v[0] = (__bf16)a.x, v[1] = (__bf16)a.y, v[2] = (__bf16)a.z, v[3] = (__bf16)a.w;
        v[4] = (__bf16)b.x, v[5] = (__bf16)b.y, v[6] = (__bf16)b.z, v[7] = (__bf16)b.w;
        return v;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = (__bf16)(j < cnt ? p[j * stride] : 0.0f);
    return v;
}
__device__ __forceinline__ bf16x8 run8(const __bf16 *p, long long stride, int cnt)
{
    if (cnt == 8 && stride == 1 && ((uintptr_t)p & 15) == 0) return *(const bf16x8 *)p;
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = j < cnt ? p[j * stride] : (__bf16)0.0f;
    return v;
}

// Stage a 64 x GK block of an operand into S[row][k] (rows: M of A, N of B).  rows_contig: the row
// index is the unit-stride one (a run of 8 rows at one k); otherwise a run of 8 k at one row.
template <typename T>
__device__ __forceinline__ void stage(const T *X, long long ld_row, long long ld_k, int nrow, int r0, int kk,
                                      int k_end, bool rows_contig, __bf16 (*S)[GLD], int tid)
{
    if (rows_contig) {
        const int c = tid >> 3, r = (tid & 7) * 8;
        const int row = r0 + r, k = kk + c;
        const int cnt = k < k_end ? max(0, min(8, nrow - row)) : 0;
        const bf16x8 v = run8(X + (cnt > 0 ? (size_t)row * ld_row + (size_t)k * ld_k : 0), ld_row, cnt);
#pragma unroll
        for (int j = 0; j < 8; ++j) S[r + j][c] = v[j];
    } else {
        const int r = tid >> 2, c = (tid & 3) * 8;
        const int row = r0 + r, k = kk + c;
        const int cnt = row < nrow ? max(0, min(8, k_end - k)) : 0;
        *(bf16x8 *)&S[r][c] = run8(X + (cnt > 0 ? (size_t)row * ld_row + (size_t)k * ld_k : 0), ld_k, cnt);
    }
}

template <typename TA, typename TB, int EPI>
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g)
{
    __shared__ __attribute__((aligned(16))) __bf16 As[64][GLD];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[64][GLD];  // transposed: Bs[n][k]
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;  // M on x: up to 2^31 rows
    const int k_begin = blockIdx.z * g.k_per_split;
    const int k_end = min(g.K, k_begin + g.k_per_split);
    const TA *A = (const TA *)g.A;
    const TB *B = (const TB *)g.B;
    const bool a_rows = g.lda_k != 1 && g.lda_m == 1, b_rows = g.ldb_k != 1 && g.ldb_n == 1;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;  // wave's 32 x 32 sub-tile
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};

    for (int kk = k_begin; kk < k_end; kk += GK) {
        stage(A, g.lda_m, g.lda_k, g.M, m0, kk, k_end, a_rows, As, tid);
        stage(B, g.ldb_n, g.ldb_k, g.N, n0, kk, k_end, b_rows, Bs, tid);
        __syncthreads();
        // lane holds row (lane % 16) and k (lane / 16) * 8 + 0..7 of each operand tile
        const int row = lane & 15, kq = (lane >> 4) * 8;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const bf16x8 a = *(const bf16x8 *)&As[wm + i * 16 + row][kq];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const bf16x8 b = *(const bf16x8 *)&Bs[wn + j * 16 + row][kq];
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i][j], 0, 0, 0);
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

// ---- fused inference: the three layers of the network for 64 rows per workgroup ----------------
// H1 = SiLU(X W1^T + b1) and H2 = SiLU(H1 W2^T + b2) stay in LDS (BF16); Y = H2 W3^T + b3 is written
// output-major in BF16, Y[n * ldy + m], through an LDS transpose so that each store is a contiguous
// run of rows.  Same MFMA operands, K order, bias, SiLU and BF16 roundings as the three-GEMM chain,
// so the outputs equal the chain's bit for bit; the weights ([n][k] rows, BF16) are read straight
// into MFMA operands (L2-resident).
constexpr int FH = 128, FO = 504, FLD = FH + 8;

struct MlpW {
    const __bf16 *w1, *w2, *w3;  // [128][8], [128][128], [504][128]
    const float *b1, *b2, *b3;   // [128], [128], [504]
};

__device__ __forceinline__ bf16x8 wfrag(const __bf16 *w, int ld, int n, int nmax, int k)
{
    bf16x8 v;
    if (n < nmax) return *(const bf16x8 *)(w + (size_t)n * ld + k);
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = (__bf16)0.0f;
    return v;
}

__global__ void __launch_bounds__(256) k_mlp_fused(MlpW w, const float *__restrict__ X, long long M,
                                                   __bf16 *__restrict__ Y, long long ldy, int vec)
{
    __shared__ __attribute__((aligned(16))) __bf16 Hs[2][64][FLD];  // H1, H2
    __shared__ __attribute__((aligned(16))) __bf16 Ts[64][64 + 8];  // an output chunk, [n][m]
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const long long m0 = (long long)blockIdx.x * 64;
    const int row = lane & 15, kq = (lane >> 4) * 8, mq = (lane >> 4) * 4;
    const bf16x8 zero8 = {};

    // layer 1 (K = 8, zero-padded to 32): wave w owns hidden units [32 w, 32 w + 32)
    {
        f32x4 acc[4][2];
        bf16x8 b[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) b[j] = kq == 0 ? wfrag(w.w1, 8, wave * 32 + j * 16 + row, FH, 0) : zero8;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long long m = m0 + i * 16 + row;
            bf16x8 a = zero8;
            if (kq == 0 && m < M) {
                const float4 x0 = *(const float4 *)(X + (size_t)m * 8), x1 = *(const float4 *)(X + (size_t)m * 8 + 4);
                a[0] = (__bf16)x0.x, a[1] = (__bf16)x0.y, a[2] = (__bf16)x0.z, a[3] = (__bf16)x0.w;
                a[4] = (__bf16)x1.x, a[5] = (__bf16)x1.y, a[6] = (__bf16)x1.z, a[7] = (__bf16)x1.w;
            }
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], (f32x4){0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int n = wave * 32 + j * 16 + row;
#pragma unroll
                for (int q = 0; q < 4; ++q) Hs[0][i * 16 + mq + q][n] = (__bf16)silu(acc[i][j][q] + w.b1[n]);
            }
    }
    __syncthreads();
    // layer 2 (K = 128)
    {
        f32x4 acc[4][2];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < FH; ks += 32) {
            bf16x8 b[2];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = wfrag(w.w2, FH, wave * 32 + j * 16 + row, FH, ks + kq);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const bf16x8 a = *(const bf16x8 *)&Hs[0][i * 16 + row][ks + kq];
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], acc[i][j], 0, 0, 0);
            }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int n = wave * 32 + j * 16 + row;
#pragma unroll
                for (int q = 0; q < 4; ++q) Hs[1][i * 16 + mq + q][n] = (__bf16)silu(acc[i][j][q] + w.b2[n]);
            }
    }
    __syncthreads();
    // layer 3 (K = 128, 504 outputs in chunks of 64; wave w owns outputs [16 w, 16 w + 16) of a chunk)
    for (int c0 = 0; c0 < FO; c0 += 64) {
        const int nb = c0 + wave * 16;
        f32x4 acc[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < FH; ks += 32) {
            const bf16x8 b = wfrag(w.w3, FH, nb + row, FO, ks + kq);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const bf16x8 a = *(const bf16x8 *)&Hs[1][i * 16 + row][ks + kq];
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i], 0, 0, 0);
            }
        }
        const int n = nb + row;
        const float bias = n < FO ? w.b3[n] : 0.0f;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) Ts[wave * 16 + row][i * 16 + mq + q] = (__bf16)(acc[i][q] + bias);
        __syncthreads();
        {   // 64 outputs x 64 rows: a thread stores 16 consecutive rows of one output
            const int nl = tid >> 2, part = (tid & 3) * 16;
            const int no = c0 + nl;
            const long long m = m0 + part;
            if (no < FO && m < M) {
                __bf16 *dst = Y + (size_t)no * ldy + m;
                if (vec && m + 16 <= M) {
                    *(bf16x8 *)dst = *(const bf16x8 *)&Ts[nl][part];
                    *(bf16x8 *)(dst + 8) = *(const bf16x8 *)&Ts[nl][part + 8];
                } else {
                    for (int e = 0; e < 16 && m + e < M; ++e) dst[e] = Ts[nl][part + e];
                }
            }
        }
        __syncthreads();
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
    const int kps = ((K + splits - 1) / splits + GK - 1) / GK * GK;
    g.k_per_split = kps > 0 ? kps : GK;
    g.split_stride = split_stride;
    hipStream_t s = (hipStream_t)stream;
    if (a_bf16 && b_bf16) return gemm_dispatch<__bf16, __bf16>(epilogue, g, splits, s);
    if (a_bf16) return gemm_dispatch<__bf16, float>(epilogue, g, splits, s);
    if (b_bf16) return gemm_dispatch<float, __bf16>(epilogue, g, splits, s);
    return gemm_dispatch<float, float>(epilogue, g, splits, s);
}

int wr_mlp_forward(const wr_surrogate *w, int64_t M, const float *X, void *Y, int64_t ldy, void *stream)
{
    if (!w || !w->w1 || !w->w2 || !w->w3 || !w->b1 || !w->b2 || !w->b3 || !X || !Y || M < 0 || ldy < M ||
        (M + 63) / 64 > 0x7fffffffLL)
        return WR_ERR_ARG;
    // the weight rows and feature rows are read as 16-byte vectors
    const uintptr_t al = (uintptr_t)w->w1 | (uintptr_t)w->w2 | (uintptr_t)w->w3 | (uintptr_t)X;
    if (al & 15) return WR_ERR_ARG;
    const int vec = (ldy & 7) == 0 && ((uintptr_t)Y & 15) == 0;  // output runs stored as 16-byte vectors
    if (M == 0) return WR_OK;
    MlpW mw;
    mw.w1 = (const __bf16 *)w->w1, mw.w2 = (const __bf16 *)w->w2, mw.w3 = (const __bf16 *)w->w3;
    mw.b1 = w->b1, mw.b2 = w->b2, mw.b3 = w->b3;
    hipLaunchKernelGGL(k_mlp_fused, dim3((unsigned)((M + 63) / 64)), dim3(256), 0, (hipStream_t)stream, mw, X,
                       (long long)M, (__bf16 *)Y, (long long)ldy, vec);
    return hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
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
