// wr_surr.hip -- the surrogate path: per tile, layer features -> BF16 MFMA MLP -> learned gas optics
// read by the mixed-precision band kernels (columns outside the training envelope use the tables).
#include "wr_launch.h"

extern "C" int wr_mlp_gemm(int32_t M, int32_t N, int32_t K, const void *A, int32_t a_bf16, int64_t lda_m,
                           int64_t lda_k, const void *B, int32_t b_bf16, int64_t ldb_k, int64_t ldb_n, float *C,
                           void *Cb, int64_t ldc_m, int64_t ldc_n, const float *bias, float *Zout, const float *Zin,
                           int32_t epilogue, int32_t splits, int64_t split_stride, void *stream);

namespace {
using namespace wr;

// normalised features X[k * n + t][8] of every radiation layer of tile column t, and the column's
// table-optics flag (1 when any layer lies outside the training min/max of any feature)
__global__ void __launch_bounds__(64) k_features(ColIn in, int col0, int n, wr_surrogate sur, float *__restrict__ X,
                                                 int *__restrict__ flag)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = col0 + t;
    const RadCol rc{in, i, in.nlay + 1};
    const int nrl = rc.nrl;
    int out = 0;
    for (int k = 0; k < nrl; ++k) {
        const int kin = nrl - 1 - k;
        float f[5];
        f[0] = (float)log_d(rc.plyr(kin));
        f[1] = (float)rc.tlyr(kin);
        f[2] = (float)log_d(rc.qlyr(kin));
        f[3] = (float)log_d(rc.olyr(kin));
        f[4] = (float)(rc.plvl(kin + 1) - rc.plvl(kin));
        float *x = X + ((size_t)k * n + t) * 8;
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            if (!(f[j] >= sur.fmin[j] && f[j] <= sur.fmax[j])) out = 1;
            x[j] = (f[j] - sur.fmean[j]) / sur.fstd[j];
        }
        x[5] = x[6] = x[7] = 0.f;
    }
    flag[t] = out;
}

__global__ void k_copy_flag(const int *__restrict__ f, int n, int32_t *__restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = f[t];
}
}  // namespace

int wr_launch_surr(wr_ctx *c, const ColIn &in, const RunPar &par, const OutPtrs &out, int col0, int n, int32_t *flag_out,
                   hipStream_t s)
{
    const int nrl = in.nlay + 1;
    const long long S = (long long)nrl * n;
    tbeg(c, KID_FEAT, s);
    hipLaunchKernelGGL(k_features, dim3((n + 63) / 64), dim3(64), 0, s, in, col0, n, c->sur, c->sx, c->sflag);
    tend(c, s);
    if (hipGetLastError() != hipSuccess) return WR_ERR_HIP;
    const wr_surrogate &w = c->sur;
    tbeg(c, KID_MLP, s);
    int rc = wr_mlp_gemm((int)S, 128, 8, c->sx, 0, 8, 1, w.w1, 1, 1, 8, nullptr, c->sh1, 128, 1, w.b1, nullptr, nullptr, 2,
                         1, 0, s);
    if (rc == WR_OK)
        rc = wr_mlp_gemm((int)S, 128, 128, c->sh1, 1, 128, 1, w.w2, 1, 1, 128, nullptr, c->sh2, 128, 1, w.b2, nullptr,
                         nullptr, 2, 1, 0, s);
    if (rc == WR_OK)  // outputs stored output-major, Y[j * S + s], for coalesced reads in the band kernels
        rc = wr_mlp_gemm((int)S, 504, 128, c->sh2, 1, 128, 1, w.w3, 1, 1, 128, c->sy, nullptr, 1, S, w.b3, nullptr,
                         nullptr, 1, 1, 0, s);
    tend(c, s);
    if (rc != WR_OK) return rc;
    if (flag_out) {
        hipLaunchKernelGGL(k_copy_flag, dim3((n + 255) / 256), dim3(256), 0, s, (const int *)c->sflag, n, flag_out);
        if (hipGetLastError() != hipSuccess) return WR_ERR_HIP;
    }
    Surr sg;
    sg.Y = gptr(c->sy), sg.tmean = gptr(w.tmean), sg.tstd = gptr(w.tstd), sg.flag = gptr(c->sflag), sg.S = S;
    return launch_tile<float, true>(c, c->tabs_f, c->sflux_f, in, par, out, col0, n, -1, s, sg);
}
