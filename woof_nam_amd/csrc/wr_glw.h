// wr_glw.h -- g-point lanes for the LW band kernels of the McICA route (rtrnmc).
//
// Layout of the work: a lane per (column, g-point).  A 64-lane wave holds 64 / NGP columns, NGP = the
// band's g-point count rounded up to a power of two (lanes g >= NG idle).  Each lane sweeps its own
// g-point down and up the column; the g-points of one column read neighbouring k-table entries
// ([ind][g] rows), so a wave's table gathers touch 64 / NGP rows instead of 64.  A lane holds one
// radiance, one McICA word and one layer's optics, so the kernels fit several waves per SIMD where
// the lane-per-column kernels hold the whole band in registers at one or two.
//
// The per-layer column setup (lwrad prep + setcoef: indices, interpolation factors, column amounts)
// is the same for every band and every g-point: k_lwrec computes it once per tile and layer into a
// record ([field][layer][column], column fastest), which every g-point lane of every LW band loads
// (the lanes of a column read the same words).  Its operations are lw_layer's, so the FP64 record
// equals what the lane-per-column kernels compute in their sweeps, bit for bit.
//
// The band's level flux is the sum over its g-points in g order (the reference's order): each lane
// leaves its radiance in LDS and the column's first lane adds them one after the other, so the FP64
// results stay bit-identical to the oracle.
//
// Opt-in (WR_LW_GLANE=1 at context creation).  On the 954x835x60 grid these kernels took about twice
// the time of the lane-per-column ones, serialised (LW sum 441 against 226 ms in FP64, and the
// mixed call 276 against 159 ms): a lane has one dependent gather chain per layer where a
// lane-per-column kernel keeps a band's g-point gathers in flight together, each layer ends in a
// barrier and an in-order g-sum by one lane, and the Planck and cloud terms, which depend on the band
// only, are evaluated once per g-point.
#pragma once
#include "wr_kernels.h"

namespace wr {

// record fields (per surface-first layer k and tile column t): one packed index word, 25 values in
// the optics precision R and 4 FP64 values
constexpr int REC_R = 25, REC_D = 4;
template <typename R> struct LwRec {
    WR_G int32_t *idx;  // [k][n]
    WR_G R *r;          // [REC_R][k][n]
    WR_G double *d;     // [REC_D][k][n]
    int nk, n;
    __device__ __forceinline__ size_t at(int f, int k, int t) const { return ((size_t)f * nk + k) * n + t; }
};
template <typename R> constexpr size_t lwrec_bytes_per_col(int nk)
{
    return (size_t)nk * (sizeof(int32_t) + REC_R * sizeof(R) + REC_D * sizeof(double));
}

template <typename R>
__device__ __forceinline__ void lwrec_put(const LwRec<R> &rec, int k, int t, const LwLay<R> &L)
{
    rec.idx[(size_t)k * rec.n + t] = L.jp | (L.jt << 6) | (L.jt1 << 9) | (L.indself << 12) | (L.indfor << 16) |
                                     (L.indminor << 19) | (L.lower << 24);
    const R v[REC_R] = {L.fac00,     L.fac01,        L.fac10,     L.fac11,     L.selffac,   L.selffrac,  L.forfac,
                        L.forfrac,   L.minorfrac,    L.scaleminor, L.scaleminorn2, L.colamt[0], L.colamt[1], L.colamt[2],
                        L.colamt[3], L.colamt[4],    L.colamt[5], L.colamt[6], L.wx[0],     L.wx[1],     L.wx[2],
                        L.wx[3],     L.coldry,       L.colbrd,    L.pavel};
#pragma unroll
    for (int f = 0; f < REC_R; ++f) rec.r[rec.at(f, k, t)] = v[f];
    rec.d[rec.at(0, k, t)] = L.tavel;
    rec.d[rec.at(1, k, t)] = L.ct;
    rec.d[rec.at(2, k, t)] = L.clwp;
    rec.d[rec.at(3, k, t)] = L.ciwp;
}

template <typename R>
__device__ __forceinline__ void lwrec_get(const LwRec<R> &rec, int k, int t, LwLay<R> &L)
{
    const int w = rec.idx[(size_t)k * rec.n + t];
    L.jp = w & 63, L.jt = (w >> 6) & 7, L.jt1 = (w >> 9) & 7, L.indself = (w >> 12) & 15;
    L.indfor = (w >> 16) & 7, L.indminor = (w >> 19) & 31, L.lower = (w >> 24) & 1;
    R v[REC_R];
#pragma unroll
    for (int f = 0; f < REC_R; ++f) v[f] = rec.r[rec.at(f, k, t)];
    L.fac00 = v[0], L.fac01 = v[1], L.fac10 = v[2], L.fac11 = v[3], L.selffac = v[4], L.selffrac = v[5];
    L.forfac = v[6], L.forfrac = v[7], L.minorfrac = v[8], L.scaleminor = v[9], L.scaleminorn2 = v[10];
#pragma unroll
    for (int m = 0; m < 7; ++m) L.colamt[m] = v[11 + m];
#pragma unroll
    for (int m = 0; m < 4; ++m) L.wx[m] = v[18 + m];
    L.coldry = v[22], L.colbrd = v[23], L.pavel = v[24];
    L.tavel = rec.d[rec.at(0, k, t)];
    L.ct = rec.d[rec.at(1, k, t)];
    L.clwp = rec.d[rec.at(2, k, t)];
    L.ciwp = rec.d[rec.at(3, k, t)];
    L.delp = 0.0;
}

// the LW layer records of a tile: lw_layer of every layer of every column (one thread per column)
template <typename R>
__global__ void __launch_bounds__(64) k_lwrec(const Tabs<R> *__restrict__ T, ColIn in, RunPar par, int col0, int n,
                                              ColScal cs, LwRec<R> rec)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= valid_cols(in, col0, n)) return;
    const RadCol rc{in, col0 + t, in.nlay + 1};
    const int nrl = rc.nrl;
    for (int k = 0; k < nrl; ++k) {
        const int kin = nrl - 1 - k;
        LwLay<R> L;
        lw_layer<R>(T, rc.lay(kin), par, cs.la[(size_t)kin * n + t], L);
        lwrec_put<R>(rec, k, t, L);
    }
}

constexpr int pow2_at_least(int v)
{
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Waves per SIMD the g-lane kernels are compiled for (register budget 512 / waves).
#ifndef WR_GLW_WAVES
#define WR_GLW_WAVES 4
#endif

// g-lane LW band kernel, McICA route.  Block: 64 lanes = 64 / NGP columns.
template <typename R, int B, bool SURR = false>
__global__ void __launch_bounds__(64, WR_GLW_WAVES) k_glw(const Tabs<R> *__restrict__ T, ColIn in, RunPar par, int col0,
                                                          int n, ColScal cs, AccT<R> acc, Surr sg, LwRec<R> rec)
{
    constexpr int NG = LW_NG[B];
    constexpr int NGP = pow2_at_least(NG);
    constexpr int CPW = 64 / NGP;
    constexpr int NS = LW_NS[B];
    __shared__ R red[64];
    const int lane = threadIdx.x;
    const int g = lane % NGP, cl = lane / NGP;
    const int t = blockIdx.x * CPW + cl;
    const int nv = valid_cols(in, col0, n);
    // lanes beyond the band's g-points or the tile's columns take part in the wave's LDS exchange only
    const bool col_ok = t < nv;
    const bool on = col_ok && g < NG;
    const bool lead = col_ok && g == 0;  // adds the column's g-points and writes its level fluxes
    const int tt = col_ok ? t : 0;
    const int i = col0 + tt;
    const RadCol rc{in, i, in.nlay + 1};
    const int nrl = rc.nrl;
    const int lcf1 = col_ok ? cs.lcf1[tt] : 0;
    const int table = SURR ? (col_ok ? sg.table(tt, B) : 1) : 1;
    const double pwv = col_ok ? cs.pwv[tt] : 0.0;
    const R sdif = (R)((B == 0 || B == 3 || B == 9) ? 1.66 : fmin(1.80, fmax(1.50, LW_A0[B] + LW_A1[B] * exp_d(LW_A2[B] * pwv))));
    const double sfe = col_ok ? in.emis[i] : 1.0;
    const R semis = (R)((sfe > EPS && sfe <= 1.0) ? sfe : 1.0);
    const uint64_t col = in.col_id ? in.col_id[i] : (uint64_t)i;
    const uint64_t seed = in.seed_dev ? *in.seed_dev : par.seed;
    const int gg = on ? g : 0;
    const uint64_t mb = on ? mcica_g(col, seed, 0, NS + gg, nrl, par.iovr,
                                     [&](int k) { return lcf1 ? rc.cfrac(nrl - 1 - k) : 0.0; })
                           : 0;
    // the column's g-sum of the lanes' values, in g order, by its first lane
    auto gsum_in_order = [&](R v) {
        red[lane] = on ? v : (R)0;
        __syncthreads();  // one wave per block: orders the LDS write before the reads
        R s = (R)0;
        if (lead) {
#pragma unroll
            for (int j = 0; j < NG; ++j) s = s + red[cl * NGP + j];
        }
        __syncthreads();  // the reads are done before the next exchange writes
        return s;
    };

    if (lead) acc.lw_dn[(size_t)nrl * n + t] = (R)0;
    R rad = (R)0;
    double tz_top = col_ok ? cs.tz[tt] : 0.0;
    R fr_sfc = (R)0;
    for (int kin = 0; kin < nrl; ++kin) {
        const int k = nrl - 1 - kin;
        const double tz_bot = col_ok ? cs.tz[(size_t)(kin + 1) * n + tt] : 0.0;
        R tau = (R)0, fr = (R)0, s_in = (R)0;
        if (col_ok) {
            LwLay<R> L;
            lwrec_get<R>(rec, k, tt, L);
            if (SURR && !table)
                surr_lw<R, B, 1>(sg, (size_t)k * n + tt, &tau, &fr, gg);
            else
                lw_optics<R, B, 1>(T, L, &tau, &fr, gg);
            const R blay = planck<R>(T, B, L.tavel);
            const R dplnkd = planck<R>(T, B, tz_bot) - blay;
            const R taucld = lw_taucld_l<R>(T, B, lcf1, L, rc, kin, par);
            const R taer = (R)rc.aer_lw_tau(B, kin);  // tautot = taug + tauaer
            const GasT<R> q = lw_gas<R>(T, sdif, tau + taer);
            const R plfrac = fr;
            const R bbdgas = plfrac * (blay + dplnkd * q.gasfac);
            const R gassrcd = bbdgas * q.atrgas;
            const R clfm = ((mb >> k) & 1) ? (R)1 : (R)0;
            if (clfm >= (R)EPS) {
                const R odcld = sdif * taucld;
                const R efclrfr = (R)1 - ((R)1 - M<R>::exp(-odcld)) * clfm;
                R totfac, atrtot;
                lw_tot<R>(T, q.odepth + odcld, totfac, atrtot);
                const R bbdtot = plfrac * (blay + dplnkd * totfac);
                const R totsrcd = bbdtot * atrtot;
                rad = rad * q.trng * efclrfr + gassrcd + clfm * (totsrcd - gassrcd);
            } else {
                rad = rad * q.trng + gassrcd;
            }
            s_in = rad;
            fr_sfc = fr;
        }
        const R s = gsum_in_order(s_in);
        if (lead) acc.lw_dn[(size_t)k * n + t] = s;
        tz_top = tz_bot;
    }
    // surface emission and reflection (fr of the surface layer, as the reference's rtrnmc)
    const R reflct = (R)1 - semis;
    R s0 = (R)0;
    if (col_ok) {
        const R pk0 = planck<R>(T, B, rc.tskn());
        const R rad0 = semis * fr_sfc * pk0;
        rad = rad0 + reflct * rad;
        s0 = rad;
    }
    s0 = gsum_in_order(s0);
    if (lead) acc.lw_up[t] = s0;
    for (int kin = nrl - 1; kin >= 0; --kin) {
        const int k = nrl - 1 - kin;
        const double tzt = col_ok ? cs.tz[(size_t)kin * n + tt] : 0.0;
        R s_in = (R)0;
        if (col_ok) {
            LwLay<R> L;
            lwrec_get<R>(rec, k, tt, L);
            R tau = (R)0, fr = (R)0;
            if (SURR && !table)
                surr_lw<R, B, 1>(sg, (size_t)k * n + tt, &tau, &fr, gg);
            else
                lw_optics<R, B, 1>(T, L, &tau, &fr, gg);
            const R blay = planck<R>(T, B, L.tavel);
            const R dplnku = planck<R>(T, B, tzt) - blay;
            const R taucld = lw_taucld_l<R>(T, B, lcf1, L, rc, kin, par);
            const R taer = (R)rc.aer_lw_tau(B, kin);  // tautot = taug + tauaer
            const GasT<R> q = lw_gas<R>(T, sdif, tau + taer);
            const R plfrac = fr;
            const R bbugas = plfrac * (blay + dplnku * q.gasfac);
            const R gassrcu = bbugas * q.atrgas;
            const R clfm = ((mb >> k) & 1) ? (R)1 : (R)0;
            if (clfm > (R)EPS) {
                const R odcld = sdif * taucld;
                const R efclrfr = (R)1 - ((R)1 - M<R>::exp(-odcld)) * clfm;
                R totfac, atrtot;
                lw_tot<R>(T, q.odepth + odcld, totfac, atrtot);
                const R bbutot = plfrac * (blay + dplnku * totfac);
                const R totsrcu = bbutot * atrtot;
                rad = rad * q.trng * efclrfr + gassrcu + clfm * (totsrcu - gassrcu);
            } else {
                rad = rad * q.trng + gassrcu;
            }
            s_in = rad;
        }
        const R s = gsum_in_order(s_in);
        if (lead) acc.lw_up[(size_t)(k + 1) * n + t] = s;
    }
}

}  // namespace wr
