// wr_aer.hip -- climatological aerosol optics of the radiation call (wr_aer_optics): setaer's prep
// (phys_rrtm_setaer.cu:86-161), grrad setup's layer relative humidity (phys_rrtm_grrad_setup.cu:
// 156-161) and aer_property with radclimaer (phys_rrtm_aer_property.cu:104-434), in the NAM's setting
// (layers from the top, ivflip = 0; 14 SW and 16 LW bands; no volcanic aerosol).
//
// One workgroup of 8 waves per 64 columns, a column per lane:
//   1. the 8 waves split the layers: hypsometric thickness and relative-humidity bracket of each
//      layer (the saturation vapour pressure is most of the arithmetic) into LDS;
//   2. wave 0 walks the column from the surface up: level heights, the aerosol domain of each layer
//      and its effective thickness under the domain's exponential profile, into LDS;
//   3. the 8 waves split the 30 bands: each streams its band down the column, writes single-scattering
//      albedo and asymmetry at once and the optical depth one layer late, after the smoothing across
//      domain boundaries (which only moves optical depth between a layer and the one above) is final.
// The humidity-independent part of the mixing-layer sum is the same in every layer of a column, so it
// is formed once per band: the sums keep the reference's order, species 1..6 first.
// FP64 on the float32 tables with the library's exp/log/pow and no contraction: every value agrees with
// the CPU oracle bit for bit.  Outputs: float32 [3][bands][nlay+1][ld], the column call's aerosol inputs.
#include <hip/hip_runtime.h>

#include "../../include/woofrad.h"
#include "wr_dev.h"

namespace {
using namespace wr;

constexpr int A_NDM = 5, A_NAE = 7, A_IMXAE = 72, A_JMXAE = 37, A_NXC = 5, A_NRHLEV = 8, A_NCM1 = 6, A_NCM = 10;
constexpr int A_NSW = 14, A_NLW = 16, A_NB = A_NSW + A_NLW, A_NL = 64, A_WAVES = 8;
// physcons of setaer (phys_rrtm_setaer.cu:38-43) and grrad setup (phys_rrtm_grrad_setup.cu:3-24)
constexpr double A_PI = 3.1415926535897931, A_RD = 287.05, A_RV = 461.5, A_G = 9.80665;
constexpr double A_FVIRT = A_RV / A_RD - 1.0, A_EPS = A_RD / A_RV, A_EPSM1 = A_RD / A_RV - 1.0;
constexpr double A_QMIN = 1.0e-10, A_QME5 = 1.0e-7, A_PRSMIN = 1.0e-6;

// GFS fpvsx (phys_rrtm_grrad_setup.cu:39-60): saturation vapour pressure (Pa) over liquid at and above
// the triple point, over ice 20 K and more below it, blended linearly between
__device__ double fpvsx_gfs(double t)
{
#pragma clang fp contract(off)
    const double ttp = 2.7316e+2, psat = 6.1078e+2, cvap = 1.8460e+3, cliq = 4.1855e+3, csol = 2.1060e+3,
                 hvap = 2.5000e+6, hfus = 3.3358e+5;
    const double tliq = ttp, tice = ttp - 20.0;
    const double dldtl = cvap - cliq, heatl = hvap;
    const double xponal = -dldtl / A_RV, xponbl = -dldtl / A_RV + heatl / (A_RV * ttp);
    const double dldti = cvap - csol, heati = hvap + hfus;
    const double xponai = -dldti / A_RV, xponbi = -dldti / A_RV + heati / (A_RV * ttp);
    const double tr = ttp / t;
    if (t >= tliq) return psat * pow_d(tr, xponal) * exp_d(xponbl * (1.0 - tr));
    if (t < tice) return psat * pow_d(tr, xponai) * exp_d(xponbi * (1.0 - tr));
    const double w = (t - tice) / (tliq - tice);
    const double pvl = psat * pow_d(tr, xponal) * exp_d(xponbl * (1.0 - tr));
    const double pvi = psat * pow_d(tr, xponai) * exp_d(xponbi * (1.0 - tr));
    return w * pvl + (1.0 - w) * pvi;
}

// GFS fpvs (phys_rrtm_grrad_setup.cu:62-71): the 7501-node fpvsx table over 180..330 K, interpolated
// linearly between its two nodes (evaluated here), the index clamped to the table
__device__ double fpvs_gfs(double t)
{
#pragma clang fp contract(off)
    const double xmin = 180.0, nx = 7501.0, xinc = (330.0 - 180.0) / (nx - 1.0);
    const double c2 = 1.0 / xinc, c1 = 1.0 - xmin * c2;
    const double xj = fmin(fmax(c1 + c2 * t, 1.0), nx);
    const int jx = (int)fmin(xj, nx - 1.0);
    const double f0 = fpvsx_gfs(xmin + (double)(jx - 1) * xinc), f1 = fpvsx_gfs(xmin + (double)jx * xinc);
    return f0 + (xj - (double)jx) * (f1 - f0);
}

struct ATab {  // wr_aer_tables with global-address-space pointers (kernel parameters are structs)
    const WR_G float *haer, *prsref, *sigref, *rhlev, *extrhi, *scarhi, *ssarhi, *asyrhi, *extrhd, *scarhd,
        *ssarhd, *asyrhd, *extstra, *cmixg, *denng;
    const WR_G int32_t *idxcg, *kprfg;
    // 1-based accessors, the reference's Fortran indices
    __device__ double haer_(int idm, int kp) const { return (double)haer[(kp - 1) * A_NDM + idm - 1]; }
    __device__ double prsref_(int idm, int kp) const { return (double)prsref[(kp - 1) * A_NDM + idm - 1]; }
    __device__ double sigref_(int idm, int kp) const { return (double)sigref[(kp - 1) * A_NDM + idm - 1]; }
    __device__ double rhlev_(int ir) const { return (double)rhlev[ir - 1]; }
    __device__ int kprf(int i, int j) const
    {
        const int k = kprfg[(j - 1) * A_IMXAE + i - 1];
        return k < 1 ? 1 : (k > A_NAE ? A_NAE : k);  // a valid table holds 1..7
    }
};
struct ACols {
    const WR_G double *p_lev, *t_lay, *q_lay;
    const WR_G float *xlon, *xlat, *slmsk;
    WR_G float *sw, *lw;
    int64_t ld;
    int ncol, nlay;
};

__device__ __forceinline__ double rhi(const WR_G float *a, int ic, int ib) { return (double)a[(ib - 1) * A_NCM1 + ic - 1]; }
__device__ __forceinline__ double rhd(const WR_G float *a, int ir, int ic1, int ib)
{
    return (double)a[((ib - 1) * 4 + ic1 - 1) * A_NRHLEV + ir - 1];
}

// level l of the radiation column (mb, setaer's prsi from the top): the added top level at PRSMIN
__device__ __forceinline__ double level(const ACols &g, int c, int l)
{
    if (l == 0) return A_PRSMIN;
    const double v = 10.0 * (g.p_lev[(size_t)(l - 1) * g.ncol + c] * 1.0e-3);
    return (l == 1 && v <= A_PRSMIN) ? 2.0 * A_PRSMIN : v;
}

struct Cell {  // aer_property's grid cell around the column (phys_rrtm_aer_property.cu:133-202)
    int i1, i2, j1, j2, kp, ii;
    double wi, wj, h1;
};

__device__ Cell cell_of(const ATab &T, double xlon, double xlat, double slm)
{
#pragma clang fp contract(off)
    Cell e;
    const double rdg = 180.0 / A_PI;
    double alon = xlon * rdg;
    if (alon < 0.0) alon = alon + 360.0;
    double alat = xlat * rdg;
    if (alat < -90.0) alat = -90.0;
    if (alat > 90.0) alat = 90.0;
    const double dltg = 360.0 / (double)A_IMXAE, hdlt = 0.5 * dltg, rdlt = 1.0 / dltg;
    int i3 = 1;
    double dtmp = alon;
    if (dtmp < 0.0) dtmp = 0.0;
    while (dtmp > dltg && i3 < A_IMXAE) {
        ++i3;
        dtmp = alon - dltg * (double)(i3 - 1);
    }
    e.i1 = i3, e.i2 = i3 % A_IMXAE + 1, e.wi = dtmp * rdlt;
    const int kpi = (dtmp <= hdlt) ? i3 : e.i2;
    int j3 = 1;
    double dtl = 90.0 - alat;
    while (dtl > dltg && j3 < A_JMXAE) {
        ++j3;
        dtl = (90.0 - dltg * (double)(j3 - 1)) - alat;
    }
    e.j1 = j3, e.j2 = (j3 + 1 > A_JMXAE) ? A_JMXAE : j3 + 1, e.wj = dtl * rdlt;
    const int kpj = (dtl <= hdlt) ? j3 : e.j2;
    int kp = T.kprf(kpi, kpj);
    const int kpa = imax(imax(T.kprf(e.i1, e.j1), T.kprf(e.i1, e.j2)), imax(T.kprf(e.i2, e.j1), T.kprf(e.i2, e.j2)));
    double h1 = T.haer_(1, kp);
    int ii = 1;
    if (kp != kpa) {
        if (kpa == 6) {  // ocean profile with a mineral overlay
            ii = 2;
            if (slm > 0.0) {
                kp = 7;
                h1 = 0.5 * (T.haer_(1, 6) + T.haer_(1, 7));
            } else {
                kp = kpa;
                h1 = T.haer_(1, 6);
            }
        } else if (kpa == 7) {  // land profile with a mineral overlay
            ii = 2;
            if (slm <= 0.0) {
                kp = 6;
                h1 = 0.5 * (T.haer_(1, 6) + T.haer_(1, 7));
            } else {
                kp = kpa;
                h1 = T.haer_(1, 7);
            }
        } else {
            h1 = T.haer_(1, kpa);
            kp = kpa;
        }
    }
    e.kp = kp, e.ii = ii, e.h1 = h1;
    return e;
}

__global__ void __launch_bounds__(512) k_aer_optics(ATab T, ACols g)
{
#pragma clang fp contract(off)
    __shared__ double s_delz[A_NL][64];  // thickness, then effective thickness (km)
    __shared__ double s_rdrh[A_NL][64];  // humidity interpolation weight
    __shared__ unsigned char s_idom[A_NL][64], s_ih[A_NL][64];  // domain; humidity levels ih1 | ih2 << 4 (<= 136)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c0 = blockIdx.x * 64 + lane;
    const bool valid = c0 < g.ncol;
    const int c = valid ? c0 : g.ncol - 1;  // idle lanes redo the last column and store nothing
    const int nrl = g.nlay + 1;
    const size_t nc = (size_t)g.ncol;

    // 1. thickness and humidity bracket, layers split over the waves
    for (int l = wave; l < nrl; l += A_WAVES) {
        const size_t s = (size_t)(l == 0 ? 0 : l - 1) * nc + c;  // the added top layer takes layer 1's state
        const double t = g.t_lay[s], q = g.q_lay[s];
        const double ql = fmax(A_QME5, q);
        const double pl_cb = 0.5 * (g.p_lev[s] + g.p_lev[s + nc]) * 1.0e-3;
        const double es = fmin(pl_cb, 0.001 * fpvs_gfs(t));
        const double qs = fmax(A_QMIN, A_EPS * es / (pl_cb + A_EPSM1 * es));
        const double rh = fmax(0.0, fmin(1.0, fmax(A_QMIN, q) / qs));
        const double lnk = (l == 0) ? log_d(0.5 * level(g, c, 1)) : log_d(level(g, c, l));
        const double lnk1 = log_d(level(g, c, l + 1));
        double dz = (0.001 * A_RD / A_G) * (lnk1 - lnk) * t * (1.0 + A_FVIRT * ql);
        if (l == 0) dz = 2.0 * dz;
        s_delz[l][lane] = dz;
        int ih2 = 1;
        while (ih2 <= A_NRHLEV && rh > T.rhlev_(ih2)) ++ih2;
        const int ih1 = imax(1, ih2 - 1);
        ih2 = imin(A_NRHLEV, ih2);
        const double drh0 = T.rhlev_(ih2) - T.rhlev_(ih1), drh1 = rh - T.rhlev_(ih1);
        s_rdrh[l][lane] = (ih1 == ih2) ? 0.0 : drh1 / drh0;
        s_ih[l][lane] = (unsigned char)(ih1 | ih2 << 4);
    }
    __syncthreads();

    const Cell e = cell_of(T, (double)g.xlon[c], (double)g.xlat[c], (double)g.slmsk[c]);
    // 2. heights, domains and effective thickness, from the surface up (phys_rrtm_aer_property.cu:251-267)
    if (wave == 0) {
        const double rps = 1.0 / level(g, c, nrl);
        int ii = 1;
        double hz_lo = 0.0;  // height of the layer's lower level
        for (int l = nrl - 1; l >= 0; --l) {
            const double dz = s_delz[l][lane];
            const double hz_hi = hz_lo + dz;
            if (ii <= A_NDM && level(g, c, l) * rps < T.sigref_(ii, e.kp)) {
                ii = ii + 1;
                if (ii == 2 && T.prsref_(2, e.kp) == T.prsref_(3, e.kp)) ii = 3;
            }
            s_idom[l][lane] = (unsigned char)ii;
            const double tmp1 = (ii > 1) ? ((ii <= A_NDM) ? T.haer_(ii, e.kp) : 0.0) : e.h1;
            if (tmp1 > 0.0) {
                const double tmp2 = 1.0 / tmp1;
                s_delz[l][lane] = tmp1 * (exp_d(-hz_lo * tmp2) - exp_d(-hz_hi * tmp2));
            }
            hz_lo = hz_hi;
        }
    }
    // the column's number densities and species mixing ratios (bilinear over the cell's corners)
    const double w11 = (1.0 - e.wi) * (1.0 - e.wj), w12 = (1.0 - e.wi) * e.wj, w21 = e.wi * (1.0 - e.wj);
    const double w22 = e.wi * e.wj;
    auto cm_at = [&](int m, int i, int j) { return (double)T.cmixg[((j - 1) * A_IMXAE + i - 1) * A_NXC + m - 1]; };
    auto ic_at = [&](int m, int i, int j) { return T.idxcg[((j - 1) * A_IMXAE + i - 1) * A_NXC + m - 1]; };
    auto dn_at = [&](int m, int i, int j) { return (double)T.denng[((j - 1) * A_IMXAE + i - 1) * 2 + m - 1]; };
    double denn1 = 0.0, denn2 = 0.0;
    denn1 = w11 * dn_at(1, e.i1, e.j1) + w12 * dn_at(1, e.i1, e.j2) + w21 * dn_at(1, e.i2, e.j1) +
            w22 * dn_at(1, e.i2, e.j2);
    if (e.ii >= 2)
        denn2 = w11 * dn_at(2, e.i1, e.j1) + w12 * dn_at(2, e.i1, e.j2) + w21 * dn_at(2, e.i2, e.j1) +
                w22 * dn_at(2, e.i2, e.j2);
    double cmix[A_NCM + 1];
    for (int n = 0; n <= A_NCM; ++n) cmix[n] = 0.0;
    for (int m = 1; m <= A_NXC; ++m) {
        const int i_[4] = {e.i1, e.i1, e.i2, e.i2}, j_[4] = {e.j1, e.j2, e.j1, e.j2};
        const double w_[4] = {w11, w12, w21, w22};
        for (int q = 0; q < 4; ++q) {
            const int ic = ic_at(m, i_[q], j_[q]);
            if (ic > 0 && ic <= A_NCM) {
#pragma unroll
                for (int n = 1; n <= A_NCM; ++n)  // register-resident: select, no dynamic index
                    if (n == ic) cmix[n] = cmix[n] + w_[q] * cm_at(m, i_[q], j_[q]);
            }
        }
    }
    __syncthreads();

    // 3. bands split over the waves; each streams down the column
    const double crt1 = 30.0, crt2 = 0.03333;
    for (int b = wave; b < A_NB; b += A_WAVES) {
        const int ib = b + 1;
        const bool sw = ib <= A_NSW;
        const int nb = sw ? A_NSW : A_NLW, bb = sw ? b : b - A_NSW;
        const size_t fs = (size_t)nb * nrl * (size_t)g.ld;  // field stride
        WR_G float *ot = (sw ? g.sw : g.lw) + (size_t)bb * nrl * (size_t)g.ld + c0;
        // mixing layer, humidity-independent species (the same in every layer)
        double e6 = 0.0, s6 = 0.0, ss6 = 0.0, as6 = 0.0;
#pragma unroll
        for (int icmp = 1; icmp <= A_NCM1; ++icmp) {
            const double cm = cmix[icmp];
            if (cm > 0.0) {
                const double ex = rhi(T.extrhi, icmp, ib), sc = rhi(T.scarhi, icmp, ib);
                const double ss = rhi(T.ssarhi, icmp, ib), as = rhi(T.asyrhi, icmp, ib);
                e6 = e6 + cm * ex;
                s6 = s6 + cm * sc;
                ss6 = ss6 + cm * ss * ex;
                as6 = as6 + cm * as * sc;
            }
        }
        // free troposphere: species 1 and 2 (humidity-independent) part
        const double ex01 = rhi(T.extrhi, 1, ib), sc01 = rhi(T.scarhi, 1, ib), ss01 = rhi(T.ssarhi, 1, ib);
        const double as01 = rhi(T.asyrhi, 1, ib), ex02 = rhi(T.extrhi, 2, ib), sc02 = rhi(T.scarhi, 2, ib);
        const double ss02 = rhi(T.ssarhi, 2, ib), as02 = rhi(T.asyrhi, 2, ib);
        const double f3e = 0.17e-3 * ex01 + 0.4 * ex02, f3s = 0.17e-3 * sc01 + 0.4 * sc02;
        const double f3ss = 0.17e-3 * ss01 * ex01 + 0.4 * ss02 * ex02, f3as = 0.17e-3 * as01 * sc01 + 0.4 * as02 * sc02;
        // mineral layer and stratosphere
        const double ex6 = rhi(T.extrhi, 6, ib), ssa6 = rhi(T.ssarhi, 6, ib), asy6 = rhi(T.asyrhi, 6, ib);
        const double exst = (double)T.extstra[b];
        const double ssa_st = sw ? 0.99 : 0.5, asy_st = sw ? 0.696 : 0.3;
        double prev = 0.0;
        for (int l = 0; l < nrl; ++l) {
            const int idom = s_idom[l][lane], code = s_ih[l][lane];
            const int ih1 = code & 15, ih2 = code >> 4;
            const double rdrh = s_rdrh[l][lane], dzk = s_delz[l][lane];
            double tau, ssa, asy;
            if (idom == 5 || idom == 4) {
                tau = (idom == 5) ? 0.0 : exst * dzk;
                ssa = ssa_st, asy = asy_st;
            } else if (idom == 3) {
                const double ex03 = rhd(T.extrhd, ih1, 1, ib) + rdrh * (rhd(T.extrhd, ih2, 1, ib) - rhd(T.extrhd, ih1, 1, ib));
                const double sc03 = rhd(T.scarhd, ih1, 1, ib) + rdrh * (rhd(T.scarhd, ih2, 1, ib) - rhd(T.scarhd, ih1, 1, ib));
                const double ss03 = rhd(T.ssarhd, ih1, 1, ib) + rdrh * (rhd(T.ssarhd, ih2, 1, ib) - rhd(T.ssarhd, ih1, 1, ib));
                const double as03 = rhd(T.asyrhd, ih1, 1, ib) + rdrh * (rhd(T.asyrhd, ih2, 1, ib) - rhd(T.asyrhd, ih1, 1, ib));
                const double ext1 = f3e + 0.59983 * ex03, sca1 = f3s + 0.59983 * sc03;
                const double ssa1 = f3ss + 0.59983 * ss03 * ex03, asy1 = f3as + 0.59983 * as03 * sc03;
                tau = ext1 * 730.0 * dzk;
                ssa = (ext1 > 0.0) ? fmin(1.0, ssa1 / ext1) : 1.0;
                asy = (sca1 > 0.0) ? fmin(1.0, asy1 / sca1) : 0.0;
            } else if (idom == 1) {
                double ext1 = e6, sca1 = s6, ssa1 = ss6, asy1 = as6;
#pragma unroll
                for (int ic1 = 1; ic1 <= A_NCM - A_NCM1; ++ic1) {
                    const double cm = cmix[A_NCM1 + ic1];
                    if (cm > 0.0) {
                        const double ex = rhd(T.extrhd, ih1, ic1, ib) + rdrh * (rhd(T.extrhd, ih2, ic1, ib) - rhd(T.extrhd, ih1, ic1, ib));
                        const double sc = rhd(T.scarhd, ih1, ic1, ib) + rdrh * (rhd(T.scarhd, ih2, ic1, ib) - rhd(T.scarhd, ih1, ic1, ib));
                        const double ss = rhd(T.ssarhd, ih1, ic1, ib) + rdrh * (rhd(T.ssarhd, ih2, ic1, ib) - rhd(T.ssarhd, ih1, ic1, ib));
                        const double as = rhd(T.asyrhd, ih1, ic1, ib) + rdrh * (rhd(T.asyrhd, ih2, ic1, ib) - rhd(T.asyrhd, ih1, ic1, ib));
                        ext1 = ext1 + cm * ex;
                        sca1 = sca1 + cm * sc;
                        ssa1 = ssa1 + cm * ss * ex;
                        asy1 = asy1 + cm * as * sc;
                    }
                }
                tau = ext1 * denn1 * dzk;
                ssa = (ext1 > 0.0) ? fmin(1.0, ssa1 / ext1) : 1.0;
                asy = (sca1 > 0.0) ? fmin(1.0, asy1 / sca1) : 0.0;
            } else if (idom == 2) {
                tau = ex6 * denn2 * dzk;
                ssa = ssa6, asy = asy6;
            } else {
                tau = 0.0, ssa = 1.0, asy = 0.0;
            }
            if (valid) {
                ot[fs + (size_t)l * g.ld] = (float)ssa;
                ot[2 * fs + (size_t)l * g.ld] = (float)asy;
            }
            if (l > 0) {  // smoothing from the top down (phys_rrtm_aer_property.cu:382-390): the layer
                          // above is final after this step
                const double ratio = (tau > 0.0) ? prev / tau : 1.0;
                const double tt0 = tau + prev, tt1 = 0.2 * tt0, tt2 = tt0 - tt1;
                if (ratio > crt1) tau = tt1, prev = tt2;
                if (ratio < crt2) tau = tt2, prev = tt1;
                if (valid) ot[(size_t)(l - 1) * g.ld] = (float)prev;
            }
            prev = tau;
        }
        if (valid) ot[(size_t)(nrl - 1) * g.ld] = (float)prev;
    }
}

}  // namespace

extern "C" int wr_aer_optics(const wr_aer_tables *tab, const wr_columns *cols, const float *xlon, const float *xlat,
                             const float *slmsk, float *aer_sw, float *aer_lw, int64_t aer_ld, void *stream)
{
    if (!tab || !cols || !xlon || !xlat || !slmsk || !aer_sw || !aer_lw) return WR_ERR_ARG;
    if (!tab->haer || !tab->prsref || !tab->sigref || !tab->rhlev || !tab->extrhi || !tab->scarhi || !tab->ssarhi ||
        !tab->asyrhi || !tab->extrhd || !tab->scarhd || !tab->ssarhd || !tab->asyrhd || !tab->extstra || !tab->cmixg ||
        !tab->denng || !tab->idxcg || !tab->kprfg)
        return WR_ERR_ARG;
    if (!cols->p_lev || !cols->t_lay || !cols->q_lay || cols->ncol < 0 || cols->nlay < 1 || cols->nlay + 1 > A_NL ||
        aer_ld < cols->ncol)
        return WR_ERR_ARG;
    if (cols->ncol == 0) return WR_OK;
    ATab T;
    T.haer = gptr(tab->haer), T.prsref = gptr(tab->prsref), T.sigref = gptr(tab->sigref), T.rhlev = gptr(tab->rhlev);
    T.extrhi = gptr(tab->extrhi), T.scarhi = gptr(tab->scarhi), T.ssarhi = gptr(tab->ssarhi);
    T.asyrhi = gptr(tab->asyrhi), T.extrhd = gptr(tab->extrhd), T.scarhd = gptr(tab->scarhd);
    T.ssarhd = gptr(tab->ssarhd), T.asyrhd = gptr(tab->asyrhd), T.extstra = gptr(tab->extstra);
    T.cmixg = gptr(tab->cmixg), T.denng = gptr(tab->denng), T.idxcg = gptr(tab->idxcg), T.kprfg = gptr(tab->kprfg);
    ACols g;
    g.p_lev = gptr(cols->p_lev), g.t_lay = gptr(cols->t_lay), g.q_lay = gptr(cols->q_lay);
    g.xlon = gptr(xlon), g.xlat = gptr(xlat), g.slmsk = gptr(slmsk), g.sw = gptr(aer_sw), g.lw = gptr(aer_lw);
    g.ld = aer_ld, g.ncol = cols->ncol, g.nlay = cols->nlay;
    hipLaunchKernelGGL(k_aer_optics, dim3((cols->ncol + 63) / 64), dim3(64 * A_WAVES), 0, (hipStream_t)stream, T, g);
    return hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
}
