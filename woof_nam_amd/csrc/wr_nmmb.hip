// wr_nmmb.hip -- the NMMB-state side of the radiation call on the device: wr_nmmb_prep (driver prep,
// driver call's condensate and subgrid cloud, grrad setup, progcld2 with gethml) and wr_nmmb_post
// (driver post).  One thread per column, FP64 on the single-precision state, the CPU oracle's
// operation sequence (compiled without contraction, own exp/log/pow), so that every value agrees
// with the oracle bit for bit.
#include <hip/hip_runtime.h>

#include "../../include/woofrad.h"
#include "wr_albd.h"
#include "wr_dev.h"

namespace {
using namespace wr;

// physcons of progcld2 (phys_rrtm_progcld2.cu:49-55), module_CONSTANTS of the driver call
// (nmmb_common.cuh:36-40)
constexpr double N_G = 9.80665, N_RD = 287.05, N_RV = 461.5, N_FVIRT = N_RV / N_RD - 1.0;
constexpr double N_T0C = 273.15, N_TTP = 273.16;
constexpr double N_RV_MC = 461.6, N_EP2 = 287.04 / 461.6, N_EPSQ = 1.0e-12, N_TRAD_ICE = 273.15 - 30.0;
constexpr double N_STBOLT = 5.67051e-8;
constexpr int N_NL = 66;  // radiation layers of a column (model layers + 1)

// FPVSX (module_MP_FER_HIRES.F90:2915-2971): saturation vapour pressure over water or ice, kPa
__device__ double fpvsx(double t)
{
#pragma clang fp contract(off)
    const double ttp = 2.7316e+2, hvap = 2.5000e+6, psat = 6.1078e+2, cliq = 4.1855e+3, cvap = 1.8460e+3,
                 cice = 2.1060e+3, hsub = 2.8340e+6;
    const double psatk = psat * 1.0e-3;
    const double dldt = cvap - cliq, xa = -dldt / N_RV_MC, xb = xa + hvap / (N_RV_MC * ttp);
    const double dldti = cvap - cice, xai = -dldti / N_RV_MC, xbi = xai + hsub / (N_RV_MC * ttp);
    const double tr = ttp / t;
    if (t >= ttp) return psatk * pow_d(tr, xa) * exp_d(xb * (1.0 - tr));
    return psatk * pow_d(tr, xai) * exp_d(xbi * (1.0 - tr));
}

// FPVS (phys_fer_fpvs.cu:31-46): the FPVSX table over 180..330 K in 0.02 K steps, interpolated
// linearly (its two nodes evaluated here), Magnus-Tetens outside
__device__ double fpvs(double t)
{
#pragma clang fp contract(off)
    const double xmin = 180.0, xmax = 330.0, nx = 7501.0;
    const double xinc = (xmax - xmin) / (nx - 1.0);
    if (t >= xmin && t <= xmax) {
        const double c2 = 1.0 / xinc, c1 = 1.0 - xmin / xinc;
        const double xj = fmin(fmax(c1 + c2 * t, 1.0), nx);
        const int jx = (int)fmin(xj, nx - 1.0);
        const double f0 = fpvsx(xmin + (double)(jx - 1) * xinc), f1 = fpvsx(xmin + (double)jx * xinc);
        return f0 + (xj - (double)jx) * (f1 - f0);
    }
    if (t > xmax) return 0.61078 * exp_d(17.2694 * (t - 273.16) / (t - 35.86));
    return 0.61078 * exp_d(21.8746 * (t - 273.16) / (t - 7.66));
}

struct NState {  // wr_nmmb_state with global-address-space pointers
    int im, jm, lm, j0, j1;
    float pt;
    const WR_G float *pd, *pdsg1, *dsg2, *psgml1, *sgml2, *t, *q, *o3, *cw, *f_ice, *f_rain, *glat;
    __device__ __forceinline__ size_t s3(int k, int j, int i) const { return ((size_t)k * jm + j) * im + i; }
    __device__ __forceinline__ size_t s2(int j, int i) const { return (size_t)j * im + i; }
};
struct NCols {
    WR_G double *p_lev, *t_lay, *q_lay, *o3_lay, *cldf, *cwp, *cip, *crp, *rew, *rei;
    WR_G float *clds;
    WR_G int32_t *mtop, *mbot;
    WR_G float *p_phy;
    WR_G double *o3_top;
};

// climatological ozone (wr_ozone with global-address-space pointers); on = 0: the state's O3
constexpr int OZ_LOZ_MAX = 32;
struct NOz {
    const WR_G float *o3r;   // [12][loz][jmr]
    const WR_G float *pstr;  // [loz] hPa
    int on, loz, jmr, k1oz, k2oz;
    double blte, dlte, facoz;
};

// the Exner function (p / 1e5 Pa)^(R/cp) as funcphys' fpkap table gives it: linear between the
// table's nodes every 10 Pa (phys_rrtm_gas_init.cu:52-72), the nodes evaluated here
__device__ double fpkap(double p)
{
#pragma clang fp contract(off)
    const double rocp = (double)(287.05f / 1004.6f);  // con_rd / con_cp, formed in single precision
    const double xinc = 110000.0 / 11000.0;
    const double xj = fmin(fmax(1.0 + p / xinc, 1.0), 11001.0);
    const int jx = (int)fmin(xj, 11000.0);
    const double t0 = pow_d((double)(jx - 1) * xinc / 1.0e5, rocp);
    const double t1 = pow_d((double)jx * xinc / 1.0e5, rocp);
    return t0 + (xj - (double)jx) * (t1 - t0);
}

// getozn (phys_rrtm_getozn.cu:57-110) at one layer: the climatology column o3i at the climatology's
// Exner levels pk, interpolated linearly in the Exner function wk; the end values outside
__device__ double ozone_at(const double *o3i, const double *pk, int loz, double wk)
{
#pragma clang fp contract(off)
    if (wk > pk[loz - 1]) return o3i[loz - 1];
    if (wk <= pk[0]) return o3i[0];
    double v = o3i[0];
    for (int k = 0; k < loz - 1; ++k) {
        if (wk > pk[k] && wk <= pk[k + 1]) {
            const double tem = (pk[k + 1] - wk) * (1.0 / (pk[k + 1] - pk[k]));
            v = tem * o3i[k] + (1.0 - tem) * o3i[k + 1];
        }
    }
    return v;
}

__global__ void __launch_bounds__(64) k_nmmb_prep(NState st, wr_nmmb_opts o, NCols w, NOz oz)
{
#pragma clang fp contract(off)
    const int nc = (st.j1 - st.j0) * st.im;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int lm = st.lm, nrl = lm + 1;
    const int i = c % st.im, j = st.j0 + c / st.im;
    const size_t snc = (size_t)nc;
    // driver prep (phys_rrtm_driver_prep.cu:82-98) in single precision as there
    const float pd = st.pd[st.s2(j, i)];
    float p8 = st.pt;
    w.p_lev[c] = (double)p8;
    for (int k = 0; k < lm; ++k) {
        p8 = p8 + st.pdsg1[k] + st.dsg2[k] * pd;
        w.p_lev[(size_t)(k + 1) * snc + c] = (double)p8;
        if (w.p_phy) w.p_phy[st.s3(k, j, i)] = st.sgml2[k] * pd + st.psgml1[k];
    }
    // climatological ozone (getozn): this latitude's column of the climatology, between the two months
    double o3i[OZ_LOZ_MAX], pk[OZ_LOZ_MAX];
    bool oz_on = oz.on != 0;
    if (oz_on) {
        const double deglat = (double)st.glat[st.s2(j, i)] * 57.29578;
        const double elte = oz.blte + (double)(oz.jmr - 1) * oz.dlte;
        int j1, j2;
        double tem1;
        if (deglat > oz.blte && deglat < elte) {
            tem1 = (deglat - oz.blte) / oz.dlte + 1.0;
            j1 = (int)tem1;
            j2 = j1 + 1;
            tem1 = tem1 - (double)j1;
        } else if (deglat <= oz.blte) {
            j1 = j2 = 1, tem1 = 1.0;
        } else if (deglat >= elte) {
            j1 = j2 = oz.jmr, tem1 = 1.0;
        } else {
            j1 = j2 = 1, tem1 = 1.0, oz_on = false;  // latitude NaN: the state's O3
        }
        const double tem2 = 1.0 - tem1;
        const double rocp = (double)(287.05f / 1004.6f);
        for (int k = 0; k < oz.loz; ++k) {
            const size_t m1 = ((size_t)(oz.k1oz - 1) * oz.loz + k) * oz.jmr, m2 = ((size_t)(oz.k2oz - 1) * oz.loz + k) * oz.jmr;
            const double tem3 = tem2 * (double)oz.o3r[m1 + j1 - 1] + tem1 * (double)oz.o3r[m1 + j2 - 1];
            const double tem4 = tem2 * (double)oz.o3r[m2 + j1 - 1] + tem1 * (double)oz.o3r[m2 + j2 - 1];
            o3i[k] = tem4 * oz.facoz + tem3 * (1.0 - oz.facoz);
            const double ps = (double)oz.pstr[k];
            pk[k] = ps >= 0.1 ? fpkap(ps * 100.0) : pow_d(ps / 1000.0, rocp);
        }
    }
    // layer pressures (hPa) as the column call forms them from the interfaces; layer 0 is the added
    // layer above the model top.  ct keeps the layer cloud fractions for gethml.
    double ct[N_NL];
    double plyr_top = 0.0;
    {
        double pl1 = 10.0 * (w.p_lev[c] * 1.0e-3);
        if (pl1 <= 1.0e-6) pl1 = 2.0e-6;
        plyr_top = 0.5 * pl1;
    }
    if (w.o3_top)  // the added top layer: getozn at its Exner function (grrad_setup's fpkapx), or none
        w.o3_top[c] = oz_on ? ozone_at(o3i, pk, oz.loz, pow_d(plyr_top * 100.0 / 1.0e5, (double)(287.05f / 1004.6f)))
                            : (double)st.o3[st.s3(0, j, i)];
    ct[0] = 0.0;
    const double recw1 = 620.35 / pow_d(200.0, 1.0 / 3.0);
    double pb = w.p_lev[c];
    for (int k = 0; k < lm; ++k) {
        const size_t kc = (size_t)k * snc + c;
        const size_t s = st.s3(k, j, i);
        const double pt_ = pb;
        pb = w.p_lev[kc + snc];
        const double t = (double)st.t[s], q = (double)st.q[s];
        w.t_lay[kc] = t;
        w.q_lay[kc] = q;
        if (oz_on) {  // the model layer's Exner function as the driver call forms it (PRSLK)
            const float prsl = (st.sgml2[k] * st.pd[st.s2(j, i)] + st.psgml1[k]) / 1000.0f;
            const double wk = pow_d((double)prsl * 0.01, (double)(287.04f / 1004.6f));
            w.o3_lay[kc] = ozone_at(o3i, pk, oz.loz, wk);
        } else {
            w.o3_lay[kc] = (double)st.o3[s];
        }
        double cw = (double)st.cw[s];
        double fice = (double)st.f_ice[s], frain = (double)st.f_rain[s];
        // driver call SUBGRID (phys_rrtm_driver_call.cu:248-275)
        if (o.subgrid) {
            const double wv = q / (1.0 - q);
            const double p1 = 0.5 * (pt_ + pb) * 1.0e-3;
            const double esat = fmin(fpvs(t), 0.99 * p1);
            const double qsat = N_EP2 * esat / (p1 - esat);
            const double rhum = wv / qsat;
            if (cw < N_EPSQ && rhum > 0.95) {
                cw = fmin(0.01e-3, fmin(0.01, rhum - 0.95) * qsat);
                fice = (t > N_TRAD_ICE) ? 0.0 : 1.0;
                frain = 0.0;
            }
        }
        const double clw = cw < N_EPSQ ? 0.0 : cw;  // grrad clouds (phys_rrtm_grrad_clouds.cu:397-412)
        // progcld2 (phys_rrtm_progcld2.cu:145-305)
        const double pl = 10.0 * (0.5 * (pt_ + pb) * 1.0e-3);
        double plo = 10.0 * (pt_ * 1.0e-3);
        if (k == 0 && plo <= 1.0e-6) plo = 2.0e-6;
        const double plp = 10.0 * (pb * 1.0e-3);
        const double qlyr = fmax(1.0e-7, q);
        const double tvly = t * (1.0 + N_FVIRT * qlyr);
        const double fc = fmax(0.0, fmin(1.0, fice)), fr = fmax(0.0, fmin(1.0, frain));
        double qcice, qrain, qcwat, clw2;
        if (t - N_T0C > -40.0) {
            qcice = clw * fc;
            const double tem1 = clw - qcice;
            qrain = tem1 * fr;
            qcwat = tem1 - qrain;
            clw2 = qcwat + qcice;
        } else {
            qcice = clw;
            qrain = 0.0;
            qcwat = 0.0;
            clw2 = clw;
        }
        double cwp = 0.0, cip = 0.0, crp = 0.0, rew = 5.0, rei = 10.0;
        if (qcwat + qcice + qrain > 1.0e-12) {
            const double rho = 100.0 * pl / (N_RD * t * (1.0 + N_FVIRT * qlyr));
            const double cpath = fabs(plp - plo) * (100000.0 / N_G);
            if (qcwat > 0.0) {
                rew = fmax(5.0, recw1 * pow_d(rho * qcwat, 1.0 / 3.0));
                cwp = cpath * qcwat;
            }
            if (qrain > 0.0) crp = cpath * qrain;
            cip = cpath * qcice;
        }
        const double tem2d = (N_G * pl) / (N_RD * (plp - plo));
        double cf = 0.0;
        if (clw2 > 1.0e-8 * (pl * 0.001)) cf = fmin(1.0, 1.0e5 * clw2 * pl * 0.001);
        if (cf < 0.001) {
            cf = 0.0;
            cwp = 0.0;
            cip = 0.0;
            crp = 0.0;
        }
        if (cip > 0.0) {
            const double tem3 = tem2d * cip / tvly;
            const double tem1 = t - N_TTP;
            double tem2;
            if (tem1 < -50.0)
                tem2 = (1250.0 / 9.917) * pow_d(tem3, 0.109);
            else if (tem1 < -40.0)
                tem2 = (1250.0 / 9.337) * pow_d(tem3, 0.08);
            else if (tem1 < -30.0)
                tem2 = (1250.0 / 9.208) * pow_d(tem3, 0.055);
            else
                tem2 = (1250.0 / 9.387) * pow_d(tem3, 0.031);
            rei = fmax(10.0, tem2 * 0.6495);
        }
        ct[k + 1] = cf;
        w.cldf[kc] = cf;
        w.cwp[kc] = cwp;
        w.cip[kc] = cip;
        w.crp[kc] = crp;
        w.rew[kc] = rew;
        w.rei[kc] = rei;
    }
    // gethml (phys_rrtm_progcld2.cu:307-450, ivflip = 0); OVCST = 1 - 1e-8 in single precision is 1
    auto plyr_at = [&](int l) {  // l: 0-based radiation layer
        if (l == 0) return plyr_top;
        return 10.0 * (0.5 * (w.p_lev[(size_t)(l - 1) * snc + c] + w.p_lev[(size_t)l * snc + c]) * 1.0e-3);
    };
    double ptop1[4];
    {
        const double p1c[4] = {1050.0, 650.0, 400.0, 0.0}, p2c[4] = {1050.0, 750.0, 500.0, 0.0};
        const double tem2 = fmax(0.0, 4.0 * fabs((double)st.glat[st.s2(j, i)]) / 3.1415926535897931 - 1.0);
        for (int id = 0; id < 4; ++id) ptop1[id] = p1c[id] + (p2c[id] - p1c[id]) * tem2;
    }
    double cl[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    int mt[4] = {0, 0, 0, 0}, mb[4] = {0, 0, 0, 0};
    auto ct1 = [&](int k) { return fmin(1.0, fmax(ct[k - 1], 0.0)); };
    {
        double cl1 = 1.0, cl2 = 1.0;
        for (int k = nrl; k >= 1; --k) {
            const double ccur = ct1(k);
            if (o.iovr == 0) {
                if (ccur >= 0.001) cl1 = cl1 * (1.0 - ccur);
                if (k == o.llyr) cl[4] = 1.0 - cl1;
            } else {
                if (ccur >= 0.001) {
                    cl2 = fmin(cl2, 1.0 - ccur);
                } else {
                    cl1 = cl1 * cl2;
                    cl2 = 1.0;
                }
                if (k == o.llyr) cl[4] = 1.0 - cl1 * cl2;
            }
        }
        cl[3] = (o.iovr == 0) ? 1.0 - cl1 : 1.0 - cl1 * cl2;
    }
    {
        double cl1 = 0.0, cl2 = 0.0;
        int kbt1 = nrl, kbt2 = nrl, kth1 = 0, kth2 = 0, idom = 1;
        mb[0] = nrl;
        mt[0] = nrl;
        double pcur = plyr_at(nrl - 1);
        for (int k = nrl; k >= 1; --k) {
            int id = idom, id1 = id + 1;
            const double ccur = ct1(k);
            double pnxt, cnxt;
            if (k > 1) {
                pnxt = plyr_at(k - 2);
                cnxt = ct1(k - 1);
            } else {
                pnxt = -1.0;
                cnxt = 0.0;
            }
            if (pcur < ptop1[id1 - 1]) {
                id += 1;
                id1 += 1;
                idom = id;
            }
            if (ccur >= 0.001) {
                if (kth2 == 0) kbt2 = k;
                kth2 += 1;
                cl2 = (o.iovr == 0) ? cl2 + ccur - cl2 * ccur : fmax(cl2, ccur);
                if (cnxt < 0.001 || pnxt < ptop1[id1 - 1]) {
                    kbt1 = (int)round((cl1 * (double)kbt1 + cl2 * (double)kbt2) / (cl1 + cl2));
                    kth1 = (int)round((cl1 * (double)kth1 + cl2 * (double)kth2) / (cl1 + cl2));
                    cl1 = cl1 + cl2 - cl1 * cl2;
                    kbt2 = k - 1;
                    kth2 = 0;
                    cl2 = 0.0;
                }
            }
            if (pnxt < ptop1[id1 - 1]) {
                cl[id - 1] = cl1;
                mt[id - 1] = imin(kbt1, kbt1 - kth1 + 1);
                mb[id - 1] = kbt1;
                cl1 = 0.0;
                kbt1 = k - 1;
                kth1 = 0;
                if (id1 <= 3) {
                    mb[id1 - 1] = kbt1;
                    mt[id1 - 1] = kbt1;
                }
            }
            pcur = pnxt;
        }
    }
    for (int n = 0; n < 5; ++n) w.clds[(size_t)n * snc + c] = (float)cl[n];
    for (int n = 0; n < 3; ++n) {
        w.mtop[(size_t)n * snc + c] = mt[n];
        w.mbot[(size_t)n * snc + c] = mb[n];
    }
}

struct NOut {
    WR_G float *rlwtt, *rswtt, *cldfra, *rlwin, *rlwtoa, *sigt4, *rswin, *rswout, *rswinc, *rswtoa, *czmean, *cfracl,
        *cfracm, *cfrach, *acfrst, *acfrcv;
    WR_G int32_t *ncfrst, *ncfrcv;
};
struct NRad {
    const WR_G double *hr_lw, *hr_sw, *flw_up, *flw_dn, *fsw_up, *fsw_dn, *fsw_dn_sfc_clr;
};
// the prep's cloud fields and the call's mean cos zenith.  Kernel parameters are structs or plain
// values only: an address-space-qualified pointer parameter mangles differently in the host pass
// (which registers the kernel by name) and the device pass (which names the code object's symbol), and
// the launch then finds no kernel.
struct NPostIn {
    const WR_G double *cldf;   // [lm][ncol]
    const WR_G float *clds;    // [5][ncol]
    const WR_G double *czmean; // [ncol]
};

// driver post (phys_rrtm_driver_post.cu:71-150): icwp /= -1, lssav, CU_CLOUD = 0 (CUCLD off)
__global__ void __launch_bounds__(64) k_nmmb_post(NState st, NPostIn pin, NRad rad, int lw_step, int sw_step, NOut o)
{
#pragma clang fp contract(off)
    const int nc = (st.j1 - st.j0) * st.im;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int lm = st.lm, nrl = lm + 1;
    const int i = c % st.im, j = st.j0 + c / st.im;
    const size_t snc = (size_t)nc, s2 = st.s2(j, i);
    if (lw_step) {
        for (int k = 0; k < lm; ++k) o.rlwtt[st.s3(k, j, i)] = (float)rad.hr_lw[(size_t)k * snc + c];
        o.rlwin[s2] = (float)rad.flw_dn[(size_t)nrl * snc + c];
        o.rlwtoa[s2] = (float)rad.flw_up[c];
        const double tl = (double)st.t[st.s3(lm - 1, j, i)];
        o.sigt4[s2] = (float)(N_STBOLT * tl * tl * tl * tl);
    }
    if (sw_step) {
        for (int k = 0; k < lm; ++k) o.rswtt[st.s3(k, j, i)] = (float)rad.hr_sw[(size_t)k * snc + c];
        o.rswin[s2] = (float)rad.fsw_dn[(size_t)nrl * snc + c];
        o.rswout[s2] = (float)rad.fsw_up[(size_t)nrl * snc + c];
        o.rswinc[s2] = (float)rad.fsw_dn_sfc_clr[c];
        o.rswtoa[s2] = (float)rad.fsw_up[c];
    }
    const WR_G double *cldf = pin.cldf;
    const WR_G float *clds = pin.clds;
    o.czmean[s2] = (float)pin.czmean[c];
    float cwmax = -3.402823466e38f;
    for (int k = 0; k < lm; ++k) cwmax = fmaxf(cwmax, st.cw[st.s3(k, j, i)]);
    const bool clear = cwmax < 1.0e-8f;
    for (int k = 0; k < lm; ++k) o.cldfra[st.s3(k, j, i)] = clear ? 0.0f : (float)cldf[(size_t)k * snc + c];
    float cs[4];
    for (int n = 0; n < 4; ++n) cs[n] = clear ? 0.0f : clds[(size_t)n * snc + c];
    o.cfracl[s2] = cs[0];
    o.cfracm[s2] = cs[1];
    o.cfrach[s2] = cs[2];
    o.acfrst[s2] = o.acfrst[s2] + cs[3];
    o.ncfrst[s2] = o.ncfrst[s2] + 1;
    o.ncfrcv[s2] = o.ncfrcv[s2] + 1;
}

// Payne (1972) open-water albedo at the NAM's atmospheric transmission 0.537 (the IALB = 2 block of the
// NMMB radiation driver, phys_rrtm_grrad_clouds.cu:363-385): linear in the zenith angle between the
// table's nodes, then in the transmission between its 0.50 and 0.55 columns
__constant__ double c_albd_za[20] = WR_ALBD_ZA;
__constant__ double c_albd_dza[19] = WR_ALBD_DZA;
__constant__ double c_albd_lo[20] = WR_ALBD_LO;
__constant__ double c_albd_hi[20] = WR_ALBD_HI;

struct NAlb {  // kernel parameters are structs or plain values (see NPostIn)
    const WR_G float *albedo, *sm;  // [jm][im]
    const WR_G double *coszen;      // [ncol]
    WR_G double *alb;               // [ncol]
    int im, j0, nc;
};

__global__ void __launch_bounds__(256) k_nmmb_albedo(NAlb g)
{
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.nc) return;
    const size_t p = (size_t)(g.j0 + c / g.im) * g.im + c % g.im;
    double a = (double)g.albedo[p];
    const double cz = g.coszen[c];
    if (cz > 0.0 && (double)g.sm[p] > 0.5) {
        const double zen = (180.0 / 3.14159265) * acos(fmax(cz, 0.0));
        int jx;
        if (zen >= 74.0)
            jx = (int)(0.5 * (90.0 - zen) + 1.0);
        else if (zen >= 50.0)
            jx = (int)(0.25 * (74.0 - zen) + 9.0);
        else
            jx = (int)(0.1 * (50.0 - zen) + 15.0);
        jx = imin(19, imax(1, jx));  // zenith 0 exactly would index past the table's last node
        const double dzen = -(zen - c_albd_za[jx - 1]) / c_albd_dza[jx - 1];
        const double a1 = c_albd_lo[jx - 1] + dzen * (c_albd_lo[jx] - c_albd_lo[jx - 1]);
        const double a2 = c_albd_hi[jx - 1] + dzen * (c_albd_hi[jx] - c_albd_hi[jx - 1]);
        a = a1 + 20.0 * (a2 - a1) * (0.537 - WR_ALBD_TRN_LO);
    }
    g.alb[c] = a;
}

int check_state(const wr_nmmb_state *st)
{
    if (!st || st->im <= 0 || st->jm <= 0 || st->lm < 2 || st->lm + 1 > N_NL) return WR_ERR_ARG;
    if (st->j0 < 0 || st->j1 > st->jm || st->j0 >= st->j1) return WR_ERR_ARG;
    if ((int64_t)(st->j1 - st->j0) * st->im > 0x7fffffff) return WR_ERR_ARG;
    if (!st->pd || !st->pdsg1 || !st->dsg2 || !st->psgml1 || !st->sgml2 || !st->t || !st->q || !st->o3 || !st->cw ||
        !st->f_ice || !st->f_rain || !st->glat)
        return WR_ERR_ARG;
    return WR_OK;
}

NState to_state(const wr_nmmb_state *s)
{
    NState n;
    n.im = s->im, n.jm = s->jm, n.lm = s->lm, n.j0 = s->j0, n.j1 = s->j1, n.pt = s->pt;
    n.pd = gptr(s->pd), n.pdsg1 = gptr(s->pdsg1), n.dsg2 = gptr(s->dsg2), n.psgml1 = gptr(s->psgml1);
    n.sgml2 = gptr(s->sgml2), n.t = gptr(s->t), n.q = gptr(s->q), n.o3 = gptr(s->o3), n.cw = gptr(s->cw);
    n.f_ice = gptr(s->f_ice), n.f_rain = gptr(s->f_rain), n.glat = gptr(s->glat);
    return n;
}

}  // namespace

extern "C" {

int wr_nmmb_prep(const wr_nmmb_state *st, const wr_nmmb_opts *opt, const wr_nmmb_cols *w, void *stream)
{
    int rc = check_state(st);
    if (rc != WR_OK) return rc;
    if (!opt || !w || opt->iovr < 0 || opt->iovr > 1 || opt->llyr < 1 || opt->llyr > st->lm + 1) return WR_ERR_ARG;
    if (!w->p_lev || !w->t_lay || !w->q_lay || !w->o3_lay || !w->cldf || !w->cwp || !w->cip || !w->crp || !w->rew ||
        !w->rei || !w->clds || !w->mtop || !w->mbot)
        return WR_ERR_ARG;
    NCols c;
    c.p_lev = gptr(w->p_lev), c.t_lay = gptr(w->t_lay), c.q_lay = gptr(w->q_lay), c.o3_lay = gptr(w->o3_lay);
    c.cldf = gptr(w->cldf), c.cwp = gptr(w->cwp), c.cip = gptr(w->cip), c.crp = gptr(w->crp), c.rew = gptr(w->rew);
    c.rei = gptr(w->rei), c.clds = gptr(w->clds), c.mtop = gptr(w->mtop), c.mbot = gptr(w->mbot);
    c.p_phy = gptr(w->p_phy);
    c.o3_top = gptr(w->o3_top);
    NOz oz{};
    if (opt->ozone) {
        const wr_ozone *z = opt->ozone;
        if (!z->o3r || !z->pstr || z->loz < 2 || z->loz > OZ_LOZ_MAX || z->jmr < 1 || !(z->dlte > 0.0f) ||
            z->k1oz < 1 || z->k1oz > 12 || z->k2oz < 1 || z->k2oz > 12 || !w->o3_top)
            return WR_ERR_ARG;
        oz.o3r = gptr(z->o3r), oz.pstr = gptr(z->pstr), oz.on = 1, oz.loz = z->loz, oz.jmr = z->jmr;
        oz.k1oz = z->k1oz, oz.k2oz = z->k2oz, oz.blte = z->blte, oz.dlte = z->dlte, oz.facoz = z->facoz;
    }
    wr_nmmb_opts ob = *opt;
    ob.ozone = nullptr;  // the kernel takes the climatology through oz
    const int nc = (st->j1 - st->j0) * st->im;
    hipLaunchKernelGGL(k_nmmb_prep, dim3((nc + 63) / 64), dim3(64), 0, (hipStream_t)stream, to_state(st), ob, c, oz);
    return hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
}

int wr_nmmb_albedo(const wr_nmmb_state *st, const float *albedo, const float *sm, const double *coszen, double *alb,
                   void *stream)
{
    if (!st || st->im <= 0 || st->jm <= 0 || st->j0 < 0 || st->j1 > st->jm || st->j0 >= st->j1) return WR_ERR_ARG;
    if ((int64_t)(st->j1 - st->j0) * st->im > 0x7fffffff) return WR_ERR_ARG;
    if (!albedo || !sm || !coszen || !alb) return WR_ERR_ARG;
    const int nc = (st->j1 - st->j0) * st->im;
    NAlb g;
    g.albedo = gptr(albedo), g.sm = gptr(sm), g.coszen = gptr(coszen), g.alb = gptr(alb);
    g.im = st->im, g.j0 = st->j0, g.nc = nc;
    hipLaunchKernelGGL(k_nmmb_albedo, dim3((nc + 255) / 256), dim3(256), 0, (hipStream_t)stream, g);
    return hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
}

int wr_nmmb_post(const wr_nmmb_state *st, const wr_nmmb_cols *w, const wr_outputs *rad, const double *czmean,
                 int32_t lw_step, int32_t sw_step, const wr_nmmb_out *o, void *stream)
{
    int rc = check_state(st);
    if (rc != WR_OK) return rc;
    if (!w || !rad || !o || !czmean || !w->cldf || !w->clds) return WR_ERR_ARG;
    if (!o->cldfra || !o->czmean || !o->cfracl || !o->cfracm || !o->cfrach || !o->acfrst || !o->ncfrst || !o->ncfrcv)
        return WR_ERR_ARG;
    if (lw_step && (!rad->hr_lw || !rad->flw_up || !rad->flw_dn || !o->rlwtt || !o->rlwin || !o->rlwtoa || !o->sigt4))
        return WR_ERR_ARG;
    if (sw_step && (!rad->hr_sw || !rad->fsw_up || !rad->fsw_dn || !rad->fsw_dn_sfc_clr || !o->rswtt || !o->rswin ||
                    !o->rswout || !o->rswinc || !o->rswtoa))
        return WR_ERR_ARG;
    NRad r;
    r.hr_lw = gptr((const double *)rad->hr_lw), r.hr_sw = gptr((const double *)rad->hr_sw);
    r.flw_up = gptr((const double *)rad->flw_up), r.flw_dn = gptr((const double *)rad->flw_dn);
    r.fsw_up = gptr((const double *)rad->fsw_up), r.fsw_dn = gptr((const double *)rad->fsw_dn);
    r.fsw_dn_sfc_clr = gptr((const double *)rad->fsw_dn_sfc_clr);
    NOut n;
    n.rlwtt = gptr(o->rlwtt), n.rswtt = gptr(o->rswtt), n.cldfra = gptr(o->cldfra), n.rlwin = gptr(o->rlwin);
    n.rlwtoa = gptr(o->rlwtoa), n.sigt4 = gptr(o->sigt4), n.rswin = gptr(o->rswin), n.rswout = gptr(o->rswout);
    n.rswinc = gptr(o->rswinc), n.rswtoa = gptr(o->rswtoa), n.czmean = gptr(o->czmean), n.cfracl = gptr(o->cfracl);
    n.cfracm = gptr(o->cfracm), n.cfrach = gptr(o->cfrach), n.acfrst = gptr(o->acfrst), n.acfrcv = gptr(o->acfrcv);
    n.ncfrst = gptr(o->ncfrst), n.ncfrcv = gptr(o->ncfrcv);
    const int nc = (st->j1 - st->j0) * st->im;
    NPostIn pin;
    pin.cldf = gptr((const double *)w->cldf), pin.clds = gptr((const float *)w->clds), pin.czmean = gptr(czmean);
    hipLaunchKernelGGL(k_nmmb_post, dim3((nc + 63) / 64), dim3(64), 0, (hipStream_t)stream, to_state(st), pin, r,
                       (int)lw_step, (int)sw_step, n);
    return hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
}

}  // extern "C"
