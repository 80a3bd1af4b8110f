// wr_solar.hip -- the solar geometry around the radiation call: sol_update (host: the sun's position and
// the solar constant for a date, once per call), coszmn (device: the mean cosine of the zenith angle of
// each column over the shortwave interval) and rdtemp (device: every model step, the zenith angle of
// the step and the radiative temperature change T += (RSWTT czen / CZMEAN + RLWTT) dt).  FP64
// throughout (the reference evaluates in single precision).
#include <hip/hip_runtime.h>
#include <math.h>

#include "../../include/woofrad.h"

namespace {

constexpr double S_PI = 3.1415926535897931;
constexpr double S_TPI = 2.0 * S_PI;
constexpr double S_DEGRAD = 180.0 / S_PI;

// Julian day number of a Gregorian date (phys_rrtm_sol_update.cu:16-22)
int jdn(int y, int m, int d)
{
    return d - 32075 + 1461 * (y + 4800 + (m - 14) / 12) / 4 + 367 * (m - 2 - (m - 14) / 12 * 12) / 12 -
           3 * ((y + 4900 + (m - 14) / 12) / 100) / 4;
}

// the sun's orbit for Julian day jd + fjd (phys_rrtm_sol_update.cu:24-100): eccentric anomaly by Newton
// iteration of Kepler's equation, then declination, right ascension and the solar lag
void orbit(int jd, double fjd, double &r1, double &sollag, double &sindec, double &cosdec)
{
    const int jdor = 2415020;  // Julian day of the epoch, 0 January 1900 at 12 UT
    const double cyear = 365.25, ccr = 1.3e-6, tpp = 1.55, svt6 = 78.035;
    const double t1 = (double)(jd - jdor) / 36525.0;
    double year = 0.25964134 + 0.304e-5 * t1;
    const double tyear = 0.24219879 - 0.614e-5 * t1;
    const double ec = 0.01675104 - (0.418e-4 + 0.126e-6 * t1) * t1;
    const double angin = 23.452294 - (0.0130125 + 0.164e-5 * t1) * t1;
    const int jdoe = (int)((double)jdor + (svt6 * cyear) / (year - tyear));
    const double deleqn = (double)(jdoe - jd) * (year - tyear) / cyear;
    year = year + 365.0;
    const double sni = sin(angin / S_DEGRAD);
    const double tini = 1.0 / tan(angin / S_DEGRAD);
    const double er = sqrt((1.0 + ec) / (1.0 - ec));
    auto kepler = [&](double mean) {
        double e1 = 1.0, d = 1.0;
        for (int it = 0; d > ccr && it <= 10; ++it) {
            const double ep = e1 - (e1 - ec * sin(e1) - mean) / (1.0 - ec * cos(e1));
            d = fabs(e1 - ep);
            e1 = ep;
        }
        return e1;
    };
    const double eq = 2.0 * atan(er * tan(0.5 * kepler(deleqn * S_TPI / year)));
    const double date = fmod((double)(jd - jdor) - tpp + fjd, year);
    const double e1 = kepler(S_TPI * date / year);
    const double w1 = 2.0 * atan(er * tan(0.5 * e1));
    r1 = 1.0 - ec * cos(e1);
    sindec = sni * sin(w1 - eq);
    cosdec = sqrt(1.0 - sindec * sindec);
    double alp = asin(tan(asin(sindec)) * tini);
    if (cos(w1 - eq) < 0.0) alp = S_PI - alp;
    if (alp < 0.0) alp = alp + S_TPI;
    double sun = S_TPI * (date - deleqn) / year;
    if (sun < 0.0) sun = sun + S_TPI;
    sollag = sun - alp - 0.03255;
}

__global__ void k_coszmn(int ncol, const double *__restrict__ glat, const double *__restrict__ glon, double solhr,
                         double dtswav, int nrads, wr_sol s, double *__restrict__ coszen, double *__restrict__ coszdg)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncol) return;
    const double pid12 = (2.0 * asin(1.0)) / 12.0;
    const double solang = pid12 * (solhr - 12.0);
    const double rnstp = dtswav / (double)nrads;
    const double sinlat = sin(glat[i]), coslat = cos(glat[i]);
    double cz = 0.0;
    int istsun = 0;
    for (int it = 1; it <= nrads; ++it) {
        const double cns = solang + (double)(it - 1) * pid12 * rnstp + s.sollag;
        const double coszn = sinlat * s.sindec + coslat * s.cosdec * cos(cns + glon[i]);
        cz += fmax(0.0, coszn);
        if (coszn > 1.0e-4) istsun += 1;
    }
    if (coszdg) coszdg[i] = cz * rnstp;
    coszen[i] = istsun > 0 ? cz / (double)istsun : cz;
}

// ZENITH's declination and hour angle for a step (drv_rdtemp.cu:38-100)
struct Sun {
    double dec, hrang;
};
Sun zenith(double times, int ihrst, int m, int d, int y, int leap)
{
    const double obliq = 23.440 * (S_PI / 180.0), deg2rd = S_PI / 180.0;
    int month[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
    if (y % 4 == 0 || leap) month[1] = 29;
    double day = 0.0;
    for (int k = 1; k <= m - 1; ++k) day = day + (double)month[k - 1];
    day = day + (double)(d - 1) + ((double)ihrst + times / 3600.0) / 24.0;
    const double dayi = (double)((int)day + 1);
    const double hour = (day - dayi + 1.0) * 24.0;
    const double yfctr = 2000.0 - (double)y;
    const int idifyr = y - 2000;
    const double addday = idifyr < 0 ? (double)(idifyr / 4) : (double)((idifyr + 3) / 4);
    const double startyr = 2451545.0 + (double)idifyr * 365.0 + addday - 0.5;
    const double difjd = startyr + day - 2451545.0;
    const double slonm = (280.460 + 0.9856474 * difjd) * deg2rd + yfctr * S_TPI;
    const double anom = (357.528 + 0.9856003 * difjd) * deg2rd;
    double slon = slonm + (1.915 * sin(anom) + 0.020 * sin(2.0 * anom)) * deg2rd;
    if (slon > S_TPI) slon = slon - S_TPI;
    Sun s;
    s.dec = asin(sin(slon) * sin(obliq));
    double ra = acos(cos(slon) / cos(s.dec));
    if (slon > S_PI) ra = S_TPI - ra;
    const double tu = (startyr + dayi - 1.0 - 2451545.0) / 36525.0;
    const double stim0 = 24110.54841 + tu * (8640184.812866 + 9.3104e-2 * tu + -6.2e-6 * tu * tu);
    double sidtim = (stim0 / 3600.0 + yfctr * 24.0 + 1.00273791 * hour) * 15.0 * deg2rd;
    if (sidtim < 0.0) sidtim = sidtim + S_TPI;
    if (sidtim > S_TPI) sidtim = sidtim - S_TPI;
    s.hrang = sidtim - ra;
    return s;
}

__global__ void k_rdtemp(int im, int jm, int lm, int i_lo, int i_hi, int j_lo, int j_hi, double dt, Sun sun,
                         const float *__restrict__ glat, const float *__restrict__ glon, const float *__restrict__ czmean,
                         const float *__restrict__ rswtt, const float *__restrict__ rlwtt, float *__restrict__ t,
                         float *__restrict__ czen)
{
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= im * jm) return;
    const int i = c % im, j = c / im;
    const double hrlcl = sun.hrang + (double)glon[c] + S_TPI;
    double sinalt = sin(sun.dec) * sin((double)glat[c]) + cos(sun.dec) * cos(hrlcl) * cos((double)glat[c]);
    if (sinalt < 0.0) sinalt = 0.0;
    czen[c] = (float)sinalt;
    if (i < i_lo || i > i_hi || j < j_lo || j > j_hi) return;
    const double factr = czmean[c] > 0.0f ? sinalt / (double)czmean[c] : 0.0;
    for (int l = 0; l < lm; ++l) {
        const size_t p = ((size_t)l * jm + j) * im + i;
        t[p] = (float)((double)t[p] + ((double)rswtt[p] * factr + (double)rlwtt[p]) * dt);
    }
}

}  // namespace

extern "C" {

int wr_sol_update(int32_t yr, int32_t mon, int32_t day, int32_t hr, int32_t mn, int32_t sec, double deltsw,
                  double deltim, double solc0, wr_sol *o)
{
#pragma clang fp contract(off)
    if (!o || !(deltim > 0.0) || !(deltsw > 0.0) || mon < 1 || mon > 12 || day < 1 || day > 31 || hr < 0 || hr > 23)
        return WR_ERR_ARG;
    int jd1 = jdn(yr, mon, day);
    double fjd1;
    if (hr < 12) {
        jd1 = jd1 - 1;
        fjd1 = 0.5 + (double)hr / 24.0 + (double)mn / 1440.0 + (double)sec / 86400.0;
    } else {
        fjd1 = (double)(hr - 12) / 24.0 + (double)mn / 1440.0 + (double)sec / 86400.0;
    }
    fjd1 = fjd1 + (double)jd1;
    const int jd = (int)fjd1;
    double r1, sollag, sindec, cosdec;
    orbit(jd, fjd1 - (double)jd, r1, sollag, sindec, cosdec);
    const double dtswh = deltsw / 3600.0;
    int nstp;
    if (deltsw >= 3600.0) {
        int nn = (int)round(3600.0 / deltim);
        nn = nn < 6 ? 6 : (nn > 12 ? 12 : nn);
        nstp = (int)round(dtswh) * nn + 1;
    } else {
        int nswr = (int)round(deltsw / deltim);
        nstp = (nswr < 2 ? 2 : (nswr > 20 ? 20 : nswr)) + 1;
    }
    o->sollag = sollag;
    o->sindec = sindec;
    o->cosdec = cosdec;
    o->solcon = solc0 / (r1 * r1);
    o->anginc = ((2.0 * asin(1.0)) / 12.0) * dtswh / (double)(nstp - 1);
    o->nstp = nstp;
    return WR_OK;
}

int wr_coszmn(int32_t ncol, const double *glat, const double *glon, double solhr, double dtswav, int32_t nrads,
              const wr_sol *sol, double *coszen, double *coszdg, void *stream)
{
    if (ncol < 0 || nrads < 1 || !sol || !(dtswav > 0.0)) return WR_ERR_ARG;
    if (ncol == 0) return WR_OK;
    if (!glat || !glon || !coszen) return WR_ERR_ARG;
    hipLaunchKernelGGL(k_coszmn, dim3((ncol + 255) / 256), dim3(256), 0, (hipStream_t)stream, ncol, glat, glon, solhr,
                       dtswav, nrads, *sol, coszen, coszdg);
    return hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
}

int wr_rdtemp(int32_t im, int32_t jm, int32_t lm, int32_t i_lo, int32_t i_hi, int32_t j_lo, int32_t j_hi, int32_t ntsd,
              double dt, int32_t julday, int32_t julyr, int32_t ihrst, int32_t leap, const float *glat,
              const float *glon, const float *czmean, const float *rswtt, const float *rlwtt, float *t, float *czen,
              void *stream)
{
#pragma clang fp contract(off)
    if (im <= 0 || jm <= 0 || lm <= 0 || julday < 1 || julday > 366 || !glat || !glon || !czmean || !rswtt || !rlwtt ||
        !t || !czen)
        return WR_ERR_ARG;
    int month[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
    if (julyr % 4 == 0 || leap) month[1] = 29;
    int jmonth = 12, jday = julday, rem = julday;  // CAL_MON_DAY
    for (int k = 0; k < 12; ++k) {
        if (rem > month[k]) {
            rem -= month[k];
        } else {
            jday = rem;
            jmonth = k + 1;
            break;
        }
    }
    const Sun sun = zenith((double)ntsd * dt, ihrst, jmonth, jday, julyr, leap);
    const int n = im * jm;
    hipLaunchKernelGGL(k_rdtemp, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, im, jm, lm, i_lo, i_hi, j_lo,
                       j_hi, dt, sun, glat, glon, czmean, rswtt, rlwtt, t, czen);
    return hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
}

}  // extern "C"
