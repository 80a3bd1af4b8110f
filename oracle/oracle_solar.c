/* oracle_solar.c -- CPU restatement of the radiation's solar geometry (TEST INFRASTRUCTURE ONLY):
 * sol_update (the sun's position and the solar constant for a date), coszmn (the mean cosine of the
 * zenith angle over the shortwave interval) and rdtemp (the radiative temperature change of every
 * model step with the zenith angle of the step).  FP64 with the C library's trigonometry; the
 * library's versions agree with it to a few units in the last place. */
#include "oracle.h"

#define S_PI 3.1415926535897931
#define S_TPI (2.0 * S_PI)
#define S_DEGRAD (180.0 / S_PI)

/* iw3jdn: Julian day number of a Gregorian date (phys_rrtm_sol_update.cu:16-22) */
static int s_iw3jdn(int y, int m, int d)
{
    return d - 32075 + 1461 * (y + 4800 + (m - 14) / 12) / 4 + 367 * (m - 2 - (m - 14) / 12 * 12) / 12 -
           3 * ((y + 4900 + (m - 14) / 12) / 100) / 4;
}

/* solar (phys_rrtm_sol_update.cu:24-100): orbit of the sun (Kepler's equation by Newton iteration) */
static void s_solar(int jd, double fjd, double *r1, double *sollag, double *sindec, double *cosdec)
{
    const int jdor = 2415020;
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
    const double qq = deleqn * S_TPI / year;
    double e1 = 1.0, cd = 1.0;
    for (int iter = 0; cd > ccr && iter <= 10; ++iter) {
        const double ep = e1 - (e1 - ec * sin(e1) - qq) / (1.0 - ec * cos(e1));
        cd = fabs(e1 - ep);
        e1 = ep;
    }
    const double eq = 2.0 * atan(er * tan(0.5 * e1));
    const double dat = (double)(jd - jdor) - tpp + fjd;
    const double date = fmod(dat, year);
    const double em = S_TPI * date / year;
    e1 = 1.0;
    double cr = 1.0;
    for (int iter = 0; cr > ccr && iter <= 10; ++iter) {
        const double ep = e1 - (e1 - ec * sin(e1) - em) / (1.0 - ec * cos(e1));
        cr = fabs(e1 - ep);
        e1 = ep;
    }
    const double w1 = 2.0 * atan(er * tan(0.5 * e1));
    *r1 = 1.0 - ec * cos(e1);
    *sindec = sni * sin(w1 - eq);
    *cosdec = sqrt(1.0 - (*sindec) * (*sindec));
    const double dlt = asin(*sindec);
    double alp = asin(tan(dlt) * tini);
    if (cos(w1 - eq) < 0.0) alp = S_PI - alp;
    if (alp < 0.0) alp = alp + S_TPI;
    double sun = S_TPI * (date - deleqn) / year;
    if (sun < 0.0) sun = sun + S_TPI;
    *sollag = sun - alp - 0.03255;
}

/* sol_update_nmmb (phys_rrtm_sol_update.cu:102-160) */
int oracle_sol_update(int yr, int mon, int day, int hr, int mn, int sec, double deltsw, double deltim, double solc0,
                      wr_sol *o)
{
    if (!o || !(deltim > 0.0) || !(deltsw > 0.0)) return 1;
    int jd1 = s_iw3jdn(yr, mon, day);
    double fjd1;
    if (hr < 12) {
        jd1 = jd1 - 1;
        fjd1 = 0.5 + (double)hr / 24.0 + (double)mn / 1440.0 + (double)sec / 86400.0;
    } else {
        fjd1 = (double)(hr - 12) / 24.0 + (double)mn / 1440.0 + (double)sec / 86400.0;
    }
    fjd1 = fjd1 + (double)jd1;
    const int jd = (int)fjd1;
    const double fjd = fjd1 - (double)jd;
    double r1, sollag, sindec, cosdec;
    s_solar(jd, fjd, &r1, &sollag, &sindec, &cosdec);
    const int nswr = (int)round(deltsw / deltim);
    const double dtswh = deltsw / 3600.0;
    int nstp;
    if (deltsw >= 3600.0) {
        const int nn = imax(6, imin(12, (int)round(3600.0 / deltim)));
        nstp = (int)round(dtswh) * nn + 1;
    } else {
        nstp = imax(2, imin(20, nswr)) + 1;
    }
    o->sollag = sollag;
    o->sindec = sindec;
    o->cosdec = cosdec;
    o->solcon = solc0 / (r1 * r1);
    o->anginc = ((2.0 * asin(1.0)) / 12.0) * dtswh / (double)(nstp - 1);
    o->nstp = nstp;
    return 0;
}

/* coszmn_nmmb (phys_rrtm_coszmn.cu:18-62): the cosine of the zenith angle at nrads instants of the
 * SW interval, averaged over the sunlit instants (coszen) and over all of them (coszdg) */
void oracle_coszmn(int ncol, const double *glat, const double *glon, double solhr, double dtswav, int nrads,
                   const wr_sol *s, double *coszen, double *coszdg)
{
    const double pid12 = (2.0 * asin(1.0)) / 12.0;
    const double solang = pid12 * (solhr - 12.0);
    const double rnstp = dtswav / (double)nrads;
    for (int i = 0; i < ncol; ++i) {
        const double sinlat = sin(glat[i]), coslat = cos(glat[i]);
        double cz = 0.0;
        int istsun = 0;
        for (int it = 1; it <= nrads; ++it) {
            const double cns = solang + (double)(it - 1) * pid12 * rnstp + s->sollag;
            const double coszn = sinlat * s->sindec + coslat * s->cosdec * cos(cns + glon[i]);
            cz += fmax(0.0, coszn);
            if (coszn > 1.0e-4) istsun += 1;
        }
        coszdg[i] = cz * rnstp;
        coszen[i] = istsun > 0 ? cz / (double)istsun : cz;
    }
}

/* rdtemp (drv_rdtemp.cu:10-145): the sun's declination and hour angle for the step (ZENITH), the
 * column's cosine of the zenith angle czen, and T += (RSWTT czen / CZMEAN + RLWTT) dt on the columns
 * i_lo..i_hi, j_lo..j_hi (0-based, inclusive; czen on every column) */
static void s_zenith_header(double times, int ihrst, int m, int d, int y, int leap, double *dec, double *hrang)
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
    const double datjul = startyr + day;
    const double difjd = datjul - 2451545.0;
    const double slonm = (280.460 + 0.9856474 * difjd) * deg2rd + yfctr * S_TPI;
    const double anom = (357.528 + 0.9856003 * difjd) * deg2rd;
    double slon = slonm + (1.915 * sin(anom) + 0.020 * sin(2.0 * anom)) * deg2rd;
    if (slon > S_TPI) slon = slon - S_TPI;
    *dec = asin(sin(slon) * sin(obliq));
    double ra = acos(cos(slon) / cos(*dec));
    if (slon > S_PI) ra = S_TPI - ra;
    const double datj0 = startyr + dayi - 1.0;
    const double tu = (datj0 - 2451545.0) / 36525.0;
    const double stim0 = 24110.54841 + tu * (8640184.812866 + 9.3104e-2 * tu + -6.2e-6 * tu * tu);
    double sidtim = stim0 / 3600.0 + yfctr * 24.0 + 1.00273791 * hour;
    sidtim = sidtim * 15.0 * deg2rd;
    if (sidtim < 0.0) sidtim = sidtim + S_TPI;
    if (sidtim > S_TPI) sidtim = sidtim - S_TPI;
    *hrang = sidtim - ra;
}

int oracle_rdtemp(int im, int jm, int lm, int i_lo, int i_hi, int j_lo, int j_hi, int ntsd, double dt, int julday,
                  int julyr, int ihrst, int leap, const float *glat, const float *glon, const float *czmean,
                  const float *rswtt, const float *rlwtt, float *t, float *czen)
{
    /* CAL_MON_DAY: month and day of the year's day julday */
    int month[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
    if (julyr % 4 == 0 || leap) month[1] = 29;
    int jmonth = 12, jday = julday, rem = julday;
    for (int k = 0; k < 12; ++k) {
        if (rem > month[k]) {
            rem -= month[k];
        } else {
            jday = rem;
            jmonth = k + 1;
            break;
        }
    }
    double dec, hrang;
    s_zenith_header((double)ntsd * dt, ihrst, jmonth, jday, julyr, leap, &dec, &hrang);
    for (int j = 0; j < jm; ++j)
        for (int i = 0; i < im; ++i) {
            const size_t ij = (size_t)j * im + i;
            const double hrlcl = hrang + (double)glon[ij] + S_TPI;
            double sinalt = sin(dec) * sin((double)glat[ij]) + cos(dec) * cos(hrlcl) * cos((double)glat[ij]);
            if (sinalt < 0.0) sinalt = 0.0;
            czen[ij] = (float)sinalt;
            if (i < i_lo || i > i_hi || j < j_lo || j > j_hi) continue;   /* T on the inner columns only */
            const double factr = czmean[ij] > 0.0f ? sinalt / (double)czmean[ij] : 0.0;
            for (int l = 0; l < lm; ++l) {
                const size_t p = ((size_t)l * jm + j) * im + i;
                t[p] = (float)((double)t[p] + ((double)rswtt[p] * factr + (double)rlwtt[p]) * dt);
            }
        }
    return 0;
}
