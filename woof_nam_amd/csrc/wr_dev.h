// wr_dev.h -- device data structures and per-layer physics of the MI355X RRTMG kernels.
//
// Layout in HBM: every input/output field is column fastest ([level][column]); the kernels map
// one thread to one column, so each wavefront load of a field is one contiguous 512-byte (FP64)
// segment.  Everything per layer (driver prep, setcoef, Planck source, cloud optics) is rebuilt
// on the fly inside the band kernels from the raw column inputs: it costs a few hundred VALU
// operations per layer, far less than re-reading a materialised per-layer record for every band.
//
// Physics follows the reference CUDA kernels it cites (recastsystems/woof-nam namgpu/kernels/),
// evaluated in the same order of operations; the layer index k is 0-based SURFACE-FIRST (the
// reference's lwrad/swrad internal order), kin = nrl-1-k the top-first radiation layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wr_math.h"

namespace wr {

constexpr int NL = 66;  // max radiation layers
constexpr int NTBL = 10000;

// physcons (phys_rrtm_lw_lwrad.cu:55-62, phys_rrtm_sw_swrad.cu:22-34)
constexpr double CON_G = 9.80665;
constexpr double CON_AVGD = 6.0221415e23;
constexpr double CON_AMD = 28.9644;
constexpr double CON_AMW = 18.0154;
constexpr double CON_AMO3 = 47.9982;
constexpr double CON_CP = 1.0046e3;
constexpr double AMDW = CON_AMD / CON_AMW;
constexpr double AMDO3 = CON_AMD / CON_AMO3;
constexpr double EPS = 1.0e-6;
constexpr double ONEMINUS = 1.0 - 1.0e-6;
constexpr double FTINY = 1.0e-12;
constexpr double STPFAC = 296.0 / 1013.0;
constexpr double HEATFAC = CON_G * 1.0e-2 / CON_CP;  // K/s per (W/m2 / hPa)
constexpr double FLUXFAC = 3.141592653589793 * 2.0e4;
constexpr double S0 = 1368.22;

// Device code reads and writes HBM through pointers held in structs (tables, column fields, work
// space).  Declared in the global address space in the device pass, so that their accesses compile
// to global loads and stores: through a generic pointer they become flat accesses, which also
// count against the LDS/scalar wait counter and so serialise the waits of every scalar load
// behind the outstanding table gathers.  The host pass sees plain pointers of the same size.
#if defined(__HIP_DEVICE_COMPILE__)
#define WR_G __attribute__((address_space(1)))
#else
#define WR_G
#endif
// a device allocation stored into a WR_G member (host code)
template <typename T> inline WR_G T *gptr(T *p) { return (WR_G T *)p; }

__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }

// Entry i of the LW/SW look-up tables (transmittance e, Pade source factor tfn, optical depth tau at
// tblind = i / 10000; the FP64 tables are built by k_luts), evaluated in FP32 where it is needed: the
// FP32 kernels spend the few operations instead of three dependent table gathers per g-point.  The
// FP64 kernels read the tables, which the CPU oracle reproduces bit for bit.
struct LutF {
    float e, tfn, tau;
};
__device__ __forceinline__ LutF lut_f32(int i, bool lw_cap)
{
    if (i <= 0) return {1.0f, 0.0f, 0.0f};
    if (i >= NTBL) return {lw_cap ? 1.0e-20f : 0.0f, 1.0f, 1.0e10f};
    const float tau = (float)(1.0 / 0.278) * ((float)i / (float)(NTBL - i));
    const float e = (lw_cap && tau >= 87.0f) ? 1.0e-20f : M<float>::exp(-tau);
    const float tfn = (tau < 0.06f) ? tau / 6.0f : 1.0f - 2.0f * ((1.0f / tau) - (e / (1.0f - e)));
    return {e, tfn, tau};
}


// ---- tables --------------------------------------------------------------------------------
// LDS (the device pass's address space 3), for tables a kernel stages there
#if defined(__HIP_DEVICE_COMPILE__)
#define WR_L __attribute__((address_space(3)))
#else
#define WR_L
#endif
// The tables of one LW band; P is the pointer type (global by default, LDS in the staged kernels),
// PB that of absb.
template <typename R, typename P = const WR_G R *, typename PB = P> struct LwBandT {
    P absa;
    PB absb;
    P selfref, forref, fracrefa, fracrefb;
    P ka_m[3], kb_m[3];
    P x1, x2;
    R refrat_planck_a, refrat_planck_b, refrat_m_a, refrat_m_b, refrat_m_a3;
};
template <typename R> using LwBand = LwBandT<R>;
// The same tables packed into one image: 14 tables (absa absb selfref forref fracrefa fracrefb
// ka_m[0..2] kb_m[0..2] x1 x2) at element offsets off[] (multiples of 4, -1 where the image does not
// hold one), n elements in all (a multiple of 4), from a 256-byte aligned base.  The FP64 image leaves
// absb out.
constexpr int LW_IMG_T = 14;
template <typename R> struct LwImg {
    const WR_G R *base;
    int n;
    int off[LW_IMG_T];
};
// The tables of one SW band (pointer types as in LwBandT) and their packed image: 12 tables (absa
// absb selfref forref rayl_v rayla raylb abso3a abso3b absch4 absh2o absco2), the FP64 image again
// without absb.
template <typename R, typename P = const WR_G R *, typename PB = P> struct SwBandT {
    P absa;
    PB absb;
    P selfref, forref;
    P rayl_v, rayla, raylb, abso3a, abso3b, absch4, absh2o, absco2;
    R rayl, strrat;
};
template <typename R> using SwBand = SwBandT<R>;
constexpr int SW_IMG_T = 12;
template <typename R> struct SwImg {
    const WR_G R *base;
    int n;
    int off[SW_IMG_T];
};
template <typename R> struct Tabs {
    // index arithmetic stays FP64 in every precision
    const WR_G double *preflog, *tref, *chi_mls, *sw_preflog, *sw_tref;
    const WR_G R *totplnk, *delwave, *absliq1, *absice1;
    const WR_G int32_t *ipat, *layreffr, *ix1, *ix2, *ibx;
    const WR_G R *exp_tbl, *tfn_tbl, *tau_tbl, *sw_exp_tbl;
    const WR_G double *sw_exp_tbl64;  // FP64 SW exp table (layers that take the FP64 R/T branch)
    const WR_G R *sf01, *sf02, *sf03, *sw_strrat, *sw_specwt;
    const WR_G R *extliq1, *ssaliq1, *asyliq1, *extice3, *ssaice3, *asyice3;
    R scalekur;
    LwBand<R> lw[16];
    SwBand<R> sw[14];
    LwImg<R> lw_img[16];
    SwImg<R> sw_img[14];
};

// ---- column inputs ---------------------------------------------------------------------------
// Every member has a default (null, 0), so a ColIn made any way holds no indeterminate pointer that a
// kernel would test and follow.
struct ColIn {
    const WR_G double *p_lev = nullptr, *t_lay = nullptr, *q_lay = nullptr, *o3_lay = nullptr, *cldf = nullptr;
    const WR_G double *qc = nullptr, *qi = nullptr, *t_sfc = nullptr, *emis = nullptr, *alb = nullptr, *cosz = nullptr;
    const WR_G uint64_t *col_id = nullptr;
    const WR_G uint64_t *seed_dev = nullptr;  // optional device copy of the McICA seed
    const WR_G int32_t *ncol_dev = nullptr;   // optional device count of valid columns (a shorter replayed tile)
    const WR_G double *cwp = nullptr, *cip = nullptr, *crp = nullptr, *rew = nullptr, *rei = nullptr;  // path form
    const WR_G float *aer_sw = nullptr, *aer_lw = nullptr;  // optional aerosol optics [3][bands][nlay+1][aer_ld]
    long aer_ld = 0;
    const WR_G double *o3_top = nullptr;                    // optional ozone of the added top layer
    int ncol = 0;  // stride of the [k][column] fields
    int nlay = 0;  // model layers
};

// True when a call carries any of the optional inputs (path-form clouds, aerosol optics, the added
// top layer's ozone).
inline bool has_ext(const ColIn &in) { return in.cwp || in.aer_sw || in.aer_lw || in.o3_top; }

// The column inputs as a band kernel reads them.  EXT = false nulls the optional inputs at compile
// time, so their tests fold away and the kernel's registers carry none of their terms; such an
// instance computes what the generic one (EXT = true) computes when those pointers are null.
template <bool EXT> __device__ __forceinline__ ColIn band_in(ColIn in)
{
    if constexpr (!EXT) {
        in.cwp = in.cip = in.crp = in.rew = in.rei = nullptr;
        in.aer_sw = in.aer_lw = nullptr;
        in.o3_top = nullptr;
    }
    return in;
}

struct RunPar {
    int mcica, iovr;
    uint64_t seed;
    double solcon, re_liq, re_ice;
    double vmr[10];
};

// Radiation-grid accessors of one column (grrad_setup phys_rrtm_grrad_setup.cu:102,
// grrad_gases .cu:56): layer kin top-first, kin = 0 is the layer added above the model top.
// columns of the tile [col0, col0 + n) that hold data: all of them, or up to *ncol_dev
__device__ __forceinline__ int valid_cols(const ColIn &in, int col0, int n)
{
    if (!in.ncol_dev) return n;
    const int v = *in.ncol_dev - col0;
    return v < 0 ? 0 : (v < n ? v : n);
}

// The inputs of one radiation layer l as the layer physics uses them: interface pressures plvl(l),
// plvl(l+1), layer pressure, temperature, water vapour, ozone and progcld2's cloud triple.  The band
// kernels load the next layer's values one iteration ahead, so that their HBM latency overlaps the
// current layer's work.
struct LayV {
    double pt, pb, pl, t, q, o3, ct, cwp, cip;
};

struct RadCol {
    ColIn in;  // the call's field pointers, held by value (uniform, in scalar registers)
    int i;    // column index into the fields
    int nrl;  // radiation layers = model layers + 1
    __device__ __forceinline__ double fld(const WR_G double *a, int k) const { return a[(size_t)k * in.ncol + i]; }
    __device__ __forceinline__ double plvl(int l) const
    {
        if (l == 0) return 1.0e-6;
        const double v = 10.0 * (fld(in.p_lev, l - 1) * 1.0e-3);
        return (l == 1 && v <= 1.0e-6) ? 2.0e-6 : v;
    }
    __device__ __forceinline__ double plyr(int l) const
    {
        if (l == 0) return 0.5 * plvl(1);
        const double prsl = 0.5 * (fld(in.p_lev, l - 1) + fld(in.p_lev, l));
        return 10.0 * (prsl * 1.0e-3);
    }
    __device__ __forceinline__ double tlyr(int l) const { return fld(in.t_lay, l == 0 ? 0 : l - 1); }
    __device__ __forceinline__ double qlyr(int l) const { return fmax(1.0e-7, fld(in.q_lay, l == 0 ? 0 : l - 1)); }
    __device__ __forceinline__ double olyr(int l) const
    {
        if (l == 0 && in.o3_top) return fmax(1.0e-10, in.o3_top[i]);
        return fmax(1.0e-10, fld(in.o3_lay, l == 0 ? 0 : l - 1));
    }
    __device__ __forceinline__ double tskn() const { return in.t_sfc[i]; }
    // level temperature by log-pressure interpolation; la(l) = log(plyr(l)), lb(l) = log(plvl(l))
    __device__ __forceinline__ double tlvl(int l, double la_l, double la_lm1, double lb_l) const
    {
        if (l == 0) return tlyr(0);
        if (l == nrl) return tskn();
        const double t = tlyr(l), tm = tlyr(l - 1);
        return t + (tm - t) * (lb_l - la_l) / (la_lm1 - la_l);
    }
    // cloud fraction alone (progcld2: ct < 0.001 -> 0); the added top layer is clear
    __device__ __forceinline__ double cfrac(int l) const
    {
        if (l == 0) return 0.0;
        const double ct = fld(in.cldf, l - 1);
        return (ct < 0.001) ? 0.0 : ct;
    }
    // progcld2 with the cloud fraction given (phys_rrtm_progcld2.cu:92-255): (frac, cwp, cip)
    __device__ __forceinline__ void cloud(int l, double &ct, double &cwp, double &cip) const
    {
        ct = 0.0, cwp = 0.0, cip = 0.0;
        if (l == 0) return;
        const int k0 = l - 1;
        ct = fld(in.cldf, k0);
        if (in.cwp) {  // path form: progcld2's paths as given
            if (ct < 0.001) {
                ct = 0.0;
                return;
            }
            cwp = fld(in.cwp, k0);
            cip = fld(in.cip, k0);
            return;
        }
        const double qc = fld(in.qc, k0), qi = fld(in.qi, k0);
        double qcwat, qcice;
        if (tlyr(l) - 273.15 > -40.0) {
            qcwat = qc;
            qcice = qi;
        } else {
            qcwat = 0.0;
            qcice = qc + qi;
        }
        if (qcwat + qcice > 1.0e-12) {
            const double cpath = fabs(plvl(l + 1) - plvl(l)) * (100000.0 / CON_G);
            if (qcwat > 0.0) cwp = cpath * qcwat;
            cip = cpath * qcice;
        }
        if (ct < 0.001) {
            ct = 0.0;
            cwp = 0.0;
            cip = 0.0;
        }
    }
    // effective radii (micron) and rain path (g/m2) of cloudy layer l: the call's radii and no rain, or
    // the path form's per-layer values
    __device__ __forceinline__ void props(int l, double re_liq, double re_ice, double &rl, double &ri, double &rp) const
    {
        rl = re_liq, ri = re_ice, rp = 0.0;
        if (in.cwp && l > 0) {
            rl = fld(in.rew, l - 1);
            ri = fld(in.rei, l - 1);
            if (in.crp) rp = fld(in.crp, l - 1);
        }
    }
    // aerosol of radiation layer l (top-first) in LW band b (0-based): the absorption optical depth
    // tau (1 - ssa) that lwrad adds to every g-point's gas optical depth (phys_rrtm_lw_lwrad.cu:214-219)
    __device__ __forceinline__ double aer_lw_tau(int b, int l) const
    {
        if (!in.aer_lw) return 0.0;
        const size_t ld = (size_t)in.aer_ld, nl = (size_t)nrl;
        const double ta = (double)in.aer_lw[((size_t)b * nl + l) * ld + i];
        const double sa = (double)in.aer_lw[((size_t)(16 + b) * nl + l) * ld + i];
        return ta * (1.0 - sa);
    }
    // aerosol optical depth, single-scattering albedo and asymmetry of radiation layer l in SW band
    // b (0 = band 16), swrad's tauae / ssaae / asyae (phys_rrtm_sw_swrad.cu:224-229)
    __device__ __forceinline__ void aer_sw3(int b, int l, double &ta, double &sa, double &ya) const
    {
        ta = 0.0, sa = 0.0, ya = 0.0;
        if (!in.aer_sw) return;
        const size_t ld = (size_t)in.aer_ld, nl = (size_t)nrl;
        ta = (double)in.aer_sw[((size_t)b * nl + l) * ld + i];
        sa = (double)in.aer_sw[((size_t)(14 + b) * nl + l) * ld + i];
        ya = (double)in.aer_sw[((size_t)(28 + b) * nl + l) * ld + i];
    }
    __device__ __forceinline__ LayV lay(int l) const
    {
        LayV v;
        v.pt = plvl(l), v.pb = plvl(l + 1), v.pl = plyr(l), v.t = tlyr(l), v.q = qlyr(l), v.o3 = olyr(l);
        cloud(l, v.ct, v.cwp, v.cip);
        return v;
    }
};
// tlvl of RadCol from the values of the layers below (t, la_l) and above (tm, la_lm1) an interface
__device__ __forceinline__ double tlvl_v(double t, double tm, double lb_l, double la_l, double la_lm1)
{
    return t + (tm - t) * (lb_l - la_l) / (la_lm1 - la_l);
}

// ---- McICA sub-columns: phys_rrtm_lw_mcica_subcol.cu:105-168 draw order and max-random rule,
// Philox4x32-10 stream keyed (seed, column), counter (column, g | kind << 16, k / 4) ------------
template <typename CF> __device__ __forceinline__ uint64_t mcica_g(uint64_t col, uint64_t seed, int kind, int g,
                                                                   int nrl, int iovr, CF cfk)
{
    uint64_t bits = 0;
    double prev = 0.0, cfm1 = 0.0;
    uint32_t out[4] = {0, 0, 0, 0};
    for (int k = 0; k < nrl; ++k) {
        if ((k & 3) == 0 && (iovr != 2 || k == 0))
            philox4x32_10((uint32_t)col, (uint32_t)(col >> 32), (uint32_t)g | ((uint32_t)kind << 16), (uint32_t)(k >> 2),
                          (uint32_t)seed, (uint32_t)(seed >> 32), out);
        const uint32_t u = out[iovr == 2 ? 0 : (k & 3)];
        const double r = (double)u * 0x1p-32;
        const double cf = cfk(k);
        double cd;
        if (iovr == 1) {
            if (k == 0)
                cd = r;
            else {
                const double tem1 = 1.0 - cfm1;
                cd = (prev > tem1) ? prev : r * tem1;
            }
            prev = cd;
        } else {
            cd = r;
        }
        if (cd >= 1.0 - cf) bits |= (uint64_t)1 << k;
        cfm1 = cf;
    }
    return bits;
}

}  // namespace wr
