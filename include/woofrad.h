/* woofrad.h -- C ABI of the MI355X RRTMG radiation library (libwoofrad.so).
 *
 * Plain pointers, sizes and integer status codes; no torch or HIP types in the signatures
 * (a HIP stream is passed as `void *`).  All field arrays are COLUMN FASTEST: element
 * (k, i) of a [nk][ncol] field is at x[k * ncol + i].  Vertical index 0 is the model top.
 *
 * Reference interfaces this ABI replaces (recastsystems/woof-nam):
 *   - the radiation hook `rrtm_call(ctx, args: dict) -> None` that
 *     namgpu/physics/driver/phys_rad_driver.py:255-262 calls on radiation steps (default
 *     namgpu/physics/rrtm/rrtm_call.py:246 `run`).  The Python shim woof_nam_amd.rrtm_call.run keeps
 *     that signature and calls wr_run_* below on the column form of the state;
 *   - the chain under it: driver prep / grrad setup / gases / clouds
 *     (phys_rrtm_driver_call.cu:83, phys_rrtm_grrad_setup.cu:102, phys_rrtm_grrad_gases.cu:56,
 *     phys_rrtm_progcld2.cu:92), lwrad (phys_rrtm_lw_lwrad.cu:98,316), cldprop, setcoef,
 *     taugb01..16, rtrnmc / rtrnmr, swrad (phys_rrtm_sw_swrad.cu:127,319), setcoef, taumol16..29,
 *     spcvrtc, and the output copy phys_rrtm_driver_post.cu:71;
 *   - phys_rad_cal_cldfra.cu:29 (wr_cal_cldfra) and the McICA generators
 *     phys_rrtm_lw_mcica_subcol.cu:105 / phys_rrtm_sw_mcica_subcol.cu:129 (wr_mcica_masks).
 */
#ifndef WOOFRAD_H
#define WOOFRAD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WR_NBANDS_LW 16
#define WR_NGPT_LW 140
#define WR_NBANDS_SW 14
#define WR_NGPT_SW 112

/* status codes; the Python shim maps them onto the reference's named exceptions */
enum {
    WR_OK = 0,
    WR_ERR_ARG = 1,        /* bad size / null pointer          -> ValueError          */
    WR_ERR_TABLES = 2,     /* table file missing or malformed  -> FileNotFoundError   */
    WR_ERR_HIP = 3,        /* HIP runtime error                -> RuntimeError        */
    WR_ERR_NOMEM = 4,      /* device allocation failed         -> MemoryError         */
    WR_ERR_UNSUPPORTED = 5,/* option not implemented           -> NotImplementedError */
    WR_ERR_TOO_LARGE = 6   /* call beyond a static work space  -> ValueError          */
};

/* Column inputs (model layers nlay, interfaces nlay+1; index 0 = model top). */
typedef struct {
    int32_t ncol;
    int32_t nlay;            /* model layers; the radiation adds one layer above the model top */
    const double *p_lev;     /* [nlay+1][ncol] interface pressure, Pa (increasing downward)    */
    const double *t_lay;     /* [nlay][ncol]   temperature, K                                    */
    const double *q_lay;     /* [nlay][ncol]   specific humidity, kg/kg                           */
    const double *o3_lay;    /* [nlay][ncol]   ozone mass mixing ratio, kg/kg                     */
    const double *cldf;      /* [nlay][ncol]   cloud fraction, 0..1                               */
    const double *qc;        /* [nlay][ncol]   cloud liquid water mixing ratio, kg/kg             */
    const double *qi;        /* [nlay][ncol]   cloud ice mixing ratio, kg/kg                      */
    const double *t_sfc;     /* [ncol] skin temperature, K                                        */
    const double *emis;      /* [ncol] surface LW emissivity (all 16 bands)                      */
    const double *alb;       /* [ncol] surface SW albedo (beam/diffuse, visible/near-IR)          */
    const double *cosz;      /* [ncol] cosine of the solar zenith angle (night when < 1e-4)      */
    const uint64_t *col_id;  /* [ncol] global column index that keys the McICA stream; NULL = i  */
    const uint64_t *seed_dev;/* optional device pointer to the McICA seed; when set it replaces
                                wr_params.seed (lets a captured HIP graph be replayed per request) */
    const int32_t *ncol_dev; /* optional device pointer to the number of valid columns (<= ncol);
                                columns from it on are skipped and their outputs left untouched
                                (lets one graph captured for ncol columns serve shorter tiles)    */
    /* Optional path form of the clouds, progcld2's outputs (wr_nmmb_prep writes them).  When cwp is
     * set, each layer's cloud is cldf with these paths and radii, and qc, qi and wr_params.re_liq /
     * re_ice are not read (qc, qi may be NULL); cip, rew and rei are then required, crp optional. */
    const double *cwp;       /* [nlay][ncol] cloud liquid water path, g/m2                        */
    const double *cip;       /* [nlay][ncol] cloud ice path, g/m2                                 */
    const double *crp;       /* [nlay][ncol] rain path, g/m2 (NULL: none)                          */
    const double *rew;       /* [nlay][ncol] liquid effective radius, micron                      */
    const double *rei;       /* [nlay][ncol] ice effective radius, micron                         */
    /* Optional aerosol optical properties per radiation layer, in the layout setaer writes
     * (phys_rrtm_setaer.cu:20-27, float32): [field][band][layer][aer_ld] with field 0 the optical
     * depth, 1 the single-scattering albedo, 2 the asymmetry factor; layer 0 the layer added above
     * the model top, then the model layers top-first (nlay + 1 layers); column i of the call at
     * entry i of the last index.  aer_sw holds SW bands 16..29 (14), aer_lw LW bands 1..16 (16; the
     * LW adds tau (1 - ssa) to the gas optical depth, lwrad).  aer_ld is the column stride, >= ncol
     * (0: ncol), so the arrays may be a slice of a wider grid.  NULL: no aerosol. */
    const float *aer_sw;
    const float *aer_lw;
    int64_t aer_ld;
    const double *o3_top;    /* optional [ncol] ozone mass mixing ratio of the layer the radiation adds
                                above the model top (climatological ozone); NULL: the top layer's   */
} wr_columns;

/* Call parameters (NAM values by default; see wr_default_params). */
typedef struct {
    int32_t mcica;           /* 1: McICA sub-columns (rtrnmc + masked SW two-stream)
                                0: NAM route, max-random fractional cloud (rtrnmr + spcvrtc)     */
    int32_t iovr;            /* cloud overlap of the McICA generator: 0 random, 1 max-random, 2 max */
    uint64_t seed;           /* McICA stream key                                                  */
    double solcon;           /* solar constant, W/m2                                              */
    double re_liq;           /* cloud liquid effective radius, micron                             */
    double re_ice;           /* cloud ice effective radius, micron                                */
    double vmr[10];          /* co2, n2o, ch4, o2, co, cfc11, cfc12, cfc22, ccl4, cfc113 (mol/mol) */
} wr_params;

/* Outputs.  Levels are the radiation levels, nlay+2 of them, index 0 = top of the added layer
 * (1e-6 hPa), index l >= 1 = model interface l-1.  Heating rates are on the model layers, K/s. */
typedef struct {
    double *hr_lw;           /* [nlay][ncol]   LW heating rate, K/s                             */
    double *hr_sw;           /* [nlay][ncol]   SW heating rate, K/s (0 at night)                */
    double *flw_up;          /* [nlay+2][ncol] LW upward flux, W/m2                              */
    double *flw_dn;          /* [nlay+2][ncol] LW downward flux                                  */
    double *fsw_up;          /* [nlay+2][ncol] SW upward flux (0 at night)                       */
    double *fsw_dn;          /* [nlay+2][ncol] SW downward flux                                  */
    double *fsw_dn_sfc_clr;  /* [ncol] optional (NULL: not written): clear-sky surface SW downward
                                flux of the NAM route (mcica = 0); 0 on the McICA route and at night */
} wr_outputs;

/* Optional diagnostics for bit-exact checks (any pointer may be NULL).  Radiation layers in the
 * reference's internal SURFACE-FIRST order, nrl = nlay+1 of them. */
typedef struct {
    int32_t *lw_jp;          /* [nrl][ncol] */
    int32_t *lw_jt;
    int32_t *lw_jt1;
    int32_t *lw_laytrop;     /* [ncol] */
    int32_t *sw_jp;          /* [nrl][ncol], -1 on night columns */
    int32_t *sw_jt;
    int32_t *sw_jt1;
    int32_t *sw_laytrop;     /* [ncol] */
    uint32_t *lw_mask;       /* [nrl][5][ncol] McICA bits of the 140 LW g-points */
    uint32_t *sw_mask;       /* [nrl][4][ncol] McICA bits of the 112 SW g-points (0 at night) */
} wr_diag;

typedef struct wr_ctx wr_ctx;

/* Fill NAM default parameters: McICA on, max-random, solcon 1368.22, radii 10/50 micron,
 * CO2 410, N2O 0.33, CH4 1.9 ppmv, O2 0.209 and the reference's default minor-gas VMRs
 * (phys_rrtm_getgases.cu:56-65). */
void wr_default_params(wr_params *p);

/* Create a context on `device`: loads the table file once and keeps it device resident,
 * allocates the work space for tiles of up to `max_cols` columns.  No later call allocates. */
int wr_create(wr_ctx **out, const char *table_path, int device, int32_t max_cols);

/* Device bytes wr_create allocates for a context of `max_cols` columns (tables, tile work space and
 * the band streams' level pools, under the same WR_POOL_GB / WR_BAND_STREAMS settings), computed on
 * the host without allocating: callers size max_cols from free device memory with it, as
 * namgpu/physics/rrtm/rrtm_call.py:183-203 (batch_count) sizes its batches.  0 for max_cols <= 0. */
uint64_t wr_context_bytes(int32_t max_cols);

/* Plain FP64 radiation call on device pointers.  Columns beyond max_cols run in tiles. */
int wr_run_fp64(wr_ctx *ctx, const wr_columns *in, const wr_params *par, const wr_outputs *out,
                const wr_diag *diag, void *stream);

/* Mixed precision: FP32 gas optics, Planck source and two-stream sweeps; FP64 for the SW
 * reflectance/transmittance where w0 > 0.999999, for level-flux differencing into heating rates
 * and for pressure/temperature index arithmetic.  Same inputs/outputs as wr_run_fp64. */
int wr_run_mixed(wr_ctx *ctx, const wr_columns *in, const wr_params *par, const wr_outputs *out,
                 const wr_diag *diag, void *stream);

/* wr_run_mixed that allocates nothing, not even through its context: no counterpart in the
 * reference.  For callers that cannot allocate device memory at call time (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  Same inputs, outputs and results
 * as wr_run_mixed (the same kernels), but the tile work space is a static device array of the library
 * (219 MB per device, held until the process ends) instead of the context's, and the band
 * kernels run on the caller's stream.  At most WR_NOALLOC_MAX_COLS columns per call (more:
 * WR_ERR_TOO_LARGE) and 63 model layers.  The context supplies the tables; its own work space is not
 * touched.  Every call on one device shares that device's array: calls that overlap on one device
 * (other streams, other threads, other contexts) race and must be serialised by the caller. */
#define WR_NOALLOC_MAX_COLS 4096
int wr_run_mixed_noalloc(wr_ctx *ctx, const wr_columns *in, const wr_params *par, const wr_outputs *out,
                         void *stream);

/* One band alone, for per-band error budgets (no counterpart in the reference): as wr_run_fp64
 * (mixed = 0) or wr_run_mixed (mixed = 1), but only band `band` contributes to the fluxes and
 * heating rates: 0..15 = LW bands 1..16 (SW outputs 0), 16..29 = SW bands 16..29 (LW outputs 0). */
int wr_run_band(wr_ctx *ctx, const wr_columns *in, const wr_params *par, const wr_outputs *out,
                int32_t band, int32_t mixed, void *stream);

/* CAL_CLDFRA (phys_rad_cal_cldfra.cu:29-55): cldfra = 1 where qc(+qi) > 1e-6, else 0, on the
 * interior points i in [i_lo, i_hi], j in [j_lo, j_hi] of [lm][jm][im] float32 fields; other
 * points untouched.  f_qc / f_qi are the microphysics flags. */
int wr_cal_cldfra(int32_t im, int32_t jm, int32_t lm, int32_t i_lo, int32_t i_hi, int32_t j_lo,
                  int32_t j_hi, int32_t f_qc, int32_t f_qi, float *cldfra, const float *qc,
                  const float *qi, void *stream);

/* McICA sub-column masks alone (counter-based Philox4x32-10 stream keyed by seed and column):
 * cldf [nrl][ncol] surface-first; lw_mask [nrl][5][ncol], sw_mask [nrl][4][ncol] (either may
 * be NULL).  Bits equal those wr_run_* use. */
int wr_mcica_masks(int32_t ncol, int32_t nrl, const double *cldf, const uint64_t *col_id,
                   uint64_t seed, int32_t iovr, uint32_t *lw_mask, uint32_t *sw_mask, void *stream);

/* ---- NMMB-state side of the call (the hook's prep and post, on the device) ----------------------
 *
 * wr_nmmb_prep turns the model state of the radiated rows into the column inputs above, with the
 * cloud properties in the path form (wr_columns.cwp ...).  It restates, per column:
 *   - driver prep (phys_rrtm_driver_prep.cu:82-98): interface pressures P8W = PT + cumulative
 *     (PDSG1 + DSG2 PD) and layer pressures P_PHY = SGML2 PD + PSGML1, in single precision as there;
 *   - driver call (phys_rrtm_driver_call.cu:200-275): total condensate CW with the Ferrier fractions
 *     F_ICE, F_RAIN, and the subgrid cloud water of layers without condensate at RH > 95 % (SUBGRID,
 *     the saturation pressure FPVS of module_MP_FER_HIRES.F90:2874-2971);
 *   - grrad setup (phys_rrtm_grrad_setup.cu:102-215): the water-vapour floor 1e-7, the virtual
 *     temperature, the layer above the model top (cloud free);
 *   - progcld2 (phys_rrtm_progcld2.cu:92-359) in the NAM configuration (top-first, no lcrick / lcnorm
 *     / lnoprec / lsashal, condensate cloud fraction): ice / liquid / rain split, cloud fraction, the
 *     paths of cloud water, ice and rain (g/m2), the diagnosed liquid and ice effective radii, and
 *     gethml's low / middle / high / total / boundary-layer cloud cover with their top and base layers.
 * The state is single precision; the prep evaluates in FP64 (the CPU oracle's operation sequence).
 * wr_nmmb_post is driver post (phys_rrtm_driver_post.cu:71-150): RLWTT, RSWTT, the surface and top
 * fluxes, CZMEAN, SIGT4, and the cloud outputs CLDFRA (progcld2's layer cloud fraction), CFRACL/M/H
 * with the ACFRST/NCFRST (and ACFRCV/NCFRCV) accumulators. */
typedef struct {
    int32_t im, jm, lm;      /* fields [lm][jm][im] (level 0 = model top), surface fields [jm][im]    */
    int32_t j0, j1;          /* radiated rows j0 <= j < j1 (0-based); column c = (j - j0) * im + i     */
    float pt;                /* pressure at the model top, Pa                                         */
    const float *pd;         /* [jm][im] PD, Pa                                                       */
    const float *pdsg1, *dsg2, *psgml1, *sgml2;  /* [lm] hybrid coordinate (pressure and sigma parts) */
    const float *t, *q, *o3; /* [lm][jm][im] temperature K, specific humidity, ozone (kg/kg)          */
    const float *cw;         /* [lm][jm][im] total condensate CW, kg/kg                               */
    const float *f_ice;      /* [lm][jm][im] ice fraction of the condensate (Ferrier F_ICE)           */
    const float *f_rain;     /* [lm][jm][im] rain fraction of the liquid (Ferrier F_RAIN)             */
    const float *glat;       /* [jm][im] latitude, radians                                            */
} wr_nmmb_state;

/* Climatological ozone (getozn, phys_rrtm_getozn.cu:44-111, the NAM's ntoz = 0 branch): the monthly
 * zonal-mean climatology gas_init reads, interpolated in latitude, between two months and in the Exner
 * function of pressure. */
typedef struct {
    const float *o3r;        /* [12][loz][jmr] ozone mass mixing ratio (device; ppmv x 1.655e-6)    */
    const float *pstr;       /* [loz] its pressure levels, hPa, top first (device)                 */
    int32_t loz, jmr;        /* levels (<= 64) and latitudes                                       */
    float blte, dlte;        /* first latitude and latitude spacing, degrees                       */
    int32_t k1oz, k2oz;      /* the two months (1..12), gas_update's time factors                  */
    float facoz;             /* weight of k2oz                                                     */
} wr_ozone;

typedef struct {
    int32_t llyr;            /* top of the boundary-layer cloud: 1-based top-first radiation layer,
                                cld_init's rule (radiation_clouds_nmmb.f:327-344, ivflip = 0)         */
    int32_t iovr;            /* gethml cloud overlap: 0 random, 1 maximum-random                      */
    int32_t subgrid;         /* 1: subgrid cloud water where RH > 95 % and no condensate (the NAM's)   */
    const wr_ozone *ozone;   /* climatological ozone for the column call (with o3_top), or NULL: the
                                state's O3 (the added top layer takes the top model layer's)          */
} wr_nmmb_opts;

typedef struct {             /* device arrays the prep writes; ncol = (j1 - j0) * im                  */
    double *p_lev;           /* [lm+1][ncol] interface pressure, Pa                                   */
    double *t_lay, *q_lay, *o3_lay;               /* [lm][ncol]                                       */
    double *cldf, *cwp, *cip, *crp, *rew, *rei;   /* [lm][ncol] progcld2 layer cloud fraction, liquid /
                                                     ice / rain paths (g/m2), liquid / ice radii (um) */
    float *clds;             /* [5][ncol] gethml cover: low, middle, high, total, boundary layer      */
    int32_t *mtop, *mbot;    /* [3][ncol] top / base layer of low, middle, high (1-based top-first
                                radiation layers, the added layer above the model top is layer 1)    */
    float *p_phy;            /* optional [lm][jm][im] state field P_PHY (radiated rows), or NULL      */
    double *o3_top;          /* [ncol] ozone of the layer added above the model top (written with the
                                climatological ozone; may be NULL otherwise)                           */
} wr_nmmb_cols;

typedef struct {             /* state outputs (float32 device arrays; int32 counters)                */
    float *rlwtt, *rswtt, *cldfra;                          /* [lm][jm][im]                           */
    float *rlwin, *rlwtoa, *sigt4;                           /* [jm][im], LW steps                    */
    float *rswin, *rswout, *rswinc, *rswtoa;                 /* [jm][im], SW steps                    */
    float *czmean;                                           /* [jm][im], every call                  */
    float *cfracl, *cfracm, *cfrach, *acfrst, *acfrcv;       /* [jm][im], every call                  */
    int32_t *ncfrst, *ncfrcv;                                /* [jm][im], every call                  */
} wr_nmmb_out;

int wr_nmmb_prep(const wr_nmmb_state *st, const wr_nmmb_opts *opt, const wr_nmmb_cols *cols, void *stream);

/* rad: the radiation call's outputs on the same columns (fsw_dn_sfc_clr required on SW steps);
 * czmean [ncol]: the call's mean cosine of the zenith angle; lw_step / sw_step: which parts ran. */
int wr_nmmb_post(const wr_nmmb_state *st, const wr_nmmb_cols *cols, const wr_outputs *rad, const double *czmean,
                 int32_t lw_step, int32_t sw_step, const wr_nmmb_out *out, void *stream);

/* NAM surface SW albedo of the radiated columns (setalb with ialbflg = 2: the NMMB radiation
 * driver's IALB = 2 block, phys_rrtm_grrad_clouds.cu:345-396, which setalb copies,
 * phys_rrtm_setalb.cu:241-247): the state's ALBEDO, except over open water (SM > 0.5) with the sun up
 * (coszen > 0), where the Payne (1972) open-water table at the fixed atmospheric transmission 0.537
 * gives it by zenith angle.  albedo, sm: [jm][im] float32 state; coszen [ncol] the call's cos zenith
 * of the radiated rows; alb [ncol] out, the column call's input.  Only st's grid fields are read. */
int wr_nmmb_albedo(const wr_nmmb_state *st, const float *albedo, const float *sm, const double *coszen, double *alb,
                   void *stream);

/* ---- climatological aerosol (the NAM's iaer = 11: the OPAC climatology) --------------------------
 * wr_aer_optics replaces the setaer call of the radiation (phys_rrtm_grrad_aerosols.py:91-124 ->
 * phys_rrtm_setaer.py:142-196 with its kernels phys_rrtm_setaer_prep, phys_rrtm_setaer.cu:86-161,
 * and phys_rrtm_aer_property_column, phys_rrtm_aer_property.cu:104-434), in the NAM's setting
 * (laswflg, lalwflg, no volcanic aerosol, layers from the top, ivflip = 0).  Per column:
 *   - the layer thickness and level heights (km) from the hypsometric equation on the column call's
 *     levels, the added top layer from the top interface to half its pressure, counted twice;
 *   - the layer relative humidity of grrad setup (phys_rrtm_grrad_setup.cu:156-161, the GFS
 *     saturation vapour pressure table fpvs), the added top layer taking the one below;
 *   - the 5-degree climatology bilinearly mapped on the column: species mixing ratios, the number
 *     densities of the mixing layer and the mineral layer, the nearest profile type with its
 *     mineral-overlay rules; the vertical domains (mixing layer, mineral transport, free
 *     troposphere, stratosphere, above) from the surface up by reference sigma, and each layer's
 *     effective thickness under the domain's exponential profile;
 *   - per band (14 SW, 16 LW) and layer: optical depth, single-scattering albedo and asymmetry from
 *     the band-mean species optics (humidity-dependent species interpolated in relative humidity),
 *     then the optical depth smoothed across domain boundaries (ratios above 30 or below 1/30).
 * cols: p_lev, t_lay, q_lay of the column call (ncol, nlay <= 63; the layer pressure is the interface
 * mean, as the column call forms it); xlon, xlat [ncol] radians; slmsk [ncol] the driver call's
 * surface type (0 sea, 1 land, 2 sea ice); aer_sw [3][14][nlay+1][aer_ld], aer_lw [3][16][nlay+1]
 * [aer_ld] out (optical depth, single-scattering albedo, asymmetry: the wr_columns.aer_sw / aer_lw
 * inputs).  FP64 arithmetic on the float32 tables; device pointers. */
typedef struct {
    const float *haer, *prsref, *sigref;             /* [7][5] profile type x domain: scale height (km),
                                                        reference pressure (mb), 0.001 prsref          */
    const float *rhlev;                              /* [8] relative-humidity levels                    */
    const float *extrhi, *scarhi, *ssarhi, *asyrhi;  /* [30][6] band x humidity-independent species     */
    const float *extrhd, *scarhd, *ssarhd, *asyrhd;  /* [30][4][8] band x humidity-dependent species x
                                                        humidity level                                  */
    const float *extstra;                            /* [30] stratospheric background extinction        */
    const float *cmixg;                              /* [37][72][5] latitude (90 N first) x longitude
                                                        (from 0 E) x component: mixing ratio            */
    const float *denng;                              /* [37][72][2] number density: mixing layer, mineral */
    const int32_t *idxcg;                            /* [37][72][5] species of each component (1..10, 0 none) */
    const int32_t *kprfg;                            /* [37][72] profile type (1..7)                    */
} wr_aer_tables;
/* Table rows of the band dimension: 0..13 the SW bands in the column call's order, 14..29 the LW bands. */

int wr_aer_optics(const wr_aer_tables *tab, const wr_columns *cols, const float *xlon, const float *xlat,
                  const float *slmsk, float *aer_sw, float *aer_lw, int64_t aer_ld, void *stream);

/* ---- solar geometry around the call ---------------------------------------------------------------
 * wr_sol_update: sol_update_nmmb (phys_rrtm_sol_update.cu:16-160): the sun's position for a date
 * (UTC) from its orbit (Kepler's equation by Newton iteration) and the solar constant solc0 / r^2 at
 * the earth-sun distance r; nstp and anginc are the time steps sol_update sets for a SW interval of
 * deltsw seconds with model step deltim.  Host computation, no device work.
 * wr_coszmn: coszmn_nmmb (phys_rrtm_coszmn.cu:18-62): per column (device arrays [ncol], radians) the
 * cosine of the zenith angle at nrads instants over the SW interval (dtswav hours from the hour solhr
 * UTC), averaged over the sunlit instants (coszen, the radiation's cos z and the driver's CZMEAN) and
 * over all of them (coszdg, optional).
 * wr_rdtemp: RDTEMP (drv_rdtemp.cu:10-145), every model step: the sun's declination and hour angle at
 * step ntsd of length dt from the start date (julday of julyr, hour ihrst; leap: February has 29
 * days), czen [jm][im] of every column, and T += (RSWTT czen / CZMEAN + RLWTT) dt on the columns
 * i_lo..i_hi, j_lo..j_hi (0-based inclusive) of the [lm][jm][im] float32 state. */
typedef struct {
    double sollag;           /* equation of time (solar lag), radians */
    double sindec, cosdec;   /* sine and cosine of the declination     */
    double anginc;           /* solar angle increment per sub-step      */
    double solcon;           /* solar constant at the date, W/m2       */
    int32_t nstp;            /* sub-steps per SW interval               */
} wr_sol;

int wr_sol_update(int32_t yr, int32_t mon, int32_t day, int32_t hr, int32_t mn, int32_t sec, double deltsw,
                  double deltim, double solc0, wr_sol *out);
int wr_coszmn(int32_t ncol, const double *glat, const double *glon, double solhr, double dtswav, int32_t nrads,
              const wr_sol *sol, double *coszen, double *coszdg, void *stream);
int wr_rdtemp(int32_t im, int32_t jm, int32_t lm, int32_t i_lo, int32_t i_hi, int32_t j_lo, int32_t j_hi, int32_t ntsd,
              double dt, int32_t julday, int32_t julyr, int32_t ihrst, int32_t leap, const float *glat,
              const float *glon, const float *czmean, const float *rswtt, const float *rlwtt, float *t, float *czen,
              void *stream);

/* ---- gas-optics surrogate (no counterpart in the reference: a learned replacement of taugb/taumol)
 *
 * wr_layer_optics: per radiation layer (surface-first k, column i, sample s = k*ncol + i) of the
 * plain FP64 path: feat[s][8] = ln p (hPa), T (K), ln q, ln o3, dp (hPa), 0, 0, 0 and, when targ is
 * not NULL, targ[s][504] = ln tau_LW (140 g), Planck fraction (140), ln tau_SW gas+Rayleigh (112),
 * SW single-scattering albedo (112).  Device pointers, float32.
 *
 * wr_mlp_gemm: C[M x N] = A[M x K] * B[K x N] on BF16 MFMA with FP32 accumulation; element (m,k)
 * of A at A[m*lda_m + k*lda_k] (float32, or bfloat16 when a_bf16), B likewise, C (float32, may be
 * NULL) and/or Cb (bfloat16, may be NULL) at m*ldc_m + n*ldc_n.  epilogue: 0 store, 1 + bias[n],
 * 2 + bias then SiLU (pre-activation to Zout when not NULL), 3 multiply by SiLU'(Zin).  splits > 1 (epilogue 0,
 * C only): split-K partial sums at C + s*split_stride, summed by wr_mlp_splitk_reduce.
 * wr_mlp_colsum: out[n] = sum_m X[m*ld + n].  wr_mlp_mse: dY = 2 (Y - T) scale, per-block sums of
 * (Y - T)^2 in partial[nblocks].  wr_mlp_adam: Adam step `step` (1-based) on FP32 weights, with an
 * optional BF16 copy wb.  wr_mlp_to_bf16: element-wise conversion. */
/* Surrogate weights and normalisation (device pointers, owned by the caller, kept by the context):
 * y = W3 SiLU(W2 SiLU(W1 x + b1) + b2) + b3, x = (feature - fmean) / fstd (5 features, padded to
 * 8 with zeros), raw output = y * tstd + tmean. */
typedef struct {
    const void *w1, *w2, *w3;                   /* bfloat16 [128][8], [128][128], [504][128]   */
    const float *b1, *b2, *b3;                  /* [128], [128], [504]                          */
    const float *fmean, *fstd, *fmin, *fmax;    /* [5]: normalisation and training envelope     */
    const float *tmean, *tstd;                  /* [504]                                        */
    uint32_t bands;                             /* bands that use the network: bit b for LW band
                                                   b + 1 (b < 16), SW band b (16 <= b < 30);
                                                   0: every band.  The others keep the tables.  */
} wr_surrogate;

/* Attach a surrogate to a context and allocate its tile work space (about 1.04 KB per layer-column
 * of max_cols x 64 layers).  wr_run_surrogate: the mixed-precision call with the LW and SW gas
 * optics (taugb/taumol) replaced by the network for every column whose layers all lie inside the
 * training envelope; other columns use the tables.  col_flag (optional, device, [ncol]) receives 1
 * for columns that fell back to the tables. */
int wr_surrogate_attach(wr_ctx *ctx, const wr_surrogate *s);
int wr_run_surrogate(wr_ctx *ctx, const wr_columns *in, const wr_params *par, const wr_outputs *out,
                     int32_t *col_flag, void *stream);

int wr_layer_optics(wr_ctx *ctx, const wr_columns *in, const wr_params *par, float *feat, float *targ,
                    void *stream);
int wr_mlp_gemm(int32_t M, int32_t N, int32_t K, const void *A, int32_t a_bf16, int64_t lda_m,
                int64_t lda_k, const void *B, int32_t b_bf16, int64_t ldb_k, int64_t ldb_n, float *C,
                void *Cb, int64_t ldc_m, int64_t ldc_n, const float *bias, float *Zout,
                const float *Zin, int32_t epilogue, int32_t splits, int64_t split_stride, void *stream);
int wr_mlp_splitk_reduce(const float *part, int32_t splits, int64_t stride, float *out, int64_t n,
                         void *stream);
int wr_mlp_colsum(const float *X, int32_t M, int32_t N, int64_t ld, float *out, void *stream);
int wr_mlp_mse(const float *Y, const float *T, int64_t n, float scale, float *dY, float *partial,
               int32_t nblocks, void *stream);
int wr_mlp_adam(float *w, void *wb, const float *grad, float *m1, float *m2, int64_t n, float lr,
                float b1, float b2, float eps, int32_t step, void *stream);
int wr_mlp_to_bf16(const float *a, void *b, int64_t n, void *stream);
/* The network's inference in one kernel (what wr_run_surrogate runs per tile): X [M][8] normalised
 * features (float32, 16-byte aligned rows) -> Y bfloat16, output-major, output j of row m at
 * Y[j*ldy + m] (ldy >= M), the hidden layers kept on chip.  Equal bit for bit to the three
 * wr_mlp_gemm layers (bias + SiLU, bias + SiLU, bias) with a BF16 output.  Weights 16-byte aligned. */
int wr_mlp_forward(const wr_surrogate *w, int64_t M, const float *X, void *Y, int64_t ldy, void *stream);

/* ---- served tiles (no counterpart in the reference: the serving loop of this build) -------------
 * A slot owns, for one context of max_cols (= tile) columns: a stream of its own (default
 * priority), staging buffers for one tile's inputs and outputs (the qc / qi form of wr_columns,
 * FP64), the tile's column ids, McICA seed and valid-count words, a completion event and, unless
 * graphs is 0, a HIP graph of a whole-tile call captured at creation.  precision: 0 FP64, 1 mixed,
 * 2 surrogate (the context must have one attached).  filler: columns (at least 1, nlay layers)
 * repeated over the staging tile at creation, so a replay never computes on uninitialised memory.
 * wr_slot_launch: columns [c0, c0 + n) of a request (req: device arrays [k][req->ncol], col_id
 * optional) through the slot: one gather kernel (inputs, ids, seed, count), the graph replay (or the
 * call), one scatter kernel into res ([k][req->ncol]; fsw_dn_sfc_clr optional), the event --
 * asynchronous, nothing allocated.  wr_slot_query: 1 when the last launch has finished, 0 while it
 * runs, negative on an error; wr_slot_sync waits for it.  One slot runs one tile at a time; the
 * context must outlive the slot. */
typedef struct wr_slot wr_slot;
int wr_slot_create(wr_slot **slot, wr_ctx *ctx, int32_t precision, const wr_params *par, int32_t graphs,
                   const wr_columns *filler);
int wr_slot_launch(wr_slot *slot, const wr_columns *req, const wr_outputs *res, int32_t c0, int32_t n,
                   uint64_t seed);
int wr_slot_query(wr_slot *slot);
int wr_slot_sync(wr_slot *slot);
void wr_slot_destroy(wr_slot *slot);

/* Per-launch kernel timing for measurement (no counterpart in the reference: the reference
 * publishes call-level times only).  wr_timing(ctx, 1) creates an event pool once and brackets
 * every later kernel launch of the context with HIP events on the launching stream (up to 8192
 * launches; reset on every wr_timing call).  wr_timing_read waits for the recorded launches and
 * returns kernel id and duration (ms) of each, then clears the record; wr_kernel_name names an id. */
int wr_timing(wr_ctx *ctx, int32_t enable);
int wr_timing_read(wr_ctx *ctx, int32_t *kernel_id, float *ms, int32_t cap, int32_t *n);
const char *wr_kernel_name(int32_t kernel_id);

const char *wr_strerror(int code);
void wr_destroy(wr_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* WOOFRAD_H */
