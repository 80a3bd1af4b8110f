// wr_host.hip -- C ABI of libwoofrad: table upload, context, tile loop, diagnostics and the small
// stand-alone kernels (CAL_CLDFRA, McICA masks).
//
// The table file (rrtmg_tables.bin, "WRTBL001") is read once per context and laid out in two
// device blobs, one per optics precision (FP64 and FP32).  Index-arithmetic tables (reference
// pressures/temperatures, chi_mls) stay FP64 in both.  The exp/transmittance look-up tables are
// generated on the device by the same IEEE operation sequence on every platform.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "wr_glw.h"
#include "wr_internal.h"

using namespace wr;

namespace {

// ---- table file -------------------------------------------------------------------------------
struct TEntry {
    uint32_t dtype;  // 0 = float32, 1 = int32
    uint64_t count;
    const uint8_t *data;
};
struct TFile {
    std::vector<uint8_t> buf;
    std::unordered_map<std::string, TEntry> e;
};

int read_tables(const char *path, TFile &f)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return WR_ERR_TABLES;
    fseek(fp, 0, SEEK_END);
    const long sz = ftell(fp);
    fseek(fp, 0, SEEK_SET);
    if (sz < 12) {
        fclose(fp);
        return WR_ERR_TABLES;
    }
    f.buf.resize((size_t)sz);
    const size_t got = fread(f.buf.data(), 1, (size_t)sz, fp);
    fclose(fp);
    if (got != (size_t)sz || memcmp(f.buf.data(), "WRTBL001", 8) != 0) return WR_ERR_TABLES;
    uint32_t n;
    memcpy(&n, f.buf.data() + 8, 4);
    const size_t rec = 48 + 4 + 4 + 16 + 8 + 8;
    if (12 + (size_t)n * rec > (size_t)sz) return WR_ERR_TABLES;
    for (uint32_t k = 0; k < n; ++k) {
        const uint8_t *p = f.buf.data() + 12 + (size_t)k * rec;
        char name[49];
        memcpy(name, p, 48);
        name[48] = 0;
        TEntry t;
        uint64_t off;
        memcpy(&t.dtype, p + 48, 4);
        memcpy(&off, p + 72, 8);
        memcpy(&t.count, p + 80, 8);
        if (t.dtype > 1 || off + t.count * 4 > (uint64_t)sz) return WR_ERR_TABLES;
        t.data = f.buf.data() + off;
        f.e[std::string(name)] = t;
    }
    return WR_OK;
}

// Device blob builder.  bind() runs twice: once with base 0 to size the blob, once with the real
// device base; both passes append the same bytes in the same order, so offsets agree.
struct Blob {
    std::vector<uint8_t> bytes;
    uintptr_t base = 0;
    bool bad = false;
    size_t align()
    {
        size_t off = (bytes.size() + 255) & ~(size_t)255;
        bytes.resize(off);
        return off;
    }
    template <typename T> const WR_G T *conv(const TEntry &e)
    {
        const size_t off = align();
        bytes.resize(off + e.count * sizeof(T));
        T *dst = (T *)(bytes.data() + off);
        if (e.dtype == 0) {
            const float *s = (const float *)e.data;
            for (uint64_t j = 0; j < e.count; ++j) dst[j] = (T)s[j];
        } else {
            const int32_t *s = (const int32_t *)e.data;
            for (uint64_t j = 0; j < e.count; ++j) dst[j] = (T)s[j];
        }
        return (const WR_G T *)(base + off);
    }
    // the tables ent[0 .. nt) (null entries skipped) one after another from a 256-byte boundary, each
    // padded with zeros to a multiple of 4 elements
    template <typename T> const WR_G T *pack(const TEntry *const *ent, int nt)
    {
        const size_t off = align();
        for (int j = 0; j < nt; ++j) {
            if (!ent[j]) continue;
            const size_t at = bytes.size(), cnt = (ent[j]->count + 3) / 4 * 4;
            bytes.resize(at + cnt * sizeof(T), 0);
            T *dst = (T *)(bytes.data() + at);
            if (ent[j]->dtype != 0) {
                bad = true;
                continue;
            }
            const float *src = (const float *)ent[j]->data;
            for (uint64_t q = 0; q < ent[j]->count; ++q) dst[q] = (T)src[q];
        }
        return (const WR_G T *)(base + off);
    }
    template <typename T> WR_G T *reserve(size_t n)
    {
        const size_t off = align();
        bytes.resize(off + n * sizeof(T), 0);
        return (WR_G T *)(base + off);
    }
};

struct Binder {
    const TFile &f;
    Blob &b;
    const TEntry *find(const std::string &nm) const
    {
        auto it = f.e.find(nm);
        return it == f.e.end() ? nullptr : &it->second;
    }
    template <typename T> const WR_G T *need(const std::string &nm)
    {
        const TEntry *e = find(nm);
        if (!e) {
            b.bad = true;
            return nullptr;
        }
        return b.conv<T>(*e);
    }
    template <typename T> const WR_G T *opt(const std::string &nm)
    {
        const TEntry *e = find(nm);
        return e ? b.conv<T>(*e) : nullptr;
    }
    double scalar(const std::string &nm, bool *ok)
    {
        const TEntry *e = find(nm);
        if (!e || e->count < 1 || e->dtype != 0) {
            *ok = false;
            return 0.0;
        }
        return (double)((const float *)e->data)[0];
    }
};

std::string bname(const char *pre, int b, const char *key, bool two)
{
    char nm[64];
    if (two)
        snprintf(nm, sizeof nm, "%s/b%02d/%s", pre, b, key);
    else
        snprintf(nm, sizeof nm, "%s/b%d/%s", pre, b, key);
    return nm;
}

// Host image of the table set in precision R; device pointers are base + offset.
template <typename R> int bind_tables(const TFile &f, Blob &b, Tabs<R> &T)
{
    memset(&T, 0, sizeof(T));
    Binder d{f, b};
    T.preflog = d.need<double>("lw/preflog");
    T.tref = d.need<double>("lw/tref");
    T.chi_mls = d.need<double>("lw/chi_mls");
    T.sw_preflog = d.need<double>("sw/preflog");
    T.sw_tref = d.need<double>("sw/tref");
    T.totplnk = d.need<R>("lw/totplnk");
    T.delwave = d.need<R>("lw/delwave");
    T.absliq1 = d.need<R>("lw/absliq1");
    T.absice1 = d.need<R>("lw/absice1");
    T.ipat = d.need<int32_t>("lw/ipat");
    T.layreffr = d.need<int32_t>("sw/layreffr");
    T.ix1 = d.need<int32_t>("sw/ix1");
    T.ix2 = d.need<int32_t>("sw/ix2");
    T.ibx = d.need<int32_t>("sw/ibx");
    T.sf01 = d.need<R>("sw/sfluxref01");
    T.sf02 = d.need<R>("sw/sfluxref02");
    T.sf03 = d.need<R>("sw/sfluxref03");
    T.sw_strrat = d.need<R>("sw/strrat");
    T.sw_specwt = d.need<R>("sw/specwt");
    T.extliq1 = d.need<R>("sw/extliq1");
    T.ssaliq1 = d.need<R>("sw/ssaliq1");
    T.asyliq1 = d.need<R>("sw/asyliq1");
    T.extice3 = d.need<R>("sw/extice3");
    T.ssaice3 = d.need<R>("sw/ssaice3");
    T.asyice3 = d.need<R>("sw/asyice3");
    bool ok = true;
    T.scalekur = (R)d.scalar("sw/scalekur", &ok);
    if (!ok) return WR_ERR_TABLES;
    T.exp_tbl = b.reserve<R>(NTBL + 1);
    T.tfn_tbl = b.reserve<R>(NTBL + 1);
    T.tau_tbl = b.reserve<R>(NTBL + 1);
    T.sw_exp_tbl = b.reserve<R>(NTBL + 1);

    for (int k = 0; k < 16; ++k) {
        LwBand<R> &B = T.lw[k];
        const int bn = k + 1;
        B.absa = d.need<R>(bname("lw", bn, "absa", true));
        B.absb = (k != 5 && k != 11 && k != 12 && k != 14) ? d.need<R>(bname("lw", bn, "absb", true))
                                                             : d.opt<R>(bname("lw", bn, "absb", true));
        B.selfref = d.need<R>(bname("lw", bn, "selfref", true));
        B.forref = d.need<R>(bname("lw", bn, "forref", true));
        B.fracrefa = d.need<R>(bname("lw", bn, "fracrefa", true));
        B.fracrefb = (k != 5 && k != 11 && k != 14) ? d.need<R>(bname("lw", bn, "fracrefb", true))
                                                    : d.opt<R>(bname("lw", bn, "fracrefb", true));
    }
    // minor-gas and cross-section slots in the order the band code reads them
    struct Slot {
        int band;
        int kind;  // 0 ka_m, 1 kb_m, 2 x1, 3 x2
        int idx;
        const char *key;
    };
    static const Slot slots[] = {
        {1, 0, 0, "ka_mn2"},  {1, 1, 0, "kb_mn2"},  {3, 0, 0, "ka_mn2o"}, {3, 1, 0, "kb_mn2o"}, {5, 0, 0, "ka_mo3"},
        {5, 2, 0, "ccl4"},    {6, 0, 0, "ka_mco2"}, {6, 2, 0, "cfc11adj"}, {6, 3, 0, "cfc12"},  {7, 0, 0, "ka_mco2"},
        {7, 1, 0, "kb_mco2"}, {8, 0, 0, "ka_mco2"}, {8, 1, 0, "kb_mco2"}, {8, 0, 1, "ka_mo3"},  {8, 0, 2, "ka_mn2o"},
        {8, 1, 2, "kb_mn2o"}, {8, 2, 0, "cfc12"},   {8, 3, 0, "cfc22adj"}, {9, 0, 0, "ka_mn2o"}, {9, 1, 0, "kb_mn2o"},
        {11, 0, 0, "ka_mo2"}, {11, 1, 0, "kb_mo2"}, {13, 0, 0, "ka_mco2"}, {13, 0, 1, "ka_mco"}, {13, 1, 0, "kb_mo3"},
        {15, 0, 0, "ka_mn2"},
    };
    for (const Slot &s : slots) {
        LwBand<R> &B = T.lw[s.band - 1];
        const WR_G R *p = d.need<R>(bname("lw", s.band, s.key, true));
        if (s.kind == 0) B.ka_m[s.idx] = p;
        if (s.kind == 1) B.kb_m[s.idx] = p;
        if (s.kind == 2) B.x1 = p;
        if (s.kind == 3) B.x2 = p;
    }
    // every LW band's tables packed into one image (k_lwl and k_lwl64 copy it to LDS), in the LwImg
    // order; the FP64 image leaves absb out
    for (int k = 0; k < 16; ++k) {
        static const char *const keys[6] = {"absa", "absb", "selfref", "forref", "fracrefa", "fracrefb"};
        const TEntry *ent[LW_IMG_T] = {};
        for (int j = 0; j < 6; ++j)
            if (j != 1 || sizeof(R) == 4) ent[j] = d.find(bname("lw", k + 1, keys[j], true));
        for (const Slot &s : slots)
            if (s.band == k + 1)
                ent[s.kind == 0 ? 6 + s.idx : s.kind == 1 ? 9 + s.idx : s.kind == 2 ? 12 : 13] =
                    d.find(bname("lw", s.band, s.key, true));
        LwImg<R> &im = T.lw_img[k];
        int n = 0;
        for (int j = 0; j < LW_IMG_T; ++j) {
            im.off[j] = ent[j] ? n : -1;
            if (ent[j]) n += (int)((ent[j]->count + 3) / 4 * 4);
        }
        im.n = n;
        im.base = b.pack<R>(ent, LW_IMG_T);
    }
    // reference ratios of the binary-species bands from chi_mls (taugbNN.cu constants)
    const TEntry *chi = d.find("lw/chi_mls");
    if (!chi || chi->count < 7 * 59) return WR_ERR_TABLES;
    const float *cm = (const float *)chi->data;
    auto rr = [&](int g1, int l1, int g2, int l2) {
        return (R)((double)cm[(l1 - 1) * 7 + (g1 - 1)] / (double)cm[(l2 - 1) * 7 + (g2 - 1)]);
    };
    T.lw[2].refrat_planck_a = rr(1, 9, 2, 9);
    T.lw[2].refrat_planck_b = rr(1, 13, 2, 13);
    T.lw[2].refrat_m_a = rr(1, 3, 2, 3);
    T.lw[2].refrat_m_b = rr(1, 13, 2, 13);
    T.lw[3].refrat_planck_a = rr(1, 11, 2, 11);
    T.lw[3].refrat_planck_b = rr(3, 13, 2, 13);
    T.lw[4].refrat_planck_a = rr(1, 5, 2, 5);
    T.lw[4].refrat_planck_b = rr(3, 43, 2, 43);
    T.lw[4].refrat_m_a = rr(1, 7, 2, 7);
    T.lw[6].refrat_planck_a = rr(1, 3, 3, 3);
    T.lw[6].refrat_m_a = rr(1, 3, 3, 3);
    T.lw[8].refrat_planck_a = rr(1, 9, 6, 9);
    T.lw[8].refrat_m_a = rr(1, 3, 6, 3);
    T.lw[11].refrat_planck_a = rr(1, 10, 2, 10);
    T.lw[12].refrat_planck_a = rr(1, 5, 4, 5);
    T.lw[12].refrat_m_a = rr(1, 1, 4, 1);
    T.lw[12].refrat_m_a3 = rr(1, 3, 4, 3);
    T.lw[14].refrat_planck_a = rr(4, 1, 2, 1);
    T.lw[14].refrat_m_a = rr(4, 1, 2, 1);
    T.lw[15].refrat_planck_a = rr(1, 6, 6, 6);

    const TEntry *srat = d.find("sw/strrat");
    if (!srat || srat->count < 14) return WR_ERR_TABLES;
    for (int k = 0; k < 14; ++k) {
        SwBand<R> &B = T.sw[k];
        const int jb = k + 16;
        B.absa = d.opt<R>(bname("sw", jb, "absa", false));
        B.absb = d.opt<R>(bname("sw", jb, "absb", false));
        B.selfref = d.opt<R>(bname("sw", jb, "selfref", false));
        B.forref = d.opt<R>(bname("sw", jb, "forref", false));
        B.rayla = d.opt<R>(bname("sw", jb, "rayla", false));
        B.raylb = d.opt<R>(bname("sw", jb, "raylb", false));
        B.abso3a = d.opt<R>(bname("sw", jb, "abso3a", false));
        B.abso3b = d.opt<R>(bname("sw", jb, "abso3b", false));
        B.absch4 = d.opt<R>(bname("sw", jb, "absch4", false));
        B.absh2o = d.opt<R>(bname("sw", jb, "absh2o", false));
        B.absco2 = d.opt<R>(bname("sw", jb, "absco2", false));
        const TEntry *ry = d.find(bname("sw", jb, "rayl", false));
        if (ry && ry->count == 1)
            B.rayl = (R)((const float *)ry->data)[0];
        else if (ry)
            B.rayl_v = b.conv<R>(*ry);
        B.strrat = (R)((const float *)srat->data)[k];
    }
    // strrat(16)/(28) come with their band files (kgb16/kgb28).  Every strrat is the single-precision
    // value of the reference (radsw_datatb_nmmb.f:1108-1110 data; phys_rrtm_sw_taumol18.py:42
    // np.float32(3.89589e+1), phys_rrtm_sw_taumol19.cu:80 5.49281e+0f, ...), as the table holds it.
    // every SW band's tables packed into one image (k_swl copies it to LDS), in the SwImg order; the
    // FP64 image leaves absb out, a scalar rayl is not a table
    for (int k = 0; k < 14; ++k) {
        static const char *const keys[SW_IMG_T] = {"absa",   "absb",   "selfref", "forref", "rayl",   "rayla",
                                                   "raylb",  "abso3a", "abso3b",  "absch4", "absh2o", "absco2"};
        const TEntry *ent[SW_IMG_T] = {};
        for (int j = 0; j < SW_IMG_T; ++j) {
            if (j == 1 && sizeof(R) != 4) continue;
            const TEntry *e = d.find(bname("sw", k + 16, keys[j], false));
            if (e && (j != 4 || e->count > 1)) ent[j] = e;
        }
        SwImg<R> &im = T.sw_img[k];
        int n = 0;
        for (int j = 0; j < SW_IMG_T; ++j) {
            im.off[j] = ent[j] ? n : -1;
            if (ent[j]) n += (int)((ent[j]->count + 3) / 4 * 4);
        }
        im.n = n;
        im.base = b.pack<R>(ent, SW_IMG_T);
    }
    bool ok16 = true, ok28 = true;
    const double s16 = d.scalar("sw/b16/strrat16", &ok16), s28 = d.scalar("sw/b28/strrat28", &ok28);
    if (ok16) T.sw[0].strrat = (R)s16;
    if (ok28) T.sw[12].strrat = (R)s28;
    return b.bad ? WR_ERR_TABLES : WR_OK;
}

template <typename R> int upload(const TFile &f, void **buf, Tabs<R> &host, Tabs<R> **dev)
{
    Blob sizing;
    int rc = bind_tables<R>(f, sizing, host);
    if (rc != WR_OK) return rc;
    if (hipMalloc(buf, sizing.bytes.size()) != hipSuccess) return WR_ERR_NOMEM;
    Blob place;
    place.base = (uintptr_t)*buf;
    rc = bind_tables<R>(f, place, host);
    if (rc != WR_OK) return rc;
    if (place.bytes.size() != sizing.bytes.size()) return WR_ERR_TABLES;
    if (hipMemcpy(*buf, place.bytes.data(), place.bytes.size(), hipMemcpyHostToDevice) != hipSuccess) return WR_ERR_HIP;
    if (hipMalloc((void **)dev, sizeof(Tabs<R>)) != hipSuccess) return WR_ERR_NOMEM;
    if (hipMemcpy(*dev, &host, sizeof(Tabs<R>), hipMemcpyHostToDevice) != hipSuccess) return WR_ERR_HIP;
    return WR_OK;
}

__global__ void k_d2f(const double *__restrict__ a, float *__restrict__ b, int n)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) b[j] = (float)a[j];
}

// ---- CAL_CLDFRA (phys_rad_cal_cldfra.cu:29-55) ----------------------------------------------------
__global__ void k_cal_cldfra(int im, int jm, int lm, int i_lo, int i_hi, int j_lo, int j_hi, int f_qc, int f_qi,
                             float *__restrict__ cldfra, const float *__restrict__ qc, const float *__restrict__ qi)
{
    const size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t tot = (size_t)im * jm * lm;
    if (p >= tot) return;
    const int i = (int)(p % im);
    const int j = (int)((p / im) % jm);
    if (i < i_lo || i > i_hi || j < j_lo || j > j_hi) return;
    float v;
    if (f_qi && f_qc)
        v = (qc[p] + qi[p] > 1.0e-6f) ? 1.0f : 0.0f;
    else if (f_qc)
        v = (qc[p] > 1.0e-6f) ? 1.0f : 0.0f;
    else
        v = 0.0f;
    cldfra[p] = v;
}

// ---- McICA masks from a given surface-first cloud fraction ---------------------------------------
__global__ void __launch_bounds__(64) k_mcica_masks(int ncol, int nrl, const double *__restrict__ cldf,
                                                    const uint64_t *__restrict__ col_id, uint64_t seed, int iovr,
                                                    uint32_t *__restrict__ lw_mask, uint32_t *__restrict__ sw_mask)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncol) return;
    const size_t nc = (size_t)ncol;
    const uint64_t col = col_id ? col_id[i] : (uint64_t)i;
    auto cf = [&](int k) { return cldf[k * nc + i]; };
    for (int kind = 0; kind < 2; ++kind) {
        uint32_t *m = kind == 0 ? lw_mask : sw_mask;
        if (!m) continue;
        const int ng = kind == 0 ? 140 : 112, words = kind == 0 ? 5 : 4;
        for (int k = 0; k < nrl; ++k)
            for (int w = 0; w < words; ++w) m[((size_t)k * words + w) * nc + i] = 0u;
        for (int g = 0; g < ng; ++g) {
            const uint64_t mb = mcica_g(col, seed, kind, g, nrl, iovr, cf);
            for (int k = 0; k < nrl; ++k)
                if ((mb >> k) & 1) m[((size_t)k * words + (g >> 5)) * nc + i] |= 1u << (g & 31);
        }
    }
}

struct DevGuard {  // make the context's device current for the call, restore afterwards
    int prev = -1;
    explicit DevGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DevGuard()
    {
        int cur;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

int check_args(const wr_ctx *c, const wr_columns *in, const wr_params *p, const wr_outputs *o)
{
    if (!c || !in || !p || !o) return WR_ERR_ARG;
    if (in->ncol < 0 || in->nlay < 2 || in->nlay + 1 > c->nl_max) return WR_ERR_ARG;
    if (in->ncol == 0) return WR_OK;
    if (!in->p_lev || !in->t_lay || !in->q_lay || !in->o3_lay || !in->cldf || !in->t_sfc || !in->emis || !in->alb ||
        !in->cosz)
        return WR_ERR_ARG;
    if (in->cwp ? (!in->cip || !in->rew || !in->rei) : (!in->qc || !in->qi)) return WR_ERR_ARG;
    if ((in->aer_sw || in->aer_lw) && in->aer_ld != 0 && in->aer_ld < in->ncol) return WR_ERR_ARG;
    if (!o->hr_lw || !o->hr_sw || !o->flw_up || !o->flw_dn || !o->fsw_up || !o->fsw_dn) return WR_ERR_ARG;
    if (p->iovr < 0 || p->iovr > 2) return WR_ERR_ARG;
    if (!(p->re_liq > 0.0) || !(p->re_ice > 0.0) || !(p->solcon > 0.0)) return WR_ERR_ARG;
    return WR_OK;
}

RunPar to_par(const wr_params *p)
{
    RunPar r;
    r.mcica = p->mcica, r.iovr = p->iovr, r.seed = p->seed;
    r.solcon = p->solcon, r.re_liq = p->re_liq, r.re_ice = p->re_ice;
    for (int g = 0; g < 10; ++g) r.vmr[g] = p->vmr[g];
    return r;
}

enum { MODE_FP64 = 0, MODE_MIXED = 1, MODE_SURR = 2 };

int run(wr_ctx *c, const wr_columns *in, const wr_params *p, const wr_outputs *o, const wr_diag *dg, void *stream,
        int mode, int band = -1, int32_t *flag_out = nullptr)
{
    int rc = check_args(c, in, p, o);
    if (rc != WR_OK || in->ncol == 0) return rc;
    if (p->mcica != 0 && p->mcica != 1) return WR_ERR_ARG;
    if (mode == MODE_SURR && (!c->has_sur || p->mcica != 1)) return p->mcica != 1 ? WR_ERR_UNSUPPORTED : WR_ERR_ARG;
    DevGuard g(c->device);
    hipStream_t s = (hipStream_t)stream;
    const ColIn ci = to_colin(in);
    const RunPar par = to_par(p);
    OutPtrs op;
    op.hr_lw = gptr(o->hr_lw), op.hr_sw = gptr(o->hr_sw), op.flw_up = gptr(o->flw_up), op.flw_dn = gptr(o->flw_dn);
    op.fsw_up = gptr(o->fsw_up), op.fsw_dn = gptr(o->fsw_dn), op.fsw_dn_sfc_clr = gptr(o->fsw_dn_sfc_clr);
    // full tiles of max_cols columns and one remainder.  Larger tiles are faster (fewer partly
    // filled rounds of band-kernel waves, longer overlap of the band streams): on 796,590 columns a
    // single tile takes 6 % (FP64) and 7 % (mixed) less time than tiles of 262,144.  Tiles that are
    // multiples of 65,536 columns (one wave per SIMD on 1,024 SIMDs) leave a partly filled round
    // only in the remainder tile
    const int per = c->max_cols;
    for (int col0 = 0; col0 < in->ncol; col0 += per) {
        const int n = (in->ncol - col0 < per) ? in->ncol - col0 : per;
        if (mode == MODE_SURR)
            rc = wr_launch_surr(c, ci, par, op, col0, n, flag_out ? flag_out + col0 : nullptr, s);
        else if (mode == MODE_MIXED)
            rc = wr_launch_mixed(c, ci, par, op, col0, n, band, s);
        else
            rc = wr_launch_fp64(c, ci, par, op, col0, n, band, s);
        if (rc != WR_OK) return rc;
        if (dg) {
            // indices and bits from the layer setup of the precision that ran (the FP32 path's own
            // instantiation for mixed calls), with the layer logs k_colprep left for this tile
            rc = mode == MODE_MIXED ? wr_launch_diag_mixed(c, ci, par, dg, col0, n, s)
                                    : wr_launch_diag_fp64(c, ci, par, dg, col0, n, s);
            if (rc != WR_OK) return rc;
        }
    }
    return WR_OK;
}

// WR_LW_GLANE=1 runs the McICA route's LW bands as the g-point-lane kernels over per-tile layer
// records (wr_glw.h) instead of the lane-per-column kernels; the records then take about 15 KB per
// column of context memory.  Off by default: measured on the 954x835x60 grid, the g-lane kernels
// took twice the time of the lane-per-column ones.
bool lw_glane_env()
{
    const char *e = getenv("WR_LW_GLANE");
    return e && atoi(e) != 0;
}

// FP32 McICA LW bands (bit B = band B + 1) that run k_lwl, the kernel with the band's tables in LDS:
// WR_LW_LDS (a hexadecimal mask) overrides the default.  A band whose table image exceeds the
// device's LDS per workgroup runs k_lwc.
// WR_LW_LDS64 does the same for the FP64 McICA LW kernels (k_lwl64 in place of k_lw).  The defaults
// are the measured choice per band on the 954x835x60 grid: every FP32 band (LW sum 101 -> 97 ms per
// call), every FP64 band but 14 (LW sum 233 -> 216 ms; band 14 ran 7 % slower).
constexpr uint32_t LW_LDS_DEFAULT = 0xffffu, LW_LDS64_DEFAULT = 0xffffu & ~(1u << 13);
// WR_SW_LDS / WR_SW_LDS64 (bit B = band B + 16) choose k_swl over k_sw the same way.  Measured per
// band (serialised, 954x835x60): FP64 SW bands 17-24 and 29 run 3-9 % faster from LDS, 25, 26, 28
// slower (the FP64 call 320 -> 312 ms with this mask); in FP32 the SW sum went 62.0 -> 65.4 ms, so
// FP32 SW keeps its global-table kernels.
constexpr uint32_t SW_LDS_DEFAULT = 0x0u, SW_LDS64_DEFAULT = 0x21feu;
// WR_LWR_LDS / WR_LWR_LDS64: the NAM route's LW kernels (k_lwrl over k_lwr), on the LW images.
constexpr uint32_t LWR_LDS_DEFAULT = 0xffffu, LWR_LDS64_DEFAULT = 0xffffu;
// WR_SWR_LDS / WR_SWR_LDS64: the NAM route's SW kernels (k_swrl over k_swr), on the SW images.
constexpr uint32_t SWR_LDS_DEFAULT = 0x0u, SWR_LDS64_DEFAULT = 0x3fffu;
template <typename R, typename I>
uint32_t lds_mask(const char *env, uint32_t dflt, const I *img, int nb, int lds_max)
{
    const char *e = getenv(env);
    uint32_t m = e ? (uint32_t)strtoul(e, nullptr, 16) : dflt;
    for (int b = 0; b < nb; ++b)
        if (img[b].n <= 0 || (size_t)img[b].n * sizeof(R) > (size_t)lds_max) m &= ~(1u << b);
    return m & ((1u << nb) - 1u);
}

// Device bytes of a context's tile work space for tiles of m columns, and the band-stream count.
struct Sizes {
    size_t acc, iws, dws, rtr, sflux_d, sflux_f, lwrec, pool;
    int naux;
    size_t total() const
    {
        return acc + iws + dws + rtr + sflux_d + sflux_f + lwrec + pool * (size_t)(naux > 1 ? naux : 1);
    }
};
Sizes sizes(int32_t max_cols)
{
    const size_t m = (size_t)max_cols;
    Sizes z;
    z.acc = sizeof(double) * (60 * (NL + 1) + 14) * m;                        // band flux planes
    z.iws = sizeof(int) * 7 * m;                                              // 6 column words + day-list segments
    z.dws = sizeof(double) * (4 + 2 * (size_t)NL) * m;                        // column words, layer logs, tz
    z.rtr = (sizeof(double) * 12 * (NL + 3) + 2 * sizeof(uint64_t)) * m;      // NAM-route overlap factors
    z.sflux_d = sizeof(double) * 112 * m;
    z.sflux_f = sizeof(float) * 112 * m;
    z.lwrec = lw_glane_env() ? lwrec_bytes_per_col<double>(64 + 1) * m + 3 * 256 : 0;  // LW layer records
    // level pool: room for every 64-column block of a tile in the largest slot layout (FP64 SW:
    // 2 values, FP32 SW: 7 values, x 16 g x levels), at most WR_POOL_GB (default 6) GB per stream
    const int nl_max = 64;
    const size_t lane_max = std::max(cache_slot_bytes(SW_NV, 16, nl_max + 1, (int)sizeof(double)),
                                     cache_slot_bytes(sw_nv_mc<float>(), 16, nl_max + 1, (int)sizeof(float)));
    const size_t need = (m + 63) / 64 * lane_max;
    const char *pe = getenv("WR_POOL_GB");
    const double gb = pe ? atof(pe) : 6.0;
    size_t cap = (size_t)(gb * 1e9);
    if (cap < lane_max) cap = lane_max;
    z.pool = need < cap ? need : cap;
    // WR_BAND_STREAMS=0 keeps the band kernels on the caller's stream
    const char *e = getenv("WR_BAND_STREAMS");
    const int na = e ? atoi(e) : WR_NAUX;
    z.naux = na < 0 ? 0 : (na > WR_NAUX ? WR_NAUX : na);
    return z;
}

// ---- work space of wr_run_mixed_noalloc ---------------------------------------------------------
// The mixed-precision tile work space for up to WR_NOALLOC_MAX_COLS columns of up to 64 radiation
// layers, in one static device array (every device the library runs on holds its own copy until the
// process ends).  Layout (256-byte aligned parts): FP32 band flux planes, column words, layer logs and
// level temperatures, NAM-route overlap factors, FP32 solar source, and one band stream's level pool
// with a slot for every 64-column block.  The kernels write every word of it that they read, so
// nothing needs clearing between calls; calls that overlap on one device race on it.
struct NaLayout {
    size_t acc, iws, dws, rtr, sflux, pool, total;
};
constexpr size_t na_al(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr NaLayout na_layout()
{
    const size_t m = WR_NOALLOC_MAX_COLS;
    NaLayout z{};
    size_t off = 0;
    z.acc = off, off += na_al(sizeof(float) * (60 * (NL + 1) + 14) * m);
    z.iws = off, off += na_al(sizeof(int) * 7 * m);
    z.dws = off, off += na_al(sizeof(double) * (4 + 2 * (size_t)NL) * m);
    z.rtr = off, off += na_al((sizeof(double) * 12 * (NL + 3) + 2 * sizeof(uint64_t)) * m);
    z.sflux = off, off += na_al(sizeof(float) * 112 * m);
    const size_t lane_max = std::max(cache_slot_bytes(sw_nv_mc<float>(), 16, 64 + 1, (int)sizeof(float)),
                                     cache_slot_bytes(LW_NV_F32, 16, 64 + 1, (int)sizeof(float)));
    z.pool = off, off += (m + 63) / 64 * lane_max;
    z.total = off;
    return z;
}
constexpr NaLayout NA = na_layout();
static_assert(NA.total <= ((size_t)256 << 20), "the no-alloc work space stays under 256 MiB");

static __device__ __align__(256) char scratch[NA.total];

// Device address of `scratch` on each device, looked up once per device (hipGetSymbolAddress with the
// device current) by wr_create, or by the first wr_run_mixed_noalloc on a device without a context.
constexpr int NA_MAXDEV = 64;
std::mutex na_mu;
char *na_ptr[NA_MAXDEV];

int noalloc_setup(int device)
{
    if (device < 0 || device >= NA_MAXDEV) return WR_ERR_ARG;
    std::lock_guard<std::mutex> lk(na_mu);
    if (na_ptr[device]) return WR_OK;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return WR_ERR_HIP;
    na_ptr[device] = (char *)p;
    return WR_OK;
}

}  // namespace

extern "C" {

int wr_create(wr_ctx **out, const char *table_path, int device, int32_t max_cols)
{
    if (!out || !table_path || max_cols <= 0) return WR_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return WR_ERR_HIP;
    TFile f;
    int rc = read_tables(table_path, f);
    if (rc != WR_OK) return rc;
    DevGuard g(device);
    wr_ctx *c = (wr_ctx *)calloc(1, sizeof(wr_ctx));
    if (!c) return WR_ERR_NOMEM;
    c->device = device;
    c->max_cols = max_cols;
    c->nl_max = 64;  // McICA bits of a g-point over the layers are one 64-bit word
    rc = upload<double>(f, &c->tab_buf_d, c->host_d, &c->tabs_d);
    if (rc == WR_OK) rc = upload<float>(f, &c->tab_buf_f, c->host_f, &c->tabs_f);
    if (rc == WR_OK) {
        // both precisions reach the FP64 SW exp table (the mixed path's FP64 R/T branch)
        c->host_d.sw_exp_tbl64 = c->host_d.sw_exp_tbl;
        c->host_f.sw_exp_tbl64 = c->host_d.sw_exp_tbl;
        if (hipMemcpy(c->tabs_d, &c->host_d, sizeof(Tabs<double>), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->tabs_f, &c->host_f, sizeof(Tabs<float>), hipMemcpyHostToDevice) != hipSuccess)
            rc = WR_ERR_HIP;
    }
    if (rc == WR_OK) rc = noalloc_setup(device);
    if (rc == WR_OK) rc = wr_init_luts(c, nullptr);
    if (rc == WR_OK) {
        const double *src[4] = {c->host_d.exp_tbl, c->host_d.tfn_tbl, c->host_d.tau_tbl, c->host_d.sw_exp_tbl};
        float *dst[4] = {(float *)c->host_f.exp_tbl, (float *)c->host_f.tfn_tbl, (float *)c->host_f.tau_tbl,
                         (float *)c->host_f.sw_exp_tbl};
        for (int q = 0; q < 4; ++q)
            hipLaunchKernelGGL(k_d2f, dim3((NTBL + 256) / 256), dim3(256), 0, nullptr, src[q], dst[q], NTBL + 1);
        if (hipGetLastError() != hipSuccess) rc = WR_ERR_HIP;
    }
    const Sizes z = sizes(max_cols);
    if (rc == WR_OK) {
        if (hipMalloc((void **)&c->acc, z.acc) != hipSuccess || hipMalloc((void **)&c->iws, z.iws) != hipSuccess ||
            hipMalloc((void **)&c->dws, z.dws) != hipSuccess || hipMalloc(&c->rtr, z.rtr) != hipSuccess ||
            hipMalloc((void **)&c->sflux_d, z.sflux_d) != hipSuccess ||
            hipMalloc((void **)&c->sflux_f, z.sflux_f) != hipSuccess ||
            (z.lwrec && hipMalloc(&c->lwrec, z.lwrec) != hipSuccess))
            rc = WR_ERR_NOMEM;
        c->lw_glane = z.lwrec > 0;
        if (hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
            hipDeviceGetAttribute(&c->lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
            rc = WR_ERR_HIP;
        c->lw_lds = lds_mask<float>("WR_LW_LDS", LW_LDS_DEFAULT, c->host_f.lw_img, 16, c->lds_max);
        c->lw_lds64 = lds_mask<double>("WR_LW_LDS64", LW_LDS64_DEFAULT, c->host_d.lw_img, 16, c->lds_max);
        c->sw_lds = lds_mask<float>("WR_SW_LDS", SW_LDS_DEFAULT, c->host_f.sw_img, 14, c->lds_max);
        c->lwr_lds = lds_mask<float>("WR_LWR_LDS", LWR_LDS_DEFAULT, c->host_f.lw_img, 16, c->lds_max);
        c->lwr_lds64 = lds_mask<double>("WR_LWR_LDS64", LWR_LDS64_DEFAULT, c->host_d.lw_img, 16, c->lds_max);
        c->swr_lds = lds_mask<float>("WR_SWR_LDS", SWR_LDS_DEFAULT, c->host_f.sw_img, 14, c->lds_max);
        c->swr_lds64 = lds_mask<double>("WR_SWR_LDS64", SWR_LDS64_DEFAULT, c->host_d.sw_img, 14, c->lds_max);
        c->sw_lds64 = lds_mask<double>("WR_SW_LDS64", SW_LDS64_DEFAULT, c->host_d.sw_img, 14, c->lds_max);
    }
    if (rc == WR_OK) {
        c->pool_bytes = z.pool;
        if (hipMalloc((void **)&c->pool[0], c->pool_bytes) != hipSuccess) rc = WR_ERR_NOMEM;
    }
    if (rc == WR_OK) {
        // WR_BAND_STREAMS=0 keeps the band kernels on the caller's stream
        c->naux = z.naux;
        for (int q = 0; q < c->naux && rc == WR_OK; ++q)
            if (hipStreamCreateWithFlags(&c->aux[q], hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&c->join[q], hipEventDisableTiming) != hipSuccess)
                rc = WR_ERR_HIP;
        // one level pool per band stream (stream 0 shares the caller's), all allocated here: no later
        // call allocates
        for (int q = 1; q < c->naux && rc == WR_OK; ++q)
            if (hipMalloc((void **)&c->pool[q], c->pool_bytes) != hipSuccess) rc = WR_ERR_NOMEM;
        if (rc == WR_OK && c->naux > 0 && hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) != hipSuccess)
            rc = WR_ERR_HIP;
    }
    if (rc == WR_OK && hipDeviceSynchronize() != hipSuccess) rc = WR_ERR_HIP;
    if (rc != WR_OK) {
        wr_destroy(c);
        return rc;
    }
    *out = c;
    return WR_OK;
}

uint64_t wr_context_bytes(int32_t max_cols)
{
    if (max_cols <= 0) return 0;
    return (uint64_t)sizes(max_cols).total() + WR_TABLE_BYTES;
}

int wr_run_fp64(wr_ctx *ctx, const wr_columns *in, const wr_params *par, const wr_outputs *out, const wr_diag *diag,
                void *stream)
{
    return run(ctx, in, par, out, diag, stream, MODE_FP64);
}

int wr_run_mixed(wr_ctx *ctx, const wr_columns *in, const wr_params *par, const wr_outputs *out, const wr_diag *diag,
                 void *stream)
{
    return run(ctx, in, par, out, diag, stream, MODE_MIXED);
}

int wr_run_mixed_noalloc(wr_ctx *ctx, const wr_columns *in, const wr_params *par, const wr_outputs *out, void *stream)
{
    if (!ctx || !in) return WR_ERR_ARG;
    if (in->ncol > WR_NOALLOC_MAX_COLS) return WR_ERR_TOO_LARGE;
    if (in->nlay + 1 > 64) return WR_ERR_ARG;
    int rc = WR_OK;
    {
        DevGuard g(ctx->device);
        rc = noalloc_setup(ctx->device);
    }
    if (rc != WR_OK) return rc;
    char *base = na_ptr[ctx->device];
    // the context's tables, streams and timing with the work space of the static array; the band
    // kernels stay on the caller's stream (one level pool)
    wr_ctx t = *ctx;
    t.max_cols = WR_NOALLOC_MAX_COLS;
    t.acc = (double *)(base + NA.acc);
    t.iws = (int *)(base + NA.iws);
    t.dws = (double *)(base + NA.dws);
    t.rtr = base + NA.rtr;
    t.sflux_d = nullptr;
    t.sflux_f = (float *)(base + NA.sflux);
    t.lwrec = nullptr;  // lane-per-column LW kernels
    t.lw_glane = 0;
    for (int q = 0; q < WR_NAUX; ++q) t.pool[q] = nullptr;
    t.pool[0] = base + NA.pool;
    t.pool_bytes = NA.total - NA.pool;
    t.naux = 0;
    rc = run(&t, in, par, out, nullptr, stream, MODE_MIXED);
    ctx->ev_n = t.ev_n;
    return rc;
}

int wr_run_band(wr_ctx *ctx, const wr_columns *in, const wr_params *par, const wr_outputs *out, int32_t band,
                int32_t mixed, void *stream)
{
    if (band < 0 || band >= WR_NBANDS_LW + WR_NBANDS_SW) return WR_ERR_ARG;
    return run(ctx, in, par, out, nullptr, stream, mixed ? MODE_MIXED : MODE_FP64, band);
}

int wr_surrogate_attach(wr_ctx *c, const wr_surrogate *w)
{
    if (!c || !w || !w->w1 || !w->w2 || !w->w3 || !w->b1 || !w->b2 || !w->b3 || !w->fmean || !w->fstd || !w->fmin ||
        !w->fmax || !w->tmean || !w->tstd)
        return WR_ERR_ARG;
    DevGuard g(c->device);
    if (!c->sx) {
        const size_t S = (size_t)c->nl_max * c->max_cols;
        if (hipMalloc((void **)&c->sx, S * 8 * sizeof(float)) != hipSuccess ||
            hipMalloc(&c->sy, (S + 64) * 504 * 2) != hipSuccess ||
            hipMalloc((void **)&c->sflag, (size_t)c->max_cols * sizeof(int)) != hipSuccess)
            return WR_ERR_NOMEM;
    }
    c->sur = *w;
    c->has_sur = 1;
    return WR_OK;
}

int wr_run_surrogate(wr_ctx *ctx, const wr_columns *in, const wr_params *par, const wr_outputs *out, int32_t *col_flag,
                     void *stream)
{
    return run(ctx, in, par, out, nullptr, stream, MODE_SURR, -1, col_flag);
}

int wr_cal_cldfra(int32_t im, int32_t jm, int32_t lm, int32_t i_lo, int32_t i_hi, int32_t j_lo, int32_t j_hi,
                  int32_t f_qc, int32_t f_qi, float *cldfra, const float *qc, const float *qi, void *stream)
{
    if (im < 0 || jm < 0 || lm < 0 || !cldfra) return WR_ERR_ARG;
    const size_t tot = (size_t)im * jm * lm;
    if (tot == 0) return WR_OK;
    if (f_qc && !qc) return WR_ERR_ARG;
    if (f_qc && f_qi && !qi) return WR_ERR_ARG;
    hipLaunchKernelGGL(k_cal_cldfra, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, (hipStream_t)stream, im, jm, lm,
                       i_lo, i_hi, j_lo, j_hi, f_qc, f_qi, cldfra, qc, qi);
    return hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
}

int wr_mcica_masks(int32_t ncol, int32_t nrl, const double *cldf, const uint64_t *col_id, uint64_t seed, int32_t iovr,
                   uint32_t *lw_mask, uint32_t *sw_mask, void *stream)
{
    if (ncol < 0 || nrl < 1 || nrl > 64 || iovr < 0 || iovr > 2) return WR_ERR_ARG;
    if (ncol == 0 || (!lw_mask && !sw_mask)) return WR_OK;
    if (!cldf) return WR_ERR_ARG;
    hipLaunchKernelGGL(k_mcica_masks, dim3((ncol + 63) / 64), dim3(64), 0, (hipStream_t)stream, ncol, nrl, cldf, col_id,
                       seed, iovr, lw_mask, sw_mask);
    return hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
}

int wr_timing(wr_ctx *c, int32_t enable)
{
    if (!c) return WR_ERR_ARG;
    DevGuard g(c->device);
    if (enable && !c->ev) {
        const int cap = 8192;
        c->ev = (hipEvent_t *)calloc(2 * (size_t)cap, sizeof(hipEvent_t));
        c->ev_kid = (int *)calloc(cap, sizeof(int));
        if (!c->ev || !c->ev_kid) return WR_ERR_NOMEM;
        for (int j = 0; j < 2 * cap; ++j)
            if (hipEventCreate(&c->ev[j]) != hipSuccess) return WR_ERR_HIP;
        c->ev_cap = cap;
    }
    c->timing = enable ? 1 : 0;
    c->ev_n = 0;
    return WR_OK;
}

int wr_timing_read(wr_ctx *c, int32_t *kid, float *ms, int32_t cap, int32_t *n)
{
    if (!c || !n || (cap > 0 && (!kid || !ms))) return WR_ERR_ARG;
    DevGuard g(c->device);
    const int m = c->ev_n < cap ? c->ev_n : cap;
    for (int j = 0; j < m; ++j) {
        if (hipEventSynchronize(c->ev[2 * j + 1]) != hipSuccess) return WR_ERR_HIP;
        float t = 0.0f;
        if (hipEventElapsedTime(&t, c->ev[2 * j], c->ev[2 * j + 1]) != hipSuccess) return WR_ERR_HIP;
        kid[j] = c->ev_kid[j];
        ms[j] = t;
    }
    *n = m;
    c->ev_n = 0;
    return WR_OK;
}

const char *wr_kernel_name(int32_t kid)
{
    static const char *lw[16] = {"k_lw_b01", "k_lw_b02", "k_lw_b03", "k_lw_b04", "k_lw_b05", "k_lw_b06",
                                 "k_lw_b07", "k_lw_b08", "k_lw_b09", "k_lw_b10", "k_lw_b11", "k_lw_b12",
                                 "k_lw_b13", "k_lw_b14", "k_lw_b15", "k_lw_b16"};
    static const char *sw[14] = {"k_sw_b16", "k_sw_b17", "k_sw_b18", "k_sw_b19", "k_sw_b20", "k_sw_b21", "k_sw_b22",
                                 "k_sw_b23", "k_sw_b24", "k_sw_b25", "k_sw_b26", "k_sw_b27", "k_sw_b28", "k_sw_b29"};
    if (kid == KID_ZERO) return "k_zero";
    if (kid == KID_COLPREP) return "k_colprep";
    if (kid == KID_DAYLIST) return "k_daylist";
    if (kid >= KID_LW0 && kid < KID_LW0 + 16) return lw[kid - KID_LW0];
    if (kid >= KID_SW0 && kid < KID_SW0 + 14) return sw[kid - KID_SW0];
    if (kid == KID_POST) return "k_post";
    if (kid == KID_DIAG) return "k_diag";
    if (kid == KID_FEAT) return "k_features";
    if (kid == KID_MLP) return "mlp_gemm";
    if (kid == KID_RTRFAC) return "k_rtr_fac";
    if (kid == KID_LWREC) return "k_lwrec";
    return "?";
}

void wr_default_params(wr_params *p)
{
    memset(p, 0, sizeof(*p));
    p->mcica = 1;
    p->iovr = 1;
    p->seed = 0;
    p->solcon = 1368.22;
    p->re_liq = 10.0;
    p->re_ice = 50.0;
    // co2, n2o, ch4, o2, co, cfc11, cfc12, cfc22, ccl4, cfc113 (phys_rrtm_getgases.cu:56-65)
    const double v[10] = {410.0e-6, 0.33e-6, 1.9e-6, 0.209, 1.50e-8, 3.520e-10, 6.358e-10, 1.500e-10, 1.397e-10, 8.2000e-11};
    for (int g = 0; g < 10; ++g) p->vmr[g] = v[g];
}

const char *wr_strerror(int code)
{
    switch (code) {
    case WR_OK: return "ok";
    case WR_ERR_ARG: return "invalid argument (size or null pointer)";
    case WR_ERR_TABLES: return "table file missing or malformed";
    case WR_ERR_HIP: return "HIP runtime error";
    case WR_ERR_NOMEM: return "device allocation failed";
    case WR_ERR_UNSUPPORTED: return "option not supported by this build";
    case WR_ERR_TOO_LARGE: return "more columns than the static work space of wr_run_mixed_noalloc holds";
    default: return "unknown error";
    }
}

void wr_destroy(wr_ctx *c)
{
    if (!c) return;
    DevGuard g(c->device);
    (void)hipDeviceSynchronize();
    (void)hipFree(c->tab_buf_d);
    (void)hipFree(c->tab_buf_f);
    (void)hipFree(c->tabs_d);
    (void)hipFree(c->tabs_f);
    (void)hipFree(c->acc);
    (void)hipFree(c->iws);
    (void)hipFree(c->dws);
    (void)hipFree(c->rtr);
    (void)hipFree(c->sflux_d);
    (void)hipFree(c->sflux_f);
    (void)hipFree(c->lwrec);
    (void)hipFree(c->sx);
    (void)hipFree(c->sy);
    (void)hipFree(c->sflag);
    for (int q = 0; q < WR_NAUX; ++q) {
        (void)hipFree(c->pool[q]);
        if (c->aux[q]) (void)hipStreamDestroy(c->aux[q]);
        if (c->join[q]) (void)hipEventDestroy(c->join[q]);
    }
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->ev)
        for (int j = 0; j < 2 * c->ev_cap; ++j)
            if (c->ev[j]) (void)hipEventDestroy(c->ev[j]);
    free(c->ev);
    free(c->ev_kid);
    free(c);
}

}  // extern "C"
