// wr_serve.hip -- served tiles: a slot is a context with its own stream, staging buffers for one tile,
// a HIP graph of a whole-tile call captured once, and a completion event.  A tile of a request is
// gathered into the staging buffers by one kernel (which also writes the request's McICA seed and the
// tile's valid column count into the device words the call's kernels read), the graph is replayed,
// one kernel scatters the outputs into the request's arrays, and the event is recorded: four stream
// operations per tile, no allocation, no host synchronisation.
#include <hip/hip_runtime.h>

#include "wr_internal.h"

namespace {
using namespace wr;

constexpr int NIN = 11;   // p_lev, t_lay, q_lay, o3_lay, cldf, qc, qi, t_sfc, emis, alb, cosz
constexpr int NOUT = 7;   // hr_lw, hr_sw, flw_up, flw_dn, fsw_up, fsw_dn, fsw_dn_sfc_clr

struct SlotDev {  // the slot's device current for a call, the previous one restored afterwards
    int prev = -1;
    explicit SlotDev(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~SlotDev()
    {
        int cur;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

// column copy of up to NIN fields (rows[f] rows each): dst[f][k * dld + j] = src[f][k * sld + c(j)],
// j < n, c(j) = c0 + j (or (c0 + j) % wrap when wrap > 0); one extra grid row writes the column ids
// (ids[c(j)], or c(j) when ids is null) and, from its first thread, the seed and valid-count words
struct Copy {
    const WR_G double *src[NIN];
    WR_G double *dst[NIN];
    int rows[NIN];
    int nf;
    long long sld, dld;
    int c0, n, wrap;
    const WR_G uint64_t *ids;
    WR_G uint64_t *ids_dst;
    WR_G uint64_t *seed_dst;
    WR_G int32_t *ncol_dst;
    uint64_t seed;
    int ncol_word;
};

__global__ void __launch_bounds__(256) k_copy_cols(Copy p)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    int r = blockIdx.y;
    int f = 0;
    while (f < p.nf && r >= p.rows[f]) r -= p.rows[f++];
    if (j < p.n) {
        const long long c = p.wrap > 0 ? (long long)((p.c0 + j) % p.wrap) : (long long)(p.c0 + j);
        if (f < p.nf) {
            if (p.src[f]) p.dst[f][(size_t)r * p.dld + j] = p.src[f][(size_t)r * p.sld + c];
        } else if (p.ids_dst) {
            p.ids_dst[j] = p.ids ? p.ids[c] : (uint64_t)c;
        }
    }
    if (f >= p.nf && j == 0) {
        if (p.seed_dst) *p.seed_dst = p.seed;
        if (p.ncol_dst) *p.ncol_dst = p.ncol_word;
    }
}

int launch_copy(const Copy &p, hipStream_t s)
{
    int rows = 0;
    for (int f = 0; f < p.nf; ++f) rows += p.rows[f];
    rows += 1;
    hipLaunchKernelGGL(k_copy_cols, dim3((p.n + 255) / 256, rows), dim3(256), 0, s, p);
    return hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
}

const double *in_field(const wr_columns *c, int f)
{
    const double *const a[NIN] = {c->p_lev, c->t_lay, c->q_lay, c->o3_lay, c->cldf, c->qc,
                                  c->qi,    c->t_sfc, c->emis,  c->alb,    c->cosz};
    return a[f];
}
double *out_field(const wr_outputs *o, int f)
{
    double *const a[NOUT] = {o->hr_lw, o->hr_sw, o->flw_up, o->flw_dn, o->fsw_up, o->fsw_dn, o->fsw_dn_sfc_clr};
    return a[f];
}
int in_rows(int f, int nlay) { return f == 0 ? nlay + 1 : (f <= 6 ? nlay : 1); }
int out_rows(int f, int nlay) { return f <= 1 ? nlay : (f <= 5 ? nlay + 2 : 1); }

}  // namespace

struct wr_slot {
    wr_ctx *ctx;
    int tile, nlay, precision, graphs;
    wr_params par;
    hipStream_t stream;
    hipEvent_t done;
    hipGraph_t graph;
    hipGraphExec_t exec;
    double *in[NIN];
    double *out[NOUT];
    uint64_t *ids, *seed;
    int32_t *ncol;
    wr_columns cols;     // the staging buffers as the call's inputs
    wr_outputs outs;
};

namespace {

int slot_call(wr_slot *sl)
{
    if (sl->precision == 0) return wr_run_fp64(sl->ctx, &sl->cols, &sl->par, &sl->outs, nullptr, sl->stream);
    if (sl->precision == 1) return wr_run_mixed(sl->ctx, &sl->cols, &sl->par, &sl->outs, nullptr, sl->stream);
    return wr_run_surrogate(sl->ctx, &sl->cols, &sl->par, &sl->outs, nullptr, sl->stream);
}

void slot_free(wr_slot *sl)
{
    if (sl->stream) (void)hipStreamSynchronize(sl->stream);
    if (sl->exec) (void)hipGraphExecDestroy(sl->exec);
    if (sl->graph) (void)hipGraphDestroy(sl->graph);
    for (int f = 0; f < NIN; ++f) (void)hipFree(sl->in[f]);
    for (int f = 0; f < NOUT; ++f) (void)hipFree(sl->out[f]);
    (void)hipFree(sl->ids);
    (void)hipFree(sl->seed);
    (void)hipFree(sl->ncol);
    if (sl->done) (void)hipEventDestroy(sl->done);
    if (sl->stream) (void)hipStreamDestroy(sl->stream);
    delete sl;
}

}  // namespace

extern "C" {

int wr_slot_create(wr_slot **out, wr_ctx *ctx, int32_t precision, const wr_params *par, int32_t graphs,
                   const wr_columns *filler)
{
    if (!out || !ctx || !par || !filler || precision < 0 || precision > 2) return WR_ERR_ARG;
    *out = nullptr;
    if (precision == 2 && !ctx->has_sur) return WR_ERR_ARG;
    const int nlay = filler->nlay;
    if (filler->ncol < 1 || nlay < 2 || nlay + 1 > ctx->nl_max) return WR_ERR_ARG;
    for (int f = 0; f < NIN; ++f)
        if (!in_field(filler, f)) return WR_ERR_ARG;
    SlotDev g(ctx->device);
    wr_slot *sl = new wr_slot();
    sl->ctx = ctx, sl->tile = ctx->max_cols, sl->nlay = nlay, sl->precision = precision, sl->graphs = graphs != 0;
    sl->par = *par;
    const size_t T = (size_t)sl->tile;
    bool ok = hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&sl->done, hipEventDisableTiming) == hipSuccess;
    for (int f = 0; ok && f < NIN; ++f) ok = hipMalloc((void **)&sl->in[f], in_rows(f, nlay) * T * sizeof(double)) == hipSuccess;
    for (int f = 0; ok && f < NOUT; ++f)
        ok = hipMalloc((void **)&sl->out[f], out_rows(f, nlay) * T * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc((void **)&sl->ids, T * sizeof(uint64_t)) == hipSuccess &&
         hipMalloc((void **)&sl->seed, sizeof(uint64_t)) == hipSuccess &&
         hipMalloc((void **)&sl->ncol, sizeof(int32_t)) == hipSuccess;
    if (!ok) {
        slot_free(sl);
        return WR_ERR_NOMEM;
    }
    wr_columns &c = sl->cols;
    c = wr_columns();
    c.ncol = sl->tile, c.nlay = nlay;
    c.p_lev = sl->in[0], c.t_lay = sl->in[1], c.q_lay = sl->in[2], c.o3_lay = sl->in[3], c.cldf = sl->in[4];
    c.qc = sl->in[5], c.qi = sl->in[6], c.t_sfc = sl->in[7], c.emis = sl->in[8], c.alb = sl->in[9], c.cosz = sl->in[10];
    c.col_id = sl->ids, c.seed_dev = sl->seed, c.ncol_dev = sl->ncol;
    wr_outputs &o = sl->outs;
    o.hr_lw = sl->out[0], o.hr_sw = sl->out[1], o.flw_up = sl->out[2], o.flw_dn = sl->out[3], o.fsw_up = sl->out[4];
    o.fsw_dn = sl->out[5], o.fsw_dn_sfc_clr = sl->out[6];
    // every staging column holds a valid column (the filler's, repeated), so a replay over the whole
    // tile never computes on garbage; the ids, seed and count words start as a full tile's
    Copy p = {};
    for (int f = 0; f < NIN; ++f) p.src[f] = gptr(in_field(filler, f)), p.dst[f] = gptr(sl->in[f]), p.rows[f] = in_rows(f, nlay);
    p.nf = NIN, p.sld = filler->ncol, p.dld = sl->tile, p.c0 = 0, p.n = sl->tile, p.wrap = filler->ncol;
    p.ids = nullptr, p.ids_dst = gptr(sl->ids), p.seed_dst = gptr(sl->seed), p.ncol_dst = gptr(sl->ncol);
    p.seed = par->seed, p.ncol_word = sl->tile;
    int rc = launch_copy(p, sl->stream);
    if (rc == WR_OK) rc = slot_call(sl);  // warm-up: first-call work stays out of the graph
    if (rc == WR_OK && hipStreamSynchronize(sl->stream) != hipSuccess) rc = WR_ERR_HIP;
    if (rc == WR_OK && sl->graphs) {
        if (hipStreamBeginCapture(sl->stream, hipStreamCaptureModeRelaxed) != hipSuccess) rc = WR_ERR_HIP;
        const int rc_call = rc == WR_OK ? slot_call(sl) : rc;
        hipGraph_t gr = nullptr;
        const bool ended = rc == WR_OK && hipStreamEndCapture(sl->stream, &gr) == hipSuccess;
        sl->graph = gr;
        if (rc == WR_OK) rc = rc_call;
        if (rc == WR_OK && (!ended || hipGraphInstantiateWithFlags(&sl->exec, gr, 0) != hipSuccess)) rc = WR_ERR_HIP;
    }
    if (rc != WR_OK) {
        slot_free(sl);
        return rc;
    }
    *out = sl;
    return WR_OK;
}

int wr_slot_launch(wr_slot *sl, const wr_columns *req, const wr_outputs *res, int32_t c0, int32_t n, uint64_t seed)
{
    if (!sl || !req || !res || n < 1 || n > sl->tile || c0 < 0 || req->nlay != sl->nlay ||
        (int64_t)c0 + n > (int64_t)req->ncol)
        return WR_ERR_ARG;
    for (int f = 0; f < NIN; ++f)
        if (!in_field(req, f)) return WR_ERR_ARG;
    for (int f = 0; f < NOUT - 1; ++f)
        if (!out_field(res, f)) return WR_ERR_ARG;
    SlotDev g(sl->ctx->device);
    Copy p = {};
    for (int f = 0; f < NIN; ++f) p.src[f] = gptr(in_field(req, f)), p.dst[f] = gptr(sl->in[f]), p.rows[f] = in_rows(f, sl->nlay);
    p.nf = NIN, p.sld = req->ncol, p.dld = sl->tile, p.c0 = c0, p.n = n, p.wrap = 0;
    p.ids = gptr(req->col_id), p.ids_dst = gptr(sl->ids), p.seed_dst = gptr(sl->seed), p.ncol_dst = gptr(sl->ncol);
    p.seed = seed, p.ncol_word = n;
    int rc = launch_copy(p, sl->stream);
    if (rc != WR_OK) return rc;
    if (sl->exec) {
        if (hipGraphLaunch(sl->exec, sl->stream) != hipSuccess) return WR_ERR_HIP;
    } else {
        rc = slot_call(sl);
        if (rc != WR_OK) return rc;
    }
    Copy q = {};   // outputs back: staging [k][tile] -> the request's [k][ncol] at columns c0..c0+n
    for (int f = 0; f < NOUT; ++f) {
        double *d = out_field(res, f);
        q.src[f] = d ? gptr((const double *)sl->out[f]) : nullptr;
        q.dst[f] = d ? gptr(d + c0) : nullptr;
        q.rows[f] = out_rows(f, sl->nlay);
    }
    q.nf = NOUT, q.sld = sl->tile, q.dld = req->ncol, q.c0 = 0, q.n = n, q.wrap = 0;
    rc = launch_copy(q, sl->stream);
    if (rc != WR_OK) return rc;
    return hipEventRecord(sl->done, sl->stream) == hipSuccess ? WR_OK : WR_ERR_HIP;
}

int wr_slot_query(wr_slot *sl)
{
    if (!sl) return -WR_ERR_ARG;
    const hipError_t e = hipEventQuery(sl->done);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return -WR_ERR_HIP;
}

int wr_slot_sync(wr_slot *sl)
{
    if (!sl) return WR_ERR_ARG;
    return hipEventSynchronize(sl->done) == hipSuccess ? WR_OK : WR_ERR_HIP;
}

void wr_slot_destroy(wr_slot *sl)
{
    if (!sl) return;
    SlotDev g(sl->ctx->device);
    slot_free(sl);
}

}  // extern "C"
