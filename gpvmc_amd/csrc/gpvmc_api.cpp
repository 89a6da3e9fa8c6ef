// C ABI of libgpvmc_hip.so (declared in include/gpvmc.h): context, device buffers,
// argument validation and kernel dispatch.  The kernels live in vmc_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gpvmc.h"
#include "vmc_kernels.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(GPVMC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Lattice tables of SquareLattice (SquareLattice.cpp:22-27): edges with x outer, y inner,
// first (x,y)->(x,y+1) then (x,y)->(x+1,y); each vertex lists its edges in creation order
// (Lattice.h:79-84).
void build_lattice(int L, std::vector<int32_t> &nbr, std::vector<int32_t> &edge4)
{
    const int Ns = L * L;
    std::vector<int> ea, eb;
    for (int x = 0; x < L; ++x)
        for (int y = 0; y < L; ++y) {
            ea.push_back(y * L + x);
            eb.push_back(((y + 1) % L) * L + x);
            ea.push_back(y * L + x);
            eb.push_back(y * L + (x + 1) % L);
        }
    nbr.assign(4 * Ns, -1);
    std::vector<int> cnt(Ns, 0);
    for (size_t e = 0; e < ea.size(); ++e) {
        nbr[4 * ea[e] + cnt[ea[e]]++] = eb[e];
        nbr[4 * eb[e] + cnt[eb[e]]++] = ea[e];
    }
    edge4.assign(4 * ea.size(), 0);
    for (size_t e = 0; e < ea.size(); ++e) {
        int a = ea[e], b = eb[e], da = -1, db = -1;
        for (int d = 0; d < 4; ++d) {
            if (nbr[4 * a + d] == b && da < 0) da = d;
            if (nbr[4 * b + d] == a && db < 0) db = d;
        }
        edge4[4 * e] = a;
        edge4[4 * e + 1] = da;
        edge4[4 * e + 2] = b;
        edge4[4 * e + 3] = db;
    }
}

int default_lpr(int prec, int N)
{
    // the first registered launch shape of (precision, N) is the default
    int res = 0, nt = 0;
    return gv_sweep_supported(prec, N, 0, &res, &nt);
}

}  // namespace

struct gpvmc_ctx {
    gpvmc_config cfg;
    int L = 0, Ns = 0, N = 0, W = 0, G = 0, prec = 0, lpr = 0, res = 0, nthreads = 0;
    int variant = GPVMC_SWEEP_DEFAULT;  // gpvmc_set_sweep_variant
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    size_t esize = 0;  // bytes per complex element of the ratio tables
    void *d_G = nullptr;          // ratio tables, working precision
    double2 *d_Ai64 = nullptr;    // FP64 inverse of the last init/refresh
    int32_t *d_site = nullptr, *d_part = nullptr, *d_group = nullptr, *d_order = nullptr;
    int32_t *d_nbr = nullptr, *d_status = nullptr;
    int32_t *d_flag = nullptr;    // one int: gv_launch_raninit's incomplete-draw flag
    int4 *d_edge = nullptr;
    double *d_logabs = nullptr, *d_phase = nullptr, *d_acc = nullptr, *d_drift = nullptr;
    uint64_t *d_counter = nullptr;
    double2 *d_occ64 = nullptr;
    float2 *d_occ32 = nullptr;
    std::vector<int> have_orb;
    std::vector<int32_t> groups;
    bool initialised = false;
    double sweep_ms = 0, refresh_ms = 0;
    int64_t n_sweep = 0, n_refresh = 0;
    // scratch for walk traces / replay
    void *d_scratch = nullptr;
    size_t scratch_bytes = 0;
    // StatSpinStruct sums [W][ss_block(Ns)], allocated on first request
    double *d_ss = nullptr;
    // working matrices of the streamed FP64 inversion (N > 128 only)
    double2 *d_gjwork = nullptr;
    // channel (exciton) contexts: N is the larger flavour size, part is [W][2][N]
    bool channel = false;
    int n_up = 0, n_dn = 0, nk = 0;
    int32_t *d_cur = nullptr, *d_cstate = nullptr;
    double *d_weight = nullptr, *d_ho = nullptr;
    double2 *d_tab64 = nullptr;
    float2 *d_tab32 = nullptr;
    float2 *d_Ai32 = nullptr;
    double2 *d_work = nullptr;
    int32_t *d_hop_n = nullptr, *d_hop_orb = nullptr, *d_hop_new = nullptr, *d_new_off = nullptr,
            *d_new_orb = nullptr, *d_hopsign = nullptr;
    std::vector<int32_t> state_orb;  // [n_chan][nk][n_up + n_dn]
    int n_chan = 1;
    int32_t *d_chan = nullptr;       // [W] channel of each walker
    std::vector<int32_t> chans;
    // register-resident channel kernel (exc_reg_kernels.hip) when the walker fits
    bool regk = false;
    int rk_cb = 0, rk_rbw = 0, nq0 = 0, nq1 = 0, rb0 = 0, rb1 = 0;
    void *d_M = nullptr, *d_tabT = nullptr, *d_zscr = nullptr;
    // measurements not yet folded into d_ho: hb_cap records of 2 nk + 1 double2 per walker
    double2 *d_hbuf = nullptr;
    int hb_cap = 0;
    void *d_Q = nullptr;  // streamed channel kernel: Q rows [W][2][nqm][N]
    void *d_Y = nullptr;  // streamed channel kernel measurement: [W][2][N + nqm][N]
    int32_t *d_qorb = nullptr;
};

static void free_ctx(gpvmc_ctx *c)
{
    if (!c) return;
    void *ptrs[] = {c->d_G, c->d_Ai64, c->d_site, c->d_part, c->d_group, c->d_order, c->d_nbr,
                    c->d_status, c->d_edge, c->d_logabs, c->d_phase, c->d_acc, c->d_drift,
                    c->d_counter, c->d_occ64, c->d_occ32, c->d_scratch, c->d_ss, c->d_gjwork,
                    c->d_cur, c->d_cstate, c->d_weight, c->d_ho, c->d_tab64, c->d_tab32, c->d_Ai32,
                    c->d_work, c->d_hop_n, c->d_hop_orb, c->d_hop_new, c->d_new_off, c->d_new_orb,
                    c->d_hopsign, c->d_chan, c->d_M, c->d_tabT, c->d_zscr, c->d_qorb, c->d_Q, c->d_Y,
                    c->d_flag, c->d_hbuf};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

static int upload_order(gpvmc_ctx *c)
{
    // Blocks are dealt round-robin to the 8 XCDs (block b -> XCD b % 8).  Walkers are
    // sorted by parameter group and XCD x receives a contiguous run of that order, so
    // walkers sharing an orbital table share an XCD's L2.
    const int W = c->W, X = 8;
    std::vector<int32_t> byg(W);
    for (int i = 0; i < W; ++i) byg[i] = i;
    std::stable_sort(byg.begin(), byg.end(),
                     [&](int a, int b) { return c->groups[a] < c->groups[b]; });
    std::vector<int32_t> order(W);
    int pos = 0;
    for (int x = 0; x < X; ++x)
        for (int b = x; b < W; b += X) order[b] = byg[pos++];
    HIPCHK(hipMemcpy(c->d_order, order.data(), sizeof(int32_t) * W, hipMemcpyHostToDevice));
    return GPVMC_OK;
}

static int ensure_scratch(gpvmc_ctx *c, size_t bytes)
{
    if (bytes <= c->scratch_bytes) return GPVMC_OK;
    if (c->d_scratch) HIPCHK(hipFree(c->d_scratch));
    c->d_scratch = nullptr;
    c->scratch_bytes = 0;
    HIPCHK(hipMalloc(&c->d_scratch, bytes));
    c->scratch_bytes = bytes;
    return GPVMC_OK;
}

static int ensure_statspin(gpvmc_ctx *c)
{
    if (c->d_ss) return GPVMC_OK;
    const size_t bytes = (size_t)c->W * gv::ss_block((size_t)c->Ns) * sizeof(double);
    HIPCHK(hipMalloc(&c->d_ss, bytes));
    HIPCHK(hipMemsetAsync(c->d_ss, 0, bytes, c->stream));
    return GPVMC_OK;
}

static int clear_statspin(gpvmc_ctx *c)
{
    if (c->d_ss)
        HIPCHK(hipMemsetAsync(c->d_ss, 0, (size_t)c->W * gv::ss_block((size_t)c->Ns) * sizeof(double),
                              c->stream));
    return GPVMC_OK;
}

static gv::ExcArgs exc_args(gpvmc_ctx *c, int mode)
{
    gv::ExcArgs a{};
    a.Ai = c->prec == GPVMC_FP64 ? (void *)c->d_Ai64 : (void *)c->d_Ai32;
    a.site = c->d_site;
    a.part = c->d_part;
    a.cur = c->d_cur;
    a.cstate = c->d_cstate;
    a.weight = c->d_weight;
    a.logabs = c->d_logabs;
    a.phase = c->d_phase;
    a.counter = c->d_counter;
    a.acc = c->d_acc;
    a.ho = c->d_ho;
    a.tab = c->prec == GPVMC_FP64 ? (const void *)c->d_tab64 : (const void *)c->d_tab32;
    a.tab64 = c->d_tab64;
    a.group = c->d_group;
    a.order = c->d_order;
    a.nbr = c->d_nbr;
    a.edge = c->d_edge;
    a.hop_n = c->d_hop_n;
    a.hop_orb = c->d_hop_orb;
    a.hop_new = c->d_hop_new;
    a.new_off = c->d_new_off;
    a.new_orb = c->d_new_orb;
    a.hopsign = c->d_hopsign;
    a.chan = c->d_chan;
    a.M = c->d_M;
    a.Q = c->d_Q;
    a.Y = c->d_Y;
    a.qorb = c->d_qorb;
    a.tabT = c->d_tabT;
    a.zscr = c->d_zscr;
    a.hbuf = c->d_hbuf;
    a.hb_cap = c->hb_cap;
    a.nq0 = c->nq0;
    a.nq1 = c->nq1;
    a.rb0 = c->rb0;
    a.rb1 = c->rb1;
    a.cb = c->rk_cb;
    a.key0 = (uint32_t)c->cfg.seed;
    a.key1 = (uint32_t)(c->cfg.seed >> 32);
    a.L = c->L;
    a.n_up = c->n_up;
    a.n_dn = c->n_dn;
    a.nk = c->nk;
    a.mode = mode;
    return a;
}

static hipError_t launch_exc(gpvmc_ctx *c, const gv::ExcArgs &a, hipStream_t st = nullptr)
{
    if (!st) st = c->stream;
    if (c->regk) return gv_launch_excr(c->prec, c->rk_rbw, &a, c->W, st);
    return gv_launch_exc(c->prec, &a, c->W, st);
}

// The register-resident channel kernel walks the segments between measured steps (steps
// 0, meas, 2 meas, ... are followed by a measurement) and measures in a launch of its own.
// Each measurement leaves a record (a_k, h_k, S) per walker; every hb_cap measurements, and
// at the end of the call, the fold kernel adds the records to the H and O sums in order.
static hipError_t launch_exc_walk(gpvmc_ctx *c, gv::ExcArgs a, int64_t n_steps, int64_t meas, hipStream_t st)
{
    a.n_total = n_steps;
    int64_t s0 = 0, slot = 0;
    int pend = 0;
    while (s0 < n_steps) {
        int64_t last = n_steps - 1;
        bool m_after = false;
        if (meas > 0) {
            const int64_t m = meas == 1 ? s0 : ((s0 + meas - 1) / meas) * meas;
            if (m < n_steps) {
                last = m;
                m_after = true;
            }
        }
        a.mode = 0;
        a.st0 = s0;
        a.n_steps = last - s0 + 1;
        hipError_t e = launch_exc(c, a, st);
        if (e != hipSuccess) return e;
        if (m_after) {
            a.mode = 5;
            a.meas_slot = slot++;
            a.hb_slot = pend;
            e = launch_exc(c, a, st);
            if (e != hipSuccess) return e;
            if (++pend == c->hb_cap) {
                e = gv_launch_excr_fold(&a, pend, c->W, st);
                if (e != hipSuccess) return e;
                pend = 0;
            }
        }
        s0 = last + 1;
    }
    if (pend > 0) return gv_launch_excr_fold(&a, pend, c->W, st);
    return hipSuccess;
}

// channel contexts: FP64 Gauss-Jordan of the current state's matrices, then the weight
// bookkeeping (mode 2: also the choice of the dominant state, mode 3: weight only)
// FP64 inversion of every walker's Slater matrices (both flavours), the role of
// SlaterDeterminant::Init / linalg::DetInv; mode 0 ground state, 1 channel
static gv::DIArgs detinv_args(gpvmc_ctx *c, int mode)
{
    gv::DIArgs g{};
    g.mode = mode;
    g.n_up = c->channel ? c->n_up : c->N;
    g.n_dn = c->channel ? c->n_dn : c->N;
    g.nstride = c->N;
    g.Ns = c->Ns;
    g.ld = gv_detinv_ld(c->N);
    g.part = c->d_part;
    g.cur = c->d_cur;
    g.group = c->d_group;
    g.occ = c->d_occ64;
    g.tab = c->d_tab64;
    g.Ai64 = c->d_Ai64;
    g.Ai32 = (c->channel && c->prec == GPVMC_MIXED && !c->regk) ? c->d_Ai32 : nullptr;
    g.logabs = c->d_logabs;
    g.phase = c->d_phase;
    g.status = c->d_status;
    g.work = c->channel ? c->d_work : c->d_gjwork;
    return g;
}

static int run_refresh_channel(gpvmc_ctx *c, int mode)
{
    gv::ExcArgs a = exc_args(c, mode);
    const gv::DIArgs g = detinv_args(c, 1);
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    HIPCHK(gv_launch_detinv(&g, 2 * c->W, c->stream));
    if (c->regk)
        HIPCHK(gv_launch_excr_mbuild(c->prec, &a, c->d_Ai64, c->W, c->stream));
    else
        HIPCHK(gv_launch_exc_qbuild(c->prec, &a, c->d_Ai64, c->W, c->stream));
    HIPCHK(launch_exc(c, a));
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->refresh_ms += ms;
    c->n_refresh += 1;
    return GPVMC_OK;
}

// FP64 Gauss-Jordan inverse of every walker's Slater matrices, then the ratio tables
// G = Ai Phi_other in FP64, stored in the working precision.
static int run_refresh(gpvmc_ctx *c, double *drift_dev)
{
    if (const size_t wb = gv_detinv_work_bytes(c->N)) {
        if (!c->d_gjwork) HIPCHK(hipMalloc(&c->d_gjwork, wb * 2 * (size_t)c->W));
    }
    const gv::DIArgs a = detinv_args(c, 0);
    gv::GBArgs b{};
    b.G = c->d_G;
    b.Ai = c->d_Ai64;
    b.part = c->d_part;
    b.occ = c->d_occ64;
    b.group = c->d_group;
    b.drift = drift_dev;
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    HIPCHK(gv_launch_detinv(&a, 2 * c->W, c->stream));
    HIPCHK(gv_launch_gbuild(c->prec, c->N, &b, c->W, c->stream));
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->refresh_ms += ms;
    c->n_refresh += 1;
    return GPVMC_OK;
}

// SlaterDeterminant::Init after the configurations (d_site, d_part) are on the device:
// cleared accumulators and counters, the FP64 inversion, the ratio tables (ground state) or
// the current-state choice (channels: every walker starts from state 0 of its channel,
// slots in increasing orbital order, and moves to its dominant state); st receives the
// status pair of every walker
static int init_uploaded(gpvmc_ctx *c, std::vector<int32_t> &st)
{
    const int W = c->W;
    int rc = GPVMC_OK;
    HIPCHK(hipMemsetAsync(c->d_acc, 0, (size_t)W * gv::ACC_STRIDE * sizeof(double), c->stream));
    HIPCHK(hipMemsetAsync(c->d_counter, 0, (size_t)W * sizeof(uint64_t), c->stream));
    if (c->channel) {
        const int nm = c->N, npt[2] = {c->n_up, c->n_dn};
        std::vector<int32_t> cur((size_t)W * 2 * nm, 0);
        for (int w = 0; w < W; ++w) {
            const int32_t *so = c->state_orb.data() + (size_t)c->chans[w] * c->nk * (c->n_up + c->n_dn);
            for (int fl = 0; fl < 2; ++fl)
                for (int f = 0; f < npt[fl]; ++f)
                    cur[(size_t)w * 2 * nm + (size_t)fl * nm + f] = so[(fl ? c->n_up : 0) + f];
        }
        HIPCHK(hipMemcpy(c->d_cur, cur.data(), cur.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemsetAsync(c->d_cstate, 0, (size_t)W * sizeof(int32_t), c->stream));
        HIPCHK(hipMemsetAsync(c->d_ho, 0, (size_t)W * 4 * c->nk * c->nk * sizeof(double), c->stream));
        rc = run_refresh_channel(c, 2);
    } else {
        rc = clear_statspin(c);
        if (!rc) rc = run_refresh(c, nullptr);
    }
    if (rc) return rc;
    st.assign((size_t)W * 2, 0);
    HIPCHK(hipMemcpy(st.data(), c->d_status, st.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    return GPVMC_OK;
}

static int init_channel(gpvmc_ctx *c, const int32_t *spins, int32_t *ok)
{
    if (!spins) return fail(GPVMC_ERR_ARG, "null spins");
    for (int w = 0; w < c->W; ++w)
        if (!c->have_orb[c->groups[w]])
            return fail(GPVMC_ERR_STATE, "orbital table of group " + std::to_string(c->groups[w]) + " not set");
    const int Ns = c->Ns, nm = c->N, W = c->W;
    const int npt[2] = {c->n_up, c->n_dn};
    std::vector<int32_t> site((size_t)W * Ns), part((size_t)W * 2 * nm, 0);
    for (int w = 0; w < W; ++w) {
        int cnt[2] = {0, 0};
        for (int v = 0; v < Ns; ++v) {
            const int fl = spins[(size_t)w * Ns + v];
            if (fl != 0 && fl != 1) return fail(GPVMC_ERR_ARG, "spins must be 0 or 1");
            if (cnt[fl] >= npt[fl])
                return fail(GPVMC_ERR_ARG, "walker " + std::to_string(w) + " has the wrong number of particles");
            site[(size_t)w * Ns + v] = (fl << 16) | cnt[fl];  // State::InitFock order
            part[(size_t)w * 2 * nm + (size_t)fl * nm + cnt[fl]] = v;
            cnt[fl]++;
        }
    }
    HIPCHK(hipSetDevice(c->cfg.device));
    HIPCHK(hipMemcpy(c->d_site, site.data(), site.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_part, part.data(), part.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    std::vector<int32_t> st;
    int rc = init_uploaded(c, st);
    if (rc) return rc;
    if (ok)
        for (int w = 0; w < W; ++w) ok[w] = (st[2 * w] == 0 && st[2 * w + 1] == 0) ? 1 : 0;
    c->initialised = true;
    return GPVMC_OK;
}

// the sweep kernel of the context's variant: delayed-update windows where this shape has
// them, else one accepted move per window
// the default variant per shape: delayed-update windows where they are measured faster
static bool windows_by_default(int prec, int N, int lpr)
{
    (void)prec;
    (void)N;
    (void)lpr;
    return false;
}

static bool uses_windows(const gpvmc_ctx *c)
{
    if (c->channel || !gv_sweep_wdu_supported(c->prec, c->N, c->lpr)) return false;
    return c->variant == GPVMC_SWEEP_WINDOWS ||
           (c->variant == GPVMC_SWEEP_DEFAULT && windows_by_default(c->prec, c->N, c->lpr));
}

static hipError_t launch_sweep(const gpvmc_ctx *c, const gv::SweepArgs &a, hipStream_t st)
{
    if (uses_windows(c)) return gv_launch_sweep_wdu(c->prec, c->N, c->lpr, &a, c->W, st);
    return gv_launch_sweep(c->prec, c->N, c->lpr, &a, c->W, st);
}

static int walk_channel(gpvmc_ctx *c, int64_t n_steps, int64_t meas, uint32_t mask, const double *uniforms,
                        int8_t *noop, int8_t *accept, double *ratio, double *eloc, double *mz, double *logw,
                        int measure_only)
{
    (void)mask;
    if (eloc || mz) return fail(GPVMC_ERR_UNSUPPORTED, "channel contexts accumulate ProjHeis H and O "
                                                      "(gpvmc_read_channel); no per-measurement E_loc or m_z");
    const size_t W = c->W;
    const int64_t nmeas = measure_only ? 1 : meas > 0 ? (meas == 1 ? n_steps : (n_steps + meas - 1) / meas) : 0;
    if (uniforms) {
        const size_t n = W * (size_t)n_steps * 4;
        for (size_t i = 0; i < n; ++i)
            if (!(uniforms[i] >= 0.0 && uniforms[i] < 1.0))
                return fail(GPVMC_ERR_ARG, "replay uniforms must lie in [0, 1)");
    }
    size_t off = 0, o_u = 0, o_no = 0, o_ac = 0, o_ra = 0, o_lw = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    if (uniforms) o_u = take(W * n_steps * 4 * sizeof(double));
    if (noop) o_no = take(W * n_steps);
    if (accept) o_ac = take(W * n_steps);
    if (ratio) o_ra = take(W * n_steps * sizeof(double));
    if (logw && nmeas) o_lw = take(W * nmeas * sizeof(double));
    HIPCHK(hipSetDevice(c->cfg.device));
    if (off) {
        int rc = ensure_scratch(c, off);
        if (rc) return rc;
    }
    char *sb = static_cast<char *>(c->d_scratch);
    gv::ExcArgs a = exc_args(c, measure_only ? 1 : 0);
    a.n_steps = measure_only ? 0 : n_steps;
    a.n_total = a.n_steps;
    a.meas = meas;
    a.n_meas_slots = nmeas;
    if (uniforms) {
        a.uniforms = reinterpret_cast<const double *>(sb + o_u);
        HIPCHK(hipMemcpyAsync(sb + o_u, uniforms, W * n_steps * 4 * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
    }
    if (noop) a.tr_noop = reinterpret_cast<int8_t *>(sb + o_no);
    if (accept) a.tr_accept = reinterpret_cast<int8_t *>(sb + o_ac);
    if (ratio) a.tr_ratio = reinterpret_cast<double *>(sb + o_ra);
    if (logw && nmeas) a.tr_logw = reinterpret_cast<double *>(sb + o_lw);
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    if (!c->regk || measure_only) {
        HIPCHK(launch_exc(c, a));
    } else {
        HIPCHK(launch_exc_walk(c, a, n_steps, meas, c->stream));
    }
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (!measure_only) {
        c->sweep_ms += ms;
        c->n_sweep += 1;
    }
    if (noop) HIPCHK(hipMemcpy(noop, a.tr_noop, W * n_steps, hipMemcpyDeviceToHost));
    if (accept) HIPCHK(hipMemcpy(accept, a.tr_accept, W * n_steps, hipMemcpyDeviceToHost));
    if (ratio) HIPCHK(hipMemcpy(ratio, a.tr_ratio, W * n_steps * sizeof(double), hipMemcpyDeviceToHost));
    if (a.tr_logw) HIPCHK(hipMemcpy(logw, a.tr_logw, W * nmeas * sizeof(double), hipMemcpyDeviceToHost));
    return GPVMC_OK;
}

extern "C" {

int gv_fail(int code, const char *msg) { return fail(code, msg ? msg : ""); }

const char *gpvmc_last_error(void) { return g_err.c_str(); }

int gpvmc_version(void) { return 1; }

int gpvmc_create(const gpvmc_config *cfg, gpvmc_ctx **out)
{
    if (!cfg || !out) return fail(GPVMC_ERR_ARG, "null argument");
    *out = nullptr;
    const int L = cfg->L;
    if (L < 4 || (L * L) % 2) return fail(GPVMC_ERR_ARG, "L must be >= 4 with L*L even");
    if (cfg->n_walkers < 1) return fail(GPVMC_ERR_ARG, "n_walkers must be >= 1");
    if (cfg->n_groups < 1) return fail(GPVMC_ERR_ARG, "n_groups must be >= 1");
    if (cfg->precision != GPVMC_FP64 && cfg->precision != GPVMC_MIXED)
        return fail(GPVMC_ERR_ARG, "precision must be GPVMC_FP64 or GPVMC_MIXED");
    const int N = L * L / 2;
    int lpr = cfg->lanes_per_row ? cfg->lanes_per_row : default_lpr(cfg->precision, N);
    int res = 0, nt = 0;
    if (!lpr || gv_sweep_supported(cfg->precision, N, lpr, &res, &nt) != lpr)
        return fail(GPVMC_ERR_UNSUPPORTED, "no sweep kernel for N=" + std::to_string(N) +
                                               " lanes_per_row=" + std::to_string(lpr));
    if (!gv_refresh_supported(cfg->precision, N))
        return fail(GPVMC_ERR_UNSUPPORTED, "no LU kernel for N=" + std::to_string(N));

    gpvmc_ctx *c = new gpvmc_ctx();
    c->cfg = *cfg;
    c->L = L;
    c->Ns = L * L;
    c->N = N;
    c->W = cfg->n_walkers;
    c->G = cfg->n_groups;
    c->prec = cfg->precision;
    c->lpr = lpr;
    c->res = res;
    c->nthreads = nt;
    c->esize = c->prec == GPVMC_FP64 ? sizeof(double2) : sizeof(float2);
    c->have_orb.assign(c->G, 0);
    c->groups.assign(c->W, 0);
    auto bail = [&](int rc) {
        free_ctx(c);
        return rc;
    };
#define CK(expr)                                                                           \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return bail(fail(GPVMC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_))); \
    } while (0)
    CK(hipSetDevice(cfg->device));
    CK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    CK(hipEventCreate(&c->ev0));
    CK(hipEventCreate(&c->ev1));
    const size_t W = c->W, Ns = c->Ns;
    CK(hipMalloc(&c->d_G, W * 2 * N * N * c->esize));
    CK(hipMalloc(&c->d_Ai64, W * 2 * N * N * sizeof(double2)));
    CK(hipMalloc(&c->d_site, W * Ns * sizeof(int32_t)));
    CK(hipMalloc(&c->d_part, W * 2 * N * sizeof(int32_t)));
    CK(hipMalloc(&c->d_group, W * sizeof(int32_t)));
    CK(hipMalloc(&c->d_order, W * sizeof(int32_t)));
    CK(hipMalloc(&c->d_status, W * 2 * sizeof(int32_t)));
    CK(hipMalloc(&c->d_logabs, W * 2 * sizeof(double)));
    CK(hipMalloc(&c->d_phase, W * 2 * sizeof(double)));
    CK(hipMalloc(&c->d_drift, W * 2 * sizeof(double)));
    CK(hipMalloc(&c->d_acc, W * gv::ACC_STRIDE * sizeof(double)));
    CK(hipMalloc(&c->d_counter, W * sizeof(uint64_t)));
    CK(hipMalloc(&c->d_occ64, (size_t)c->G * 2 * Ns * N * sizeof(double2)));
    if (c->prec == GPVMC_MIXED) CK(hipMalloc(&c->d_occ32, (size_t)c->G * 2 * Ns * N * sizeof(float2)));
    CK(hipMemset(c->d_group, 0, W * sizeof(int32_t)));
    CK(hipMemset(c->d_acc, 0, W * gv::ACC_STRIDE * sizeof(double)));
    CK(hipMemset(c->d_counter, 0, W * sizeof(uint64_t)));
    std::vector<int32_t> nbr, edge4;
    build_lattice(L, nbr, edge4);
    CK(hipMalloc(&c->d_nbr, nbr.size() * sizeof(int32_t)));
    CK(hipMalloc(&c->d_edge, edge4.size() * sizeof(int32_t)));
    CK(hipMemcpy(c->d_nbr, nbr.data(), nbr.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    CK(hipMemcpy(c->d_edge, edge4.data(), edge4.size() * sizeof(int32_t), hipMemcpyHostToDevice));
#undef CK
    int rc = upload_order(c);
    if (rc) return bail(rc);
    *out = c;
    return GPVMC_OK;
}

int gpvmc_destroy(gpvmc_ctx *ctx)
{
    if (!ctx) return fail(GPVMC_ERR_ARG, "null context");
    (void)hipSetDevice(ctx->cfg.device);
    (void)hipStreamSynchronize(ctx->stream);
    free_ctx(ctx);
    return GPVMC_OK;
}

int gpvmc_set_orbitals(gpvmc_ctx *c, int32_t group, const double *occ)
{
    if (!c || !occ) return fail(GPVMC_ERR_ARG, "null argument");
    if (c->channel) return fail(GPVMC_ERR_STATE, "channel contexts take the full table (gpvmc_set_orbital_table)");
    if (group < 0 || group >= c->G) return fail(GPVMC_ERR_ARG, "group out of range");
    const size_t n = (size_t)2 * c->Ns * c->N;
    for (size_t i = 0; i < 2 * n; ++i)
        if (!std::isfinite(occ[i])) return fail(GPVMC_ERR_ARG, "non-finite orbital value");
    HIPCHK(hipSetDevice(c->cfg.device));
    HIPCHK(hipMemcpy(c->d_occ64 + (size_t)group * n, occ, n * sizeof(double2), hipMemcpyHostToDevice));
    if (c->prec == GPVMC_MIXED) {
        std::vector<float2> f(n);
        for (size_t i = 0; i < n; ++i) f[i] = make_float2((float)occ[2 * i], (float)occ[2 * i + 1]);
        HIPCHK(hipMemcpy(c->d_occ32 + (size_t)group * n, f.data(), n * sizeof(float2), hipMemcpyHostToDevice));
    }
    c->have_orb[group] = 1;
    return GPVMC_OK;
}

int gpvmc_set_groups(gpvmc_ctx *c, const int32_t *g)
{
    if (!c || !g) return fail(GPVMC_ERR_ARG, "null argument");
    for (int w = 0; w < c->W; ++w)
        if (g[w] < 0 || g[w] >= c->G) return fail(GPVMC_ERR_ARG, "group index out of range");
    HIPCHK(hipSetDevice(c->cfg.device));
    c->groups.assign(g, g + c->W);
    HIPCHK(hipMemcpy(c->d_group, g, sizeof(int32_t) * c->W, hipMemcpyHostToDevice));
    return upload_order(c);
}

int gpvmc_init_walkers(gpvmc_ctx *c, const int32_t *spins, int32_t *ok)
{
    if (c && c->channel) return init_channel(c, spins, ok);
    if (!c || !spins) return fail(GPVMC_ERR_ARG, "null argument");
    for (int w = 0; w < c->W; ++w)
        if (!c->have_orb[c->groups[w]])
            return fail(GPVMC_ERR_STATE, "orbitals of group " + std::to_string(c->groups[w]) + " not set");
    const int Ns = c->Ns, N = c->N, W = c->W;
    std::vector<int32_t> site((size_t)W * Ns), part((size_t)W * 2 * N);
    for (int w = 0; w < W; ++w) {
        int cnt[2] = {0, 0};
        for (int v = 0; v < Ns; ++v) {
            int fl = spins[(size_t)w * Ns + v];
            if (fl != 0 && fl != 1) return fail(GPVMC_ERR_ARG, "spins must be 0 or 1");
            if (cnt[fl] >= N) return fail(GPVMC_ERR_ARG, "walker " + std::to_string(w) + " is not at Sz=0");
            // State::InitFock (State.cpp:29-41): labels in increasing site order
            site[(size_t)w * Ns + v] = (fl << 16) | cnt[fl];
            part[(size_t)w * 2 * N + (size_t)fl * N + cnt[fl]] = v;
            cnt[fl]++;
        }
    }
    HIPCHK(hipSetDevice(c->cfg.device));
    HIPCHK(hipMemcpy(c->d_site, site.data(), site.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_part, part.data(), part.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    std::vector<int32_t> st;
    const int rc = init_uploaded(c, st);
    if (rc) return rc;
    if (ok)
        for (int w = 0; w < W; ++w) ok[w] = (st[2 * w] == 0 && st[2 * w + 1] == 0) ? 1 : 0;
    c->initialised = true;
    return GPVMC_OK;
}

int gpvmc_init_random(gpvmc_ctx *c, uint64_t seed, int32_t max_attempts, int32_t *ok, int32_t *attempts)
{
    if (!c) return fail(GPVMC_ERR_ARG, "null context");
    if (max_attempts < 1 || max_attempts > 1000000) return fail(GPVMC_ERR_ARG, "max_attempts must be in [1, 1e6]");
    for (int w = 0; w < c->W; ++w)
        if (!c->have_orb[c->groups[w]])
            return fail(GPVMC_ERR_STATE, "orbitals of group " + std::to_string(c->groups[w]) + " not set");
    const int W = c->W;
    HIPCHK(hipSetDevice(c->cfg.device));
    if (!c->d_flag) HIPCHK(hipMalloc(&c->d_flag, sizeof(int32_t)));
    gv::RIArgs a{};
    a.W = W;
    a.Ns = c->Ns;
    a.nstride = c->N;
    a.n_up = c->channel ? c->n_up : c->N;
    a.n_dn = c->channel ? c->n_dn : c->N;
    a.key0 = (uint32_t)seed;
    a.key1 = (uint32_t)(seed >> 32);
    a.site = c->d_site;
    a.part = c->d_part;
    a.incomplete = c->d_flag;
    std::vector<int32_t> st, used(W, 1);
    for (int t = 0; t < max_attempts; ++t) {
        // attempt 0 draws every walker, later ones only the walkers still singular
        a.attempt = t;
        a.status = t ? c->d_status : nullptr;
        HIPCHK(hipMemsetAsync(c->d_flag, 0, sizeof(int32_t), c->stream));
        HIPCHK(gv_launch_raninit(&a, c->stream));
        int32_t incomplete = 0;
        HIPCHK(hipMemcpyAsync(&incomplete, c->d_flag, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (incomplete) return fail(GPVMC_ERR_STATE, "a RanInit draw did not complete within 2^20 uniforms");
        const int rc = init_uploaded(c, st);
        if (rc) return rc;
        bool all = true;
        for (int w = 0; w < W; ++w)
            if (st[2 * w] || st[2 * w + 1]) {
                all = false;
                if (t + 1 < max_attempts) used[w] = t + 2;
            }
        if (all) break;
    }
    for (int w = 0; w < W; ++w) {
        if (ok) ok[w] = (st[2 * w] == 0 && st[2 * w + 1] == 0) ? 1 : 0;
        if (attempts) attempts[w] = used[w];
    }
    c->initialised = true;
    return GPVMC_OK;
}

int gpvmc_sweep_shapes(int32_t precision, int32_t L, int32_t *lanes_per_row, int32_t max)
{
    if (L < 4 || (L * L) % 2) return fail(GPVMC_ERR_ARG, "L must be >= 4 with L*L even");
    if (precision != GPVMC_FP64 && precision != GPVMC_MIXED)
        return fail(GPVMC_ERR_ARG, "precision must be GPVMC_FP64 or GPVMC_MIXED");
    if (max < 0 || (max > 0 && !lanes_per_row)) return fail(GPVMC_ERR_ARG, "null output");
    int tmp[32];
    const int n = gv_sweep_shapes(precision, L * L / 2, tmp, 32);
    for (int i = 0; i < n && i < max && i < 32; ++i) lanes_per_row[i] = tmp[i];
    return n;
}

int gpvmc_set_sweep_variant(gpvmc_ctx *c, int32_t variant)
{
    if (!c) return fail(GPVMC_ERR_ARG, "null context");
    if (c->channel) return fail(GPVMC_ERR_UNSUPPORTED, "channel contexts have one sweep kernel per shape");
    if (variant != GPVMC_SWEEP_DEFAULT && variant != GPVMC_SWEEP_ONE_MOVE && variant != GPVMC_SWEEP_WINDOWS)
        return fail(GPVMC_ERR_ARG, "variant must be GPVMC_SWEEP_DEFAULT, _ONE_MOVE or _WINDOWS");
    if (variant == GPVMC_SWEEP_WINDOWS && !gv_sweep_wdu_supported(c->prec, c->N, c->lpr))
        return fail(GPVMC_ERR_UNSUPPORTED, "no delayed-update window kernel for this precision and N");
    c->variant = variant;
    return GPVMC_OK;
}

int gpvmc_sweep_variant(const gpvmc_ctx *c)
{
    if (!c) return -1;
    return uses_windows(c) ? GPVMC_SWEEP_WINDOWS : GPVMC_SWEEP_ONE_MOVE;
}

int gpvmc_refresh(gpvmc_ctx *c, double *drift)
{
    if (!c) return fail(GPVMC_ERR_ARG, "null context");
    if (!c->initialised) return fail(GPVMC_ERR_STATE, "walkers not initialised");
    HIPCHK(hipSetDevice(c->cfg.device));
    if (c->channel) {
        if (drift)
            for (int w = 0; w < c->W; ++w) drift[w] = 0.0;
        return run_refresh_channel(c, 3);
    }
    int rc = run_refresh(c, drift ? c->d_drift : nullptr);
    if (rc) return rc;
    if (drift) {
        std::vector<double> d((size_t)c->W * 2);
        HIPCHK(hipMemcpy(d.data(), c->d_drift, d.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int w = 0; w < c->W; ++w) drift[w] = std::max(d[2 * w], d[2 * w + 1]);
    }
    return GPVMC_OK;
}

static gv::SweepArgs sweep_args(gpvmc_ctx *c, int64_t n_steps, int64_t meas, uint32_t mask, int64_t nmeas,
                                int measure_only)
{
    gv::SweepArgs a{};
    a.G = c->d_G;
    a.site = c->d_site;
    a.part = c->d_part;
    a.logabs = c->d_logabs;
    a.phase = c->d_phase;
    a.counter = c->d_counter;
    a.acc = c->d_acc;
    a.occ = c->prec == GPVMC_FP64 ? (const void *)c->d_occ64 : (const void *)c->d_occ32;
    a.group = c->d_group;
    a.order = c->d_order;
    a.nbr = c->d_nbr;
    a.edge = c->d_edge;
    a.n_steps = n_steps;
    a.meas = meas;
    a.n_meas_slots = nmeas;
    a.mask = mask;
    a.key0 = (uint32_t)c->cfg.seed;
    a.key1 = (uint32_t)(c->cfg.seed >> 32);
    a.L = c->L;
    a.measure_only = measure_only;
    a.ss = c->d_ss;
    return a;
}

// steps of one ground-state sweep launch at most (the kernel counts them in 32 bits);
// longer untraced walks are split into launches of whole measurement intervals
static constexpr int64_t kMaxLaunchSteps = int64_t(1) << 30;

static int walk_impl(gpvmc_ctx *c, int64_t n_steps, int64_t meas, uint32_t mask,
                     const double *uniforms, int8_t *noop, int8_t *accept, double *ratio,
                     double *eloc, double *mz, double *logw, int measure_only)
{
    if (!c) return fail(GPVMC_ERR_ARG, "null context");
    if (!c->initialised) return fail(GPVMC_ERR_STATE, "walkers not initialised");
    if (n_steps < 0 || meas < 0) return fail(GPVMC_ERR_ARG, "negative n_steps or meas");
    if (mask & ~(uint32_t)(GPVMC_MEAS_PROJHEIS | GPVMC_MEAS_STAGMAGNZ | GPVMC_MEAS_OVERLAP |
                           GPVMC_MEAS_STATSPIN))
        return fail(GPVMC_ERR_ARG, "unknown estimator bit in mask");
    if (n_steps == 0 && !measure_only) return GPVMC_OK;
    if (c->channel) return walk_channel(c, n_steps, meas, mask, uniforms, noop, accept, ratio, eloc, mz, logw,
                                        measure_only);
    if (meas > kMaxLaunchSteps) return fail(GPVMC_ERR_ARG, "meas above 2^30");
    if (n_steps > kMaxLaunchSteps) {
        if (uniforms || noop || accept || ratio || eloc || mz || logw)
            return fail(GPVMC_ERR_ARG, "a replayed or traced walk takes at most 2^30 steps per call");
        const int64_t chunk = meas > 0 ? (kMaxLaunchSteps / meas) * meas : kMaxLaunchSteps;
        for (int64_t done = 0; done < n_steps; done += chunk) {
            const int64_t n = n_steps - done < chunk ? n_steps - done : chunk;
            int rc = walk_impl(c, n, meas, mask, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0);
            if (rc) return rc;
        }
        return GPVMC_OK;
    }
    const size_t W = c->W;
    const int64_t nmeas = measure_only ? 1
                          : meas > 0 ? (meas == 1 ? n_steps : (n_steps + meas - 1) / meas) : 0;
    if (uniforms) {
        const size_t n = W * (size_t)n_steps * 4;
        for (size_t i = 0; i < n; ++i)
            if (!(uniforms[i] >= 0.0 && uniforms[i] < 1.0))
                return fail(GPVMC_ERR_ARG, "replay uniforms must lie in [0, 1)");
    }
    // scratch layout: uniforms | noop | accept | ratio | eloc | mz | logw
    size_t off = 0, o_u = 0, o_no = 0, o_ac = 0, o_ra = 0, o_el = 0, o_mz = 0, o_lw = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    if (uniforms) o_u = take(W * n_steps * 4 * sizeof(double));
    if (noop) o_no = take(W * n_steps);
    if (accept) o_ac = take(W * n_steps);
    if (ratio) o_ra = take(W * n_steps * sizeof(double));
    if (eloc && nmeas) o_el = take(W * nmeas * 2 * sizeof(double));
    if (mz && nmeas) o_mz = take(W * nmeas * 2 * sizeof(double));
    if (logw && nmeas) o_lw = take(W * nmeas * sizeof(double));
    HIPCHK(hipSetDevice(c->cfg.device));
    if (off) {
        int rc = ensure_scratch(c, off);
        if (rc) return rc;
    }
    if ((mask & GPVMC_MEAS_STATSPIN) && !measure_only && nmeas) {
        int rc = ensure_statspin(c);
        if (rc) return rc;
    }
    char *sb = static_cast<char *>(c->d_scratch);
    gv::SweepArgs a = sweep_args(c, n_steps, meas, mask, nmeas, measure_only);
    if (uniforms) {
        a.uniforms = reinterpret_cast<const double *>(sb + o_u);
        HIPCHK(hipMemcpyAsync(sb + o_u, uniforms, W * n_steps * 4 * sizeof(double),
                              hipMemcpyHostToDevice, c->stream));
    }
    if (noop) a.tr_noop = reinterpret_cast<int8_t *>(sb + o_no);
    if (accept) a.tr_accept = reinterpret_cast<int8_t *>(sb + o_ac);
    if (ratio) a.tr_ratio = reinterpret_cast<double *>(sb + o_ra);
    if (eloc && nmeas) a.tr_eloc = reinterpret_cast<double *>(sb + o_el);
    if (mz && nmeas) a.tr_mz = reinterpret_cast<double *>(sb + o_mz);
    if (logw && nmeas) a.tr_logw = reinterpret_cast<double *>(sb + o_lw);
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    HIPCHK(launch_sweep(c, a, c->stream));
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (!measure_only) {
        c->sweep_ms += ms;
        c->n_sweep += 1;
    }
    if (noop) HIPCHK(hipMemcpy(noop, a.tr_noop, W * n_steps, hipMemcpyDeviceToHost));
    if (accept) HIPCHK(hipMemcpy(accept, a.tr_accept, W * n_steps, hipMemcpyDeviceToHost));
    if (ratio) HIPCHK(hipMemcpy(ratio, a.tr_ratio, W * n_steps * sizeof(double), hipMemcpyDeviceToHost));
    if (a.tr_eloc) HIPCHK(hipMemcpy(eloc, a.tr_eloc, W * nmeas * 2 * sizeof(double), hipMemcpyDeviceToHost));
    if (a.tr_mz) HIPCHK(hipMemcpy(mz, a.tr_mz, W * nmeas * 2 * sizeof(double), hipMemcpyDeviceToHost));
    if (a.tr_logw) HIPCHK(hipMemcpy(logw, a.tr_logw, W * nmeas * sizeof(double), hipMemcpyDeviceToHost));
    return GPVMC_OK;
}

int gpvmc_walk(gpvmc_ctx *c, int64_t n_steps, int64_t meas, uint32_t mask, const double *uniforms,
               int8_t *noop, int8_t *accept, double *ratio, double *eloc, double *mz, double *logw)
{
    return walk_impl(c, n_steps, meas, mask, uniforms, noop, accept, ratio, eloc, mz, logw, 0);
}

int gpvmc_measure(gpvmc_ctx *c, uint32_t mask, double *eloc, double *mz, double *logw)
{
    return walk_impl(c, 0, 0, mask, nullptr, nullptr, nullptr, nullptr, eloc, mz, logw, 1);
}

int gpvmc_read_accumulators(gpvmc_ctx *c, double *acc)
{
    if (!c || !acc) return fail(GPVMC_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(c->cfg.device));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(acc, c->d_acc, (size_t)c->W * gv::ACC_STRIDE * sizeof(double), hipMemcpyDeviceToHost));
    return GPVMC_OK;
}

int gpvmc_reset_accumulators(gpvmc_ctx *c)
{
    if (!c) return fail(GPVMC_ERR_ARG, "null context");
    HIPCHK(hipSetDevice(c->cfg.device));
    HIPCHK(hipMemsetAsync(c->d_acc, 0, (size_t)c->W * gv::ACC_STRIDE * sizeof(double), c->stream));
    if (c->d_ho)
        HIPCHK(hipMemsetAsync(c->d_ho, 0, (size_t)c->W * 4 * c->nk * c->nk * sizeof(double), c->stream));
    int rc = clear_statspin(c);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return GPVMC_OK;
}

// Expands one raw StatSpinStruct block (szz | smp | spm diagonal | count) into the
// reference's five Ns x Ns complex matrices (szz, spm, smp, spp, smm), StatSpinStruct.cpp:77-81.
static void expand_statspin(const double *raw, size_t Ns, double *out)
{
    const size_t S = Ns * Ns;
    const double *zz = raw, *mp = raw + S, *pd = raw + 3 * S;
    double *szz = out, *spm = out + 2 * S, *smp = out + 4 * S, *spp = out + 6 * S, *smm = out + 8 * S;
    for (size_t e = 0; e < S; ++e) {
        szz[2 * e] = zz[e];
        szz[2 * e + 1] = 0.0;
        smp[2 * e] = mp[2 * e];
        smp[2 * e + 1] = mp[2 * e + 1];
        spp[2 * e] = spp[2 * e + 1] = 0.0;
        smm[2 * e] = smm[2 * e + 1] = 0.0;
    }
    for (size_t x = 0; x < Ns; ++x)
        for (size_t y = 0; y < Ns; ++y) {
            const size_t e = x * Ns + y;
            if (x == y) {
                spm[2 * e] = pd[x];
                spm[2 * e + 1] = 0.0;
            } else {
                spm[2 * e] = mp[2 * (y * Ns + x)];
                spm[2 * e + 1] = mp[2 * (y * Ns + x) + 1];
            }
        }
}

int gpvmc_read_statspin(gpvmc_ctx *c, int32_t w0, int32_t nw, double *acc, double *n_meas)
{
    if (!c) return fail(GPVMC_ERR_ARG, "null context");
    if (w0 < 0 || nw < 0 || (int64_t)w0 + nw > c->W) return fail(GPVMC_ERR_ARG, "walker range out of bounds");
    const size_t Ns = c->Ns, per = gv::ss_block(Ns), S = Ns * Ns;
    HIPCHK(hipSetDevice(c->cfg.device));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<double> raw(per, 0.0);
    for (int32_t k = 0; k < nw; ++k) {
        if (c->d_ss)
            HIPCHK(hipMemcpy(raw.data(), c->d_ss + (size_t)(w0 + k) * per, per * sizeof(double),
                             hipMemcpyDeviceToHost));
        if (acc) expand_statspin(raw.data(), Ns, acc + (size_t)k * 10 * S);
        if (n_meas) n_meas[k] = raw[per - 1];
    }
    return GPVMC_OK;
}

int gpvmc_read_statspin_groups(gpvmc_ctx *c, double *acc, double *n_meas)
{
    if (!c) return fail(GPVMC_ERR_ARG, "null context");
    const size_t Ns = c->Ns, per = gv::ss_block(Ns), S = Ns * Ns;
    const int G = c->G;
    HIPCHK(hipSetDevice(c->cfg.device));
    std::vector<double> raw((size_t)G * per, 0.0);
    if (c->d_ss) {
        // walkers listed group by group
        std::vector<int32_t> idx, off(G + 1, 0);
        for (int g = 0; g < G; ++g) {
            off[g] = (int32_t)idx.size();
            for (int w = 0; w < c->W; ++w)
                if (c->groups[w] == g) idx.push_back(w);
        }
        off[G] = (int32_t)idx.size();
        const size_t ib = (idx.size() + off.size()) * sizeof(int32_t);
        const size_t ob = ((ib + 255) & ~(size_t)255);
        int rc = ensure_scratch(c, ob + raw.size() * sizeof(double));
        if (rc) return rc;
        char *sb = static_cast<char *>(c->d_scratch);
        int32_t *d_idx = reinterpret_cast<int32_t *>(sb);
        int32_t *d_off = d_idx + idx.size();
        double *d_out = reinterpret_cast<double *>(sb + ob);
        if (!idx.empty())
            HIPCHK(hipMemcpyAsync(d_idx, idx.data(), idx.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(d_off, off.data(), off.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
        HIPCHK(gv_launch_ss_reduce(c->d_ss, d_idx, d_off, G, per, d_out, c->stream));
        HIPCHK(hipMemcpyAsync(raw.data(), d_out, raw.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    for (int g = 0; g < G; ++g) {
        if (acc) expand_statspin(raw.data() + (size_t)g * per, Ns, acc + (size_t)g * 10 * S);
        if (n_meas) n_meas[g] = raw[(size_t)g * per + per - 1];
    }
    return GPVMC_OK;
}

int gpvmc_read_state(gpvmc_ctx *c, int32_t *spins, int32_t *part, double *logabs, double *phase)
{
    if (!c) return fail(GPVMC_ERR_ARG, "null context");
    HIPCHK(hipSetDevice(c->cfg.device));
    HIPCHK(hipStreamSynchronize(c->stream));
    const size_t W = c->W, Ns = c->Ns, N = c->N;
    if (spins) {
        std::vector<int32_t> site(W * Ns);
        HIPCHK(hipMemcpy(site.data(), c->d_site, site.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < site.size(); ++i) spins[i] = site[i] >> 16;
    }
    if (part) HIPCHK(hipMemcpy(part, c->d_part, W * 2 * N * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<double> la(W * 2), ph(W * 2);
    HIPCHK(hipMemcpy(la.data(), c->d_logabs, la.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ph.data(), c->d_phase, ph.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t w = 0; w < W; ++w) {
        if (logabs) logabs[w] = la[2 * w] + la[2 * w + 1];
        if (phase) phase[w] = std::remainder(ph[2 * w] + ph[2 * w + 1], 2 * M_PI);
    }
    return GPVMC_OK;
}

int gpvmc_read_inverse(gpvmc_ctx *c, int32_t walker, double *ai)
{
    if (!c || !ai) return fail(GPVMC_ERR_ARG, "null argument");
    if (walker < 0 || walker >= c->W) return fail(GPVMC_ERR_ARG, "walker out of range");
    HIPCHK(hipSetDevice(c->cfg.device));
    HIPCHK(hipStreamSynchronize(c->stream));
    const size_t n = (size_t)2 * c->N * c->N;
    HIPCHK(hipMemcpy(ai, c->d_Ai64 + (size_t)walker * n, n * sizeof(double2), hipMemcpyDeviceToHost));
    return GPVMC_OK;
}

int gpvmc_read_ratio_table(gpvmc_ctx *c, int32_t walker, double *g)
{
    if (!c || !g) return fail(GPVMC_ERR_ARG, "null argument");
    if (c->channel) return fail(GPVMC_ERR_UNSUPPORTED, "channel contexts keep inverses, not ratio tables");
    if (walker < 0 || walker >= c->W) return fail(GPVMC_ERR_ARG, "walker out of range");
    HIPCHK(hipSetDevice(c->cfg.device));
    HIPCHK(hipStreamSynchronize(c->stream));
    const size_t n = (size_t)2 * c->N * c->N;
    const char *src = static_cast<const char *>(c->d_G) + (size_t)walker * n * c->esize;
    if (c->prec == GPVMC_FP64) {
        HIPCHK(hipMemcpy(g, src, n * sizeof(double2), hipMemcpyDeviceToHost));
    } else {
        std::vector<float2> f(n);
        HIPCHK(hipMemcpy(f.data(), src, n * sizeof(float2), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) {
            g[2 * i] = f[i].x;
            g[2 * i + 1] = f[i].y;
        }
    }
    return GPVMC_OK;
}

int gpvmc_kernel_time(gpvmc_ctx *c, double *sweep_ms, int64_t *sweep_launches, double *refresh_ms,
                      int64_t *refresh_launches)
{
    if (!c) return fail(GPVMC_ERR_ARG, "null context");
    if (sweep_ms) *sweep_ms = c->sweep_ms;
    if (sweep_launches) *sweep_launches = c->n_sweep;
    if (refresh_ms) *refresh_ms = c->refresh_ms;
    if (refresh_launches) *refresh_launches = c->n_refresh;
    return GPVMC_OK;
}

int gpvmc_reset_kernel_time(gpvmc_ctx *c)
{
    if (!c) return fail(GPVMC_ERR_ARG, "null context");
    c->sweep_ms = c->refresh_ms = 0;
    c->n_sweep = c->n_refresh = 0;
    return GPVMC_OK;
}

int gpvmc_create_channel(const gpvmc_config *cfg, const gpvmc_channel *ch, gpvmc_ctx **out)
{
    return gpvmc_create_channels(cfg, 1, ch, out);
}

int gpvmc_create_channels(const gpvmc_config *cfg, int32_t n_channels, const gpvmc_channel *chs, gpvmc_ctx **out)
{
    if (!cfg || !chs || !out) return fail(GPVMC_ERR_ARG, "null argument");
    *out = nullptr;
    if (n_channels < 1) return fail(GPVMC_ERR_ARG, "n_channels must be >= 1");
    const gpvmc_channel *ch = chs;
    const int L = cfg->L;
    if (L < 4 || (L * L) % 2) return fail(GPVMC_ERR_ARG, "L must be >= 4 with L*L even");
    if (cfg->n_walkers < 1) return fail(GPVMC_ERR_ARG, "n_walkers must be >= 1");
    if (cfg->n_groups < 1) return fail(GPVMC_ERR_ARG, "n_groups must be >= 1");
    if (cfg->precision != GPVMC_FP64 && cfg->precision != GPVMC_MIXED)
        return fail(GPVMC_ERR_ARG, "precision must be GPVMC_FP64 or GPVMC_MIXED");
    const int Ns = L * L, nk = ch->nk, nu = ch->n_up, nd = ch->n_dn;
    if (nu < 1 || nd < 1 || nu + nd != Ns) return fail(GPVMC_ERR_ARG, "n_up + n_dn must equal L*L");
    if (nk < 1) return fail(GPVMC_ERR_ARG, "nk must be >= 1");
    const int nm = std::max(nu, nd);
    if (nm > 256) return fail(GPVMC_ERR_UNSUPPORTED, "at most 256 particles per flavour");
    if (cfg->lanes_per_row != 0 && cfg->lanes_per_row != -1)
        return fail(GPVMC_ERR_ARG, "channel contexts take lanes_per_row 0 (register-resident kernel where the "
                                   "walker fits) or -1 (streamed kernel)");
    const size_t nch = (size_t)n_channels, kk = (size_t)nk * nk;
    std::vector<int32_t> hop_orb(nch * kk * 8), hop_n(nch * kk * 2), hopsign(nch * kk), state_orb;
    // new-orbital lists per (channel, current state, flavour) and the index of every hop's target
    std::vector<int32_t> hop_new(nch * kk * 4, 0), new_off(nch * (2 * nk + 1), 0), new_orb;
    for (size_t q = 0; q < nch; ++q) {
      ch = chs + q;
      if (ch->nk != nk || ch->n_up != nu || ch->n_dn != nd)
          return fail(GPVMC_ERR_ARG, "all channels of a context must have the same n_up, n_dn and nk");
      if (!ch->state_orb || !ch->hop_n || !ch->hop_orb || !ch->hopsign) return fail(GPVMC_ERR_ARG, "null channel array");
      for (int i = 0; i < nk * (nu + nd); ++i)
          if (ch->state_orb[i] < 0 || ch->state_orb[i] >= Ns) return fail(GPVMC_ERR_ARG, "state orbital out of range");
      state_orb.insert(state_orb.end(), ch->state_orb, ch->state_orb + (size_t)nk * (nu + nd));
      std::copy(ch->hop_orb, ch->hop_orb + kk * 8, hop_orb.begin() + q * kk * 8);
      std::copy(ch->hop_n, ch->hop_n + kk * 2, hop_n.begin() + q * kk * 2);
      std::copy(ch->hopsign, ch->hopsign + kk, hopsign.begin() + q * kk);
      for (int cc = 0; cc < nk; ++cc)
        for (int fl = 0; fl < 2; ++fl) {
            std::vector<int32_t> lst;
            for (int k = 0; k < nk; ++k) {
                const int hn = ch->hop_n[((size_t)cc * nk + k) * 2 + fl];
                if (hn < 0 || hn > 2) return fail(GPVMC_ERR_ARG, "at most two orbital hops per flavour");
                for (int h = 0; h < hn; ++h) {
                    const int32_t *ho = ch->hop_orb + ((((size_t)cc * nk + k) * 2 + fl) * 2 + h) * 2;
                    if (ho[0] < 0 || ho[0] >= Ns || ho[1] < 0 || ho[1] >= Ns)
                        return fail(GPVMC_ERR_ARG, "hop orbital out of range");
                    if (std::find(lst.begin(), lst.end(), ho[1]) == lst.end()) lst.push_back(ho[1]);
                }
            }
            std::sort(lst.begin(), lst.end());
            for (int k = 0; k < nk; ++k) {
                const int hn = ch->hop_n[((size_t)cc * nk + k) * 2 + fl];
                for (int h = 0; h < hn; ++h) {
                    const int32_t to = ch->hop_orb[((((size_t)cc * nk + k) * 2 + fl) * 2 + h) * 2 + 1];
                    hop_new[q * kk * 4 + (((size_t)cc * nk + k) * 2 + fl) * 2 + h] =
                        (int32_t)(std::lower_bound(lst.begin(), lst.end(), to) - lst.begin());
                }
            }
            new_off[q * (2 * nk + 1) + (size_t)cc * 2 + fl] = (int32_t)new_orb.size();
            new_orb.insert(new_orb.end(), lst.begin(), lst.end());
        }
      new_off[q * (2 * nk + 1) + (size_t)nk * 2] = (int32_t)new_orb.size();
    }
    if (new_orb.empty()) new_orb.push_back(0);
    // rows of the register-resident kernel: the new orbitals of any current state
    int nqf[2] = {0, 0};
    for (size_t q = 0; q < nch; ++q)
        for (int cc = 0; cc < nk; ++cc)
            for (int fl = 0; fl < 2; ++fl) {
                const size_t o = q * (2 * nk + 1) + (size_t)cc * 2 + fl;
                nqf[fl] = std::max(nqf[fl], new_off[o + 1] - new_off[o]);
            }
    const int rb0 = (nu + nqf[0] + 15) / 16, rb1 = (nd + nqf[1] + 15) / 16;
    int kcb = 0, krbw = 0;
    const bool regk = cfg->lanes_per_row == 0 &&
                      gv_excr_shape(cfg->precision, (nm + 3) / 4, rb0 + rb1, &kcb, &krbw) &&
                      gv_excr_lds_bytes(cfg->precision, L, nu, nd, nqf[0], nqf[1], rb0, rb1, kcb, nk) <= 160 * 1024;
    const int nqm_all = std::max(1, std::max(nqf[0], nqf[1]));
    if (!regk && gv_exc_lds_bytes(Ns, nm, nk, nqm_all) > 156 * 1024)
        return fail(GPVMC_ERR_UNSUPPORTED, "channel too large for one workgroup's LDS");

    gpvmc_ctx *c = new gpvmc_ctx();
    c->cfg = *cfg;
    c->channel = true;
    c->L = L;
    c->Ns = Ns;
    c->N = nm;
    c->n_up = nu;
    c->n_dn = nd;
    c->nk = nk;
    c->W = cfg->n_walkers;
    c->G = cfg->n_groups;
    c->prec = cfg->precision;
    c->have_orb.assign(c->G, 0);
    c->groups.assign(c->W, 0);
    c->n_chan = n_channels;
    c->state_orb = state_orb;
    c->regk = regk;
    c->rk_cb = kcb;
    c->rk_rbw = krbw;
    c->nq0 = nqf[0];
    c->nq1 = nqf[1];
    c->rb0 = rb0;
    c->rb1 = rb1;
    auto bail = [&](int rc) {
        free_ctx(c);
        return rc;
    };
#define CK(expr)                                                                           \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return bail(fail(GPVMC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_))); \
    } while (0)
    CK(hipSetDevice(cfg->device));
    CK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    CK(hipEventCreate(&c->ev0));
    CK(hipEventCreate(&c->ev1));
    const size_t W = c->W;
    CK(hipMalloc(&c->d_Ai64, W * 2 * nm * nm * sizeof(double2)));
    CK(hipMemset(c->d_Ai64, 0, W * 2 * nm * nm * sizeof(double2)));
    const size_t ts = c->prec == GPVMC_FP64 ? sizeof(double2) : sizeof(float2);
    if (regk) {
        const size_t rows = (size_t)16 * (rb0 + rb1), ncol = (size_t)4 * kcb, zt = (size_t)16 * ((nm + 15) / 16);
        const size_t nqm = (size_t)std::max(1, std::max(nqf[0], nqf[1]));
        CK(hipMalloc(&c->d_M, W * rows * ncol * ts));
        CK(hipMemset(c->d_M, 0, W * rows * ncol * ts));
        // the measurement's HBM tiles, only where its LDS tiles do not fit
        if (gv_excr_meas_lds_bytes(cfg->precision, L, nu, nd, nqf[0], nqf[1], rb0, rb1, kcb, nk) > 160 * 1024)
            CK(hipMalloc(&c->d_zscr, W * zt * rows * ts));
        const size_t rec = (size_t)(2 * nk + 1) * sizeof(double2);
        c->hb_cap = (int)std::max<size_t>(1, std::min<size_t>(16, 64 * 1024 / rec));
        CK(hipMalloc(&c->d_hbuf, W * (size_t)c->hb_cap * rec));
        CK(hipMalloc(&c->d_qorb, W * 2 * nqm * sizeof(int32_t)));
        CK(hipMemset(c->d_qorb, 0xff, W * 2 * nqm * sizeof(int32_t)));
    } else {
        if (c->prec == GPVMC_MIXED) {
            CK(hipMalloc(&c->d_Ai32, W * 2 * nm * nm * sizeof(float2)));
            CK(hipMemset(c->d_Ai32, 0, W * 2 * nm * nm * sizeof(float2)));
        }
        // Q rows of the streamed kernel, built with the inverse at every init and refresh
        CK(hipMalloc(&c->d_Q, W * 2 * (size_t)nqm_all * nm * ts));
        CK(hipMemset(c->d_Q, 0, W * 2 * (size_t)nqm_all * nm * ts));
        CK(hipMalloc(&c->d_qorb, W * 2 * (size_t)nqm_all * sizeof(int32_t)));
        CK(hipMemset(c->d_qorb, 0xff, W * 2 * (size_t)nqm_all * sizeof(int32_t)));
        CK(hipMalloc(&c->d_Y, W * 2 * ((size_t)nm + nqm_all) * nm * ts));
    }
    // [G][fl][site][orbital] orbital table of both channel kernels
    CK(hipMalloc(&c->d_tabT, (size_t)c->G * 2 * Ns * Ns * ts));
    if (const size_t wb = gv_detinv_work_bytes(nm)) CK(hipMalloc(&c->d_work, wb * 2 * W));
    CK(hipMalloc(&c->d_site, W * Ns * sizeof(int32_t)));
    CK(hipMalloc(&c->d_part, W * 2 * nm * sizeof(int32_t)));
    CK(hipMalloc(&c->d_cur, W * 2 * nm * sizeof(int32_t)));
    CK(hipMalloc(&c->d_cstate, W * sizeof(int32_t)));
    CK(hipMalloc(&c->d_weight, W * sizeof(double)));
    CK(hipMalloc(&c->d_ho, W * 4 * nk * nk * sizeof(double)));
    CK(hipMalloc(&c->d_group, W * sizeof(int32_t)));
    CK(hipMalloc(&c->d_order, W * sizeof(int32_t)));
    CK(hipMalloc(&c->d_status, W * 2 * sizeof(int32_t)));
    CK(hipMalloc(&c->d_logabs, W * 2 * sizeof(double)));
    CK(hipMalloc(&c->d_phase, W * 2 * sizeof(double)));
    CK(hipMalloc(&c->d_drift, W * 2 * sizeof(double)));
    CK(hipMalloc(&c->d_acc, W * gv::ACC_STRIDE * sizeof(double)));
    CK(hipMalloc(&c->d_counter, W * sizeof(uint64_t)));
    CK(hipMalloc(&c->d_tab64, (size_t)c->G * 2 * Ns * Ns * sizeof(double2)));
    if (c->prec == GPVMC_MIXED) CK(hipMalloc(&c->d_tab32, (size_t)c->G * 2 * Ns * Ns * sizeof(float2)));
    CK(hipMemset(c->d_group, 0, W * sizeof(int32_t)));
    CK(hipMemset(c->d_acc, 0, W * gv::ACC_STRIDE * sizeof(double)));
    CK(hipMemset(c->d_ho, 0, W * 4 * nk * nk * sizeof(double)));
    CK(hipMemset(c->d_counter, 0, W * sizeof(uint64_t)));
    std::vector<int32_t> chan(W);
    for (size_t w = 0; w < W; ++w) chan[w] = (int32_t)(w % nch);
    c->chans.assign(chan.begin(), chan.end());
    auto up = [&](int32_t **d, const std::vector<int32_t> &h) -> hipError_t {
        hipError_t e = hipMalloc(d, h.size() * sizeof(int32_t));
        if (e != hipSuccess) return e;
        return hipMemcpy(*d, h.data(), h.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    };
    CK(up(&c->d_hop_n, hop_n));
    CK(up(&c->d_hop_orb, hop_orb));
    CK(up(&c->d_hop_new, hop_new));
    CK(up(&c->d_new_off, new_off));
    CK(up(&c->d_new_orb, new_orb));
    CK(up(&c->d_hopsign, hopsign));
    CK(up(&c->d_chan, chan));
    std::vector<int32_t> nbr, edge4;
    build_lattice(L, nbr, edge4);
    CK(hipMalloc(&c->d_nbr, nbr.size() * sizeof(int32_t)));
    CK(hipMalloc(&c->d_edge, edge4.size() * sizeof(int32_t)));
    CK(hipMemcpy(c->d_nbr, nbr.data(), nbr.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    CK(hipMemcpy(c->d_edge, edge4.data(), edge4.size() * sizeof(int32_t), hipMemcpyHostToDevice));
#undef CK
    int rc = upload_order(c);
    if (rc) return bail(rc);
    *out = c;
    return GPVMC_OK;
}

int gpvmc_set_orbital_table(gpvmc_ctx *c, int32_t group, const double *table)
{
    if (!c || !table) return fail(GPVMC_ERR_ARG, "null argument");
    if (!c->channel) return fail(GPVMC_ERR_STATE, "ground-state contexts take occupied orbitals (gpvmc_set_orbitals)");
    if (group < 0 || group >= c->G) return fail(GPVMC_ERR_ARG, "group out of range");
    const size_t n = (size_t)2 * c->Ns * c->Ns;
    for (size_t i = 0; i < 2 * n; ++i)
        if (!std::isfinite(table[i])) return fail(GPVMC_ERR_ARG, "non-finite orbital value");
    HIPCHK(hipSetDevice(c->cfg.device));
    HIPCHK(hipMemcpy(c->d_tab64 + (size_t)group * n, table, n * sizeof(double2), hipMemcpyHostToDevice));
    if (c->prec == GPVMC_MIXED) {
        std::vector<float2> f(n);
        for (size_t i = 0; i < n; ++i) f[i] = make_float2((float)table[2 * i], (float)table[2 * i + 1]);
        HIPCHK(hipMemcpy(c->d_tab32 + (size_t)group * n, f.data(), n * sizeof(float2), hipMemcpyHostToDevice));
    }
    if (c->channel) {
        // [fl][site][orbital] copy in the working precision
        const size_t Ns = c->Ns;
        if (c->prec == GPVMC_FP64) {
            std::vector<double2> t(n);
            for (size_t fl = 0; fl < 2; ++fl)
                for (size_t o = 0; o < Ns; ++o)
                    for (size_t v = 0; v < Ns; ++v) {
                        const size_t i = (fl * Ns + o) * Ns + v;
                        t[(fl * Ns + v) * Ns + o] = make_double2(table[2 * i], table[2 * i + 1]);
                    }
            HIPCHK(hipMemcpy(static_cast<double2 *>(c->d_tabT) + (size_t)group * n, t.data(), n * sizeof(double2),
                             hipMemcpyHostToDevice));
        } else {
            std::vector<float2> t(n);
            for (size_t fl = 0; fl < 2; ++fl)
                for (size_t o = 0; o < Ns; ++o)
                    for (size_t v = 0; v < Ns; ++v) {
                        const size_t i = (fl * Ns + o) * Ns + v;
                        t[(fl * Ns + v) * Ns + o] = make_float2((float)table[2 * i], (float)table[2 * i + 1]);
                    }
            HIPCHK(hipMemcpy(static_cast<float2 *>(c->d_tabT) + (size_t)group * n, t.data(), n * sizeof(float2),
                             hipMemcpyHostToDevice));
        }
    }
    c->have_orb[group] = 1;
    return GPVMC_OK;
}

// ---- batches of contexts swept together ------------------------------------------------

struct gpvmc_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, fork = nullptr;
    struct Member {
        gpvmc_ctx *c;
        int64_t n_steps, meas;
        uint32_t mask;
        hipStream_t s;
        hipEvent_t join;
    };
    std::vector<Member> m;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static void drop_graph(gpvmc_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// one sweep of one member, kernels only (no host synchronisation, no traces), on stream st
static hipError_t enqueue_member(const gpvmc_batch::Member &mb, hipStream_t st)
{
    gpvmc_ctx *c = mb.c;
    const int64_t nmeas = mb.meas > 0 ? (mb.meas == 1 ? mb.n_steps : (mb.n_steps + mb.meas - 1) / mb.meas) : 0;
    if (c->channel) {
        gv::ExcArgs a = exc_args(c, 0);
        a.n_steps = mb.n_steps;
        a.n_total = mb.n_steps;
        a.meas = mb.meas;
        a.n_meas_slots = nmeas;
        if (c->regk) return launch_exc_walk(c, a, mb.n_steps, mb.meas, st);
        return launch_exc(c, a, st);
    }
    gv::SweepArgs a = sweep_args(c, mb.n_steps, mb.meas, mb.mask, nmeas, 0);
    return launch_sweep(c, a, st);
}

int gpvmc_batch_create(int32_t device, gpvmc_batch **out)
{
    if (!out) return fail(GPVMC_ERR_ARG, "null argument");
    *out = nullptr;
    gpvmc_batch *b = new gpvmc_batch();
    b->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    if (e != hipSuccess) {
        gpvmc_batch_destroy(b);
        return fail(GPVMC_ERR_HIP, std::string("batch creation: ") + hipGetErrorString(e));
    }
    *out = b;
    return GPVMC_OK;
}

int gpvmc_batch_add(gpvmc_batch *b, gpvmc_ctx *c, int64_t n_steps, int64_t meas, uint32_t mask)
{
    if (!b || !c) return fail(GPVMC_ERR_ARG, "null argument");
    if (c->cfg.device != b->device) return fail(GPVMC_ERR_ARG, "context on another device");
    if (!c->initialised) return fail(GPVMC_ERR_STATE, "walkers not initialised");
    if (n_steps < 1 || meas < 0) return fail(GPVMC_ERR_ARG, "n_steps must be >= 1 and meas >= 0");
    if (!c->channel && (n_steps > kMaxLaunchSteps || meas > kMaxLaunchSteps))
        return fail(GPVMC_ERR_ARG, "a ground-state batch member takes at most 2^30 steps per sweep");
    if (mask & ~(uint32_t)(GPVMC_MEAS_PROJHEIS | GPVMC_MEAS_STAGMAGNZ | GPVMC_MEAS_OVERLAP | GPVMC_MEAS_STATSPIN))
        return fail(GPVMC_ERR_ARG, "unknown estimator bit in mask");
    for (const auto &mb : b->m)
        if (mb.c == c) return fail(GPVMC_ERR_ARG, "context already in the batch");
    HIPCHK(hipSetDevice(b->device));
    if (!c->channel && (mask & GPVMC_MEAS_STATSPIN) && meas > 0) {
        int rc = ensure_statspin(c);
        if (rc) return rc;
    }
    gpvmc_batch::Member mb{c, n_steps, meas, mask, nullptr, nullptr};
    HIPCHK(hipStreamCreateWithFlags(&mb.s, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&mb.join, hipEventDisableTiming));
    b->m.push_back(mb);
    drop_graph(b);
    return GPVMC_OK;
}

int gpvmc_batch_sweep(gpvmc_batch *b, int32_t n_sweeps, double *device_ms)
{
    if (!b) return fail(GPVMC_ERR_ARG, "null batch");
    if (n_sweeps < 0) return fail(GPVMC_ERR_ARG, "negative n_sweeps");
    if (b->m.empty()) return fail(GPVMC_ERR_STATE, "empty batch");
    HIPCHK(hipSetDevice(b->device));
    // members' own streams may hold earlier work (init, refresh, walks)
    for (const auto &mb : b->m) HIPCHK(hipStreamSynchronize(mb.c->stream));
    if (!b->exec) {
        HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipEventRecord(b->fork, b->stream);
        for (size_t i = 0; e == hipSuccess && i < b->m.size(); ++i) {
            const auto &mb = b->m[i];
            e = hipStreamWaitEvent(mb.s, b->fork, 0);
            if (e == hipSuccess) e = enqueue_member(mb, mb.s);
            if (e == hipSuccess) e = hipEventRecord(mb.join, mb.s);
            if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, mb.join, 0);
        }
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(b->stream, &g);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return fail(GPVMC_ERR_HIP, std::string("sweep capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
        }
        b->graph = g;
        HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    HIPCHK(hipEventRecord(b->ev0, b->stream));
    for (int i = 0; i < n_sweeps; ++i) HIPCHK(hipGraphLaunch(b->exec, b->stream));
    HIPCHK(hipEventRecord(b->ev1, b->stream));
    HIPCHK(hipEventSynchronize(b->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    if (device_ms) *device_ms = ms;
    return GPVMC_OK;
}

int gpvmc_batch_destroy(gpvmc_batch *b)
{
    if (!b) return GPVMC_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (auto &mb : b->m) {
        if (mb.join) (void)hipEventDestroy(mb.join);
        if (mb.s) (void)hipStreamDestroy(mb.s);
    }
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return GPVMC_OK;
}

int gpvmc_channel_kernel(gpvmc_ctx *c, int32_t *register_resident)
{
    if (!c || !register_resident) return fail(GPVMC_ERR_ARG, "null argument");
    if (!c->channel) return fail(GPVMC_ERR_STATE, "not a channel context");
    *register_resident = c->regk ? 1 : 0;
    return GPVMC_OK;
}

int gpvmc_read_amplitudes(gpvmc_ctx *c, double *amp)
{
    if (!c || !amp) return fail(GPVMC_ERR_ARG, "null argument");
    if (!c->channel) return fail(GPVMC_ERR_STATE, "not a channel context");
    if (!c->initialised) return fail(GPVMC_ERR_STATE, "walkers not initialised");
    HIPCHK(hipSetDevice(c->cfg.device));
    const size_t bytes = (size_t)c->W * c->nk * 2 * sizeof(double);
    int rc = ensure_scratch(c, bytes);
    if (rc) return rc;
    gv::ExcArgs a = exc_args(c, 4);
    a.amp_out = static_cast<double *>(c->d_scratch);
    HIPCHK(launch_exc(c, a));
    HIPCHK(hipMemcpyAsync(amp, a.amp_out, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return GPVMC_OK;
}

int gpvmc_read_channel(gpvmc_ctx *c, int32_t w0, int32_t nw, double *ho, int32_t *cstate, double *weight)
{
    if (!c) return fail(GPVMC_ERR_ARG, "null context");
    if (!c->channel) return fail(GPVMC_ERR_STATE, "not a channel context");
    if (w0 < 0 || nw < 0 || (int64_t)w0 + nw > c->W) return fail(GPVMC_ERR_ARG, "walker range out of bounds");
    HIPCHK(hipSetDevice(c->cfg.device));
    HIPCHK(hipStreamSynchronize(c->stream));
    const size_t per = (size_t)4 * c->nk * c->nk;
    if (ho) HIPCHK(hipMemcpy(ho, c->d_ho + (size_t)w0 * per, (size_t)nw * per * sizeof(double), hipMemcpyDeviceToHost));
    if (cstate) HIPCHK(hipMemcpy(cstate, c->d_cstate + w0, (size_t)nw * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (weight) HIPCHK(hipMemcpy(weight, c->d_weight + w0, (size_t)nw * sizeof(double), hipMemcpyDeviceToHost));
    return GPVMC_OK;
}

}  // extern "C"
