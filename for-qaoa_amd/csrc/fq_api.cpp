// fq_api.cpp -- host side of libforqaoa: handles, validation, cost plans, the
// layer schedules and the C ABI declared in include/forqaoa.h.
#include "forqaoa.h"
#include "fq_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <vector>

using namespace fq;

// ---- errors ------------------------------------------------------------------

static thread_local char g_err[512] = "";

static int set_err(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define FQ_HIP(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return set_err(FQ_ERR_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_));   \
    } while (0)

extern "C" const char *fq_last_error(void) { return g_err; }

// ---- HBM budget (states and slabs) ----------------------------------------------

#include <mutex>
static std::mutex g_budget_mu;
static size_t g_budget = SIZE_MAX;
static size_t g_in_use = 0;

static bool budget_take(size_t bytes)
{
    std::lock_guard<std::mutex> lk(g_budget_mu);
    if (g_in_use + bytes > g_budget || g_in_use + bytes < g_in_use)
        return false;
    g_in_use += bytes;
    return true;
}

static void budget_give(size_t bytes)
{
    std::lock_guard<std::mutex> lk(g_budget_mu);
    g_in_use = bytes > g_in_use ? 0 : g_in_use - bytes;
}

extern "C" int fq_set_memory_budget(size_t bytes)
{
    std::lock_guard<std::mutex> lk(g_budget_mu);
    g_budget = bytes ? bytes : SIZE_MAX;
    return FQ_OK;
}

extern "C" int fq_memory_in_use(size_t *bytes)
{
    if (!bytes)
        return set_err(FQ_ERR_ARG, "null output pointer");
    std::lock_guard<std::mutex> lk(g_budget_mu);
    *bytes = g_in_use;
    return FQ_OK;
}
extern "C" int fq_version(void) { return 100; }

// ---- RX coefficients ---------------------------------------------------------------

void fq::rx_coeffs(double beta, int dtype, double *c, double *s)
{
    double sn, cs;
    sincos(beta / 2, &sn, &cs);
    if (dtype != FQ_C64) {
        *c = cs;
        *s = sn;
        return;
    }
    // Search FP32 pairs within 3 ulps of the rounded |cos|, |sin| for the
    // smallest |c^2 + s^2 - 1| + |angle error| (both exact or nearly so in FP64:
    // a product of two FP32 values is exact).  Working on magnitudes and
    // restoring the signs afterwards makes RX(-beta) the exact inverse-transpose
    // partner of RX(beta), as the adjoint's uncompute needs.
    const double ac = std::fabs(cs), as = std::fabs(sn);
    const double half = std::atan2(as, ac);
    const float c0 = (float)ac, s0 = (float)as;
    float cb = c0, sb = s0;
    double best = HUGE_VAL;
    for (int i = -3; i <= 3; ++i) {
        float ci = c0;
        for (int k = 0; k < std::abs(i); ++k)
            ci = std::nextafterf(ci, i < 0 ? 0.0f : 2.0f);
        if (ci < 0.0f || ci > 1.0f)
            continue;
        for (int j = -3; j <= 3; ++j) {
            float sj = s0;
            for (int k = 0; k < std::abs(j); ++k)
                sj = std::nextafterf(sj, j < 0 ? 0.0f : 2.0f);
            if (sj < 0.0f || sj > 1.0f)
                continue;
            const double dn = (double)ci * ci + (double)sj * sj - 1.0;
            const double da = std::atan2((double)sj, (double)ci) - half;
            const double score = std::fabs(dn) + std::fabs(da);
            if (score < best) {
                best = score;
                cb = ci;
                sb = sj;
            }
        }
    }
    *c = std::copysign((double)cb, cs);
    *s = std::copysign((double)sb, sn);
}

void fq::rx_tform(double beta, int dtype, bool flip, double *t, double *sigma_re, double *sigma_im)
{
    double sn, cs;
    sincos(beta / 2, &sn, &cs);
    const bool f = flip && std::fabs(sn) > std::fabs(cs);
    double tt = f ? -cs / sn : sn / cs;
    if (dtype == FQ_C64)
        tt = (double)(float)tt;
    *t = tt;
    const double mag = 1.0 / std::sqrt(1.0 + tt * tt);
    *sigma_re = f ? 0.0 : std::copysign(mag, cs);
    *sigma_im = f ? -std::copysign(mag, sn) : 0.0;
}

// RX(beta) as applied with rx_coeffs is sqrt(c^2 + s^2) times a unitary, a
// global factor per gate.  A P-layer evaluation gives every qubit one RX per
// layer, so its state is prod_p (c_p^2 + s_p^2)^(N/2) times the state of a
// unitary circuit; scaling the launch-control amplitude by the inverse removes
// that factor exactly (1 for complex128, where c^2 + s^2 - 1 is ~1e-16).
static double rx_norm_scale(int N, int P, const double *betas, int dtype)
{
    if (dtype != FQ_C64)
        return 1.0;
    double sc = 1.0;
    for (int p = 0; p < P; ++p) {
        double c, s;
        rx_coeffs(betas[p], dtype, &c, &s);
        sc *= std::pow(c * c + s * s, -0.5 * N);
    }
    return sc;
}

// ---- streams the library created -----------------------------------------------
//
// States and slabs run on streams the library creates (a state may be moved to
// a caller's stream instead).  A cached cost plan that is replaced must first
// be released by every stream it was handed to; the library can order that on
// the device only for its own live streams (a destroyed one finished its work
// when it was destroyed), so it keeps a registry of them.

static std::mutex g_streams_mu;
static std::set<hipStream_t> g_own_streams;

static void own_stream(hipStream_t s, bool live)
{
    std::lock_guard<std::mutex> lk(g_streams_mu);
    if (live)
        g_own_streams.insert(s);
    else
        g_own_streams.erase(s);
}

static bool is_own_stream(hipStream_t s)
{
    std::lock_guard<std::mutex> lk(g_streams_mu);
    return g_own_streams.count(s) != 0;
}

// ---- handles -------------------------------------------------------------------

struct PlanEntry {
    void *dmem = nullptr;
    size_t cap = 0;               // bytes of dmem and staging
    PlanDev dev{};
    void *staging = nullptr;      // pinned copy of the blob (asynchronous upload)
    hipEvent_t ready = nullptr;   // upload done
    hipStream_t stream = nullptr; // stream the upload was queued on
    std::vector<hipStream_t> users; // streams the plan was handed to
    bool foreign = false;         // handed to a stream the library did not create
    uint64_t last_use = 0;        // fq_graph::use_clock at the last lookup
};

// Releases a plan's memory; the caller has synchronized the device (hipFree
// waits for the whole device anyway).
static void plan_free(PlanEntry &e)
{
    hipFree(e.dmem);
    hipHostFree(e.staging);
    if (e.ready)
        hipEventDestroy(e.ready);
    e = PlanEntry();
}

// Plans cached per graph; beyond this the least recently used plan's buffers
// are reused for the next one.
#define FQ_PLAN_CACHE 64

struct fq_graph {
    int N = 0;
    std::vector<double> W;  // N*N, W[min*N+max], current labels
    // key: tile geometry + the bytes of W, so a hit is an exact content match
    std::map<std::string, PlanEntry> plans;
    uint64_t version = 1;       // bumped by every change of W
    uint64_t plan_epoch = 0;    // bumped when a cached plan is dropped or reused
    uint64_t use_clock = 0;
};

struct TimingRec {
    int kind;
    hipEvent_t e0, e1;
    double bytes;
};

struct fq_state {
    int N = 0, D = 0, B = 0, C = 0, dtype = 0;
    size_t elem = 16, bytes = 0;
    void *sv = nullptr;
    hipStream_t own = nullptr, stream = nullptr;
    int qmap[FQ_MAXN];
    double init_amp = 0;
    // product of the RX factors of tile passes that left theirs to a later
    // pass (run_layer); 1 between entry points
    double pending_re = 1.0, pending_im = 0.0;
    // inside fq_run: every RX applies one angle to all qubits of an
    // X^(x)N-symmetric state, so rx_tform may drop its X (see rx_tform)
    bool in_run = false;
    double *d_part = nullptr;  // expect partials
    size_t part_cap = 0;
    double *d_out2 = nullptr;
    double *d_cand_p = nullptr;
    uint64_t *d_cand_i = nullptr;
    int cand_blocks = 0;
    double2 *d_phase = nullptr;  // per-launch cost phase factors
    size_t phase_cap = 0;        // bytes
    void *aux = nullptr;         // adjoint state (lambda), allocated on first use
    double *d_grad = nullptr;    // adjoint per-pass sums
    size_t grad_cap = 0;
    bool timing = false;
    std::vector<TimingRec> pending;
    std::vector<hipEvent_t> pool;
    double t_ms[FQ_K_COUNT] = {0};
    int64_t t_cnt[FQ_K_COUNT] = {0};
    double t_bytes[FQ_K_COUNT] = {0};
};

// ---- timing ------------------------------------------------------------------

static hipEvent_t pool_get(fq_state *st)
{
    if (!st->pool.empty()) {
        hipEvent_t e = st->pool.back();
        st->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess)
        return nullptr;
    return e;
}

static int timing_collect(fq_state *st)
{
    size_t done = 0;
    int rc = FQ_OK;
    for (; done < st->pending.size(); ++done) {
        const TimingRec &r = st->pending[done];
        hipError_t e = hipEventSynchronize(r.e1);
        float ms = 0;
        if (e == hipSuccess)
            e = hipEventElapsedTime(&ms, r.e0, r.e1);
        if (e != hipSuccess) {
            rc = set_err(FQ_ERR_DEVICE, "timing event failed: %s", hipGetErrorString(e));
            break;
        }
        st->t_ms[r.kind] += ms;
        st->t_cnt[r.kind] += 1;
        st->t_bytes[r.kind] += r.bytes;
        st->pool.push_back(r.e0);
        st->pool.push_back(r.e1);
    }
    // collected records leave the list (a failed one stays for a later call)
    st->pending.erase(st->pending.begin(), st->pending.begin() + (std::ptrdiff_t)done);
    return rc;
}

// Wrap one launch with events when timing is on.
template <typename F>
static int timed(fq_state *st, int kind, double bytes, F &&launch)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (st->timing) {
        if (st->pending.size() > 4096) {
            int rc = timing_collect(st);
            if (rc)
                return rc;
        }
        e0 = pool_get(st);
        e1 = pool_get(st);
        if (!e0 || !e1 || hipEventRecord(e0, st->stream) != hipSuccess) {
            if (e0)
                st->pool.push_back(e0);
            if (e1)
                st->pool.push_back(e1);
            return set_err(FQ_ERR_DEVICE, "timing event unavailable");
        }
    }
    hipError_t e = launch();
    if (e == hipSuccess && st->timing)
        e = hipEventRecord(e1, st->stream);
    if (e != hipSuccess) {
        if (st->timing) {
            st->pool.push_back(e0);
            st->pool.push_back(e1);
        }
        return set_err(FQ_ERR_DEVICE, "kernel launch failed: %s", hipGetErrorString(e));
    }
    if (st->timing)
        st->pending.push_back({kind, e0, e1, bytes});
    return FQ_OK;
}

// ---- geometry ------------------------------------------------------------------

struct Geom {
    int N, Q, a_log, nthr, nh;
    uint8_t tpos[16];
    uint8_t hpos[FQ_MAXN];
};

// tile: positions in order (thread bits first, the last a_log are register bits)
static Geom make_geom(int N, const std::vector<int> &tile, int a_log)
{
    Geom g{};
    g.N = N;
    g.Q = (int)tile.size();
    g.a_log = a_log;
    g.nthr = g.Q - a_log;
    g.nh = N - g.Q;
    bool in_tile[FQ_MAXN] = {false};
    for (int i = 0; i < g.Q; ++i) {
        g.tpos[i] = (uint8_t)tile[i];
        in_tile[tile[i]] = true;
    }
    int h = 0;
    for (int p = 0; p < N; ++p)
        if (!in_tile[p])
            g.hpos[h++] = (uint8_t)p;
    return g;
}

// Group consecutive index bits that land on consecutive positions; returns
// the number of runs (may exceed FQ_MAX_RUNS: then r is not usable).
static int make_runs(const uint8_t *pos, int n, BitRuns &r)
{
    int cnt = 0;
    r.n = 0;
    for (int i = 0; i < n;) {
        int j = i + 1;
        while (j < n && pos[j] == pos[j - 1] + 1)
            ++j;
        if (cnt < FQ_MAX_RUNS)
            r.run[cnt] = (uint32_t)i | ((uint32_t)(j - i) << 8) | ((uint32_t)pos[i] << 16);
        cnt++;
        i = j;
    }
    r.n = cnt < FQ_MAX_RUNS ? cnt : FQ_MAX_RUNS;
    return cnt;
}

static int geom_runs(const Geom &g)
{
    BitRuns r;
    return std::max(make_runs(g.hpos, g.nh, r), make_runs(g.tpos, g.nthr, r));
}

// Tile order of the persistent passes (tile_at): 2, one contiguous range of
// tiles per work-group (measured fastest on MI355X for the contiguous and
// the strided passes); FQ_TILE_ORDER (environment, read per call) selects
// another.
static uint32_t tile_order()
{
    const char *e = getenv("FQ_TILE_ORDER");
    return (uint32_t)(e ? std::max(0, std::min(3, atoi(e))) : 2);
}

static void geom_to_args(const Geom &g, TileArgs &a)
{
    memset(&a, 0, sizeof(a));
    a.scale_re = 1.0;
    a.order = tile_order();
    a.N = g.N;
    a.Q = g.Q;
    a.nthr = g.nthr;
    a.nh = g.nh;
    make_runs(g.hpos, g.nh, a.hruns);
    make_runs(g.tpos, g.nthr, a.truns);
    a.rpos_w = 0;
    for (int r = 0; r < g.a_log; ++r) {
        a.rpos[r] = g.tpos[g.nthr + r];
        a.rpos_w |= (uint32_t)g.tpos[g.nthr + r] << (6 * r);
    }
}

static std::string geom_key(const Geom &g)
{
    std::string k;
    k.push_back((char)g.N);
    k.push_back((char)g.Q);
    k.push_back((char)g.a_log);
    k.append(reinterpret_cast<const char *>(g.tpos), g.Q);
    return k;
}

// RX groups: gate bits among the thread bits, a_log per LDS phase.
// lane_swap: thread bits 4 and 5 (gated) are exchanged with register bits 1
// and 0 by cross-lane swaps and get their RX there; the LDS phases take the
// other gated thread bits.
static int set_gates(TileArgs &a, const Geom &g, uint32_t gate_mask, bool lane_swap = false)
{
    a.gate_mask = gate_mask;
    a.lane_swap = lane_swap ? 1u : 0u;
    // gated thread bits, lowest first: the last LDS phase gets the highest
    // ones, so that it usually holds none of the six lane bits and the tile
    // can be stored straight from its layout (direct_store)
    std::vector<int> pend;
    for (int i = 0; i < g.nthr; ++i)
        if (((gate_mask >> i) & 1u) && !(lane_swap && (i == 4 || i == 5)))
            pend.push_back(i);
    a.ngroups = 0;
    a.grp_mask_w = 0;
    a.grp_gates_w = 0;
    a.direct_store = 0;
    size_t pos = 0;
    while (pos < pend.size()) {
        if (a.ngroups >= FQ_MAX_GROUPS)
            return set_err(FQ_ERR_ARG, "too many LDS phases");
        std::vector<int> bits;
        for (int r = 0; r < g.a_log && pos < pend.size(); ++r)
            bits.push_back(pend[pos++]);
        const int real = (int)bits.size();
        // fill the phase with tile bits that receive no gate in it, from the
        // top (register bits first)
        for (int i = g.Q - 1; (int)bits.size() < g.a_log && i >= 0; --i)
            if (std::find(bits.begin(), bits.end(), i) == bits.end())
                bits.push_back(i);
        std::vector<int> sorted = bits;
        std::sort(sorted.begin(), sorted.end());
        uint16_t gm = 0;
        uint8_t gg = 0;
        for (int r = 0; r < g.a_log; ++r) {
            gm |= (uint16_t)(1u << sorted[r]);
            for (int q = 0; q < real; ++q)
                if (bits[q] == sorted[r])
                    gg |= (uint8_t)(1u << r);
        }
        a.grp_mask[a.ngroups] = gm;
        a.grp_gates[a.ngroups] = gg;
        if (a.ngroups < 4)
            a.grp_mask_w |= (uint64_t)gm << (16 * a.ngroups);
        a.grp_gates_w |= (uint32_t)(gg & 0x1f) << (5 * a.ngroups);
        a.ngroups++;
    }
    a.dpos_w = 0;
    if (a.ngroups > 0 && g.nthr >= 6 && (a.grp_mask[a.ngroups - 1] & 0x3f) == 0) {
        a.direct_store = 1;
        // register r of the last phase is its r-th lowest tile bit
        for (int i = 0, r = 0; i < g.Q; ++i)
            if ((a.grp_mask[a.ngroups - 1] >> i) & 1u)
                a.dpos_w |= (uint32_t)g.tpos[i] << (6 * r++);
    }
    return FQ_OK;
}

// ---- cost plans ------------------------------------------------------------------

static size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

// The cut plan of graph G for tile geometry g as one blob (doubles first,
// then ints, each array 16-byte aligned); dev receives the plan's scalars
// (place_plan fills in the arrays' device addresses).
struct PlanBlob {
    std::vector<char> bytes;
    PlanDev dev{};
    size_t o_cH = 0, o_om = 0, o_cw = 0, o_hw = 0, o_thr = 0, o_off = 0, o_ch = 0, o_hu = 0, o_hv = 0;
};

static void build_plan(const fq_graph *G, const Geom &g, PlanBlob &pb)
{
    const int N = G->N;
    int ti[FQ_MAXN], hi[FQ_MAXN];
    for (int p = 0; p < N; ++p)
        ti[p] = hi[p] = -1;
    for (int i = 0; i < g.Q; ++i)
        ti[g.tpos[i]] = i;
    for (int h = 0; h < g.nh; ++h)
        hi[g.hpos[h]] = h;

    std::vector<std::vector<std::pair<int, double>>> csr(g.Q);
    std::vector<double> cH(g.Q, 0.0), omega(std::max(g.nh, 1), 0.0);
    std::vector<int> hh_u, hh_v;
    std::vector<double> hh_w;
    struct TE {
        int a, b;
        double w;
    };
    std::vector<TE> te;
    double wtot = 0;
    for (int i = 0; i < N; ++i)
        for (int j = i + 1; j < N; ++j) {
            const double w = G->W[(size_t)i * N + j];
            if (w == 0)
                continue;
            wtot += w;
            if (ti[i] < 0 && ti[j] < 0) {
                hh_u.push_back(hi[i]);
                hh_v.push_back(hi[j]);
                hh_w.push_back(w);
            } else if (ti[i] >= 0 && ti[j] < 0) {
                csr[ti[i]].push_back({hi[j], w});
                cH[ti[i]] += w;
                omega[hi[j]] += w;
            } else if (ti[i] < 0 && ti[j] >= 0) {
                csr[ti[j]].push_back({hi[i], w});
                cH[ti[j]] += w;
                omega[hi[i]] += w;
            } else {
                te.push_back({ti[i], ti[j], w});
            }
        }
    const int nthr = g.nthr, A = 1 << g.a_log, T = 1 << nthr;
    std::vector<double> thr((size_t)(1 + g.a_log) * T, 0.0);
    for (int t = 0; t < T; ++t) {
        double base = 0;
        for (const TE &e : te) {
            const bool at = e.a < nthr, bt = e.b < nthr;
            if (at && bt) {
                base += e.w * (double)(((t >> e.a) ^ (t >> e.b)) & 1);
            } else if (at != bt) {
                const int th = at ? e.a : e.b;
                const int r = (at ? e.b : e.a) - nthr;
                const int z = (t >> th) & 1;
                base += e.w * z;
                thr[(size_t)(1 + r) * T + t] += e.w * (1 - 2 * z);
            }
        }
        thr[t] = base;
    }
    PlanDev &dev = pb.dev;
    for (int j = 0; j < FQ_MAX_REGS; ++j)
        dev.rr[j] = 0;
    for (int j = 0; j < A; ++j) {
        double v = 0;
        for (const TE &e : te)
            if (e.a >= nthr && e.b >= nthr)
                v += e.w * (double)(((j >> (e.a - nthr)) ^ (j >> (e.b - nthr))) & 1);
        dev.rr[j] = v;
    }
    dev.wtot = wtot;
    dev.n_hh = (int)hh_w.size();

    std::vector<int> off(g.Q + 1, 0);
    std::vector<int> ch;
    std::vector<double> cw;
    for (int i = 0; i < g.Q; ++i) {
        off[i] = (int)ch.size();
        for (auto &pr : csr[i]) {
            ch.push_back(pr.first);
            cw.push_back(pr.second);
        }
    }
    off[g.Q] = (int)ch.size();

    // blob layout
    pb.o_cH = 0;
    pb.o_om = align16(pb.o_cH + sizeof(double) * g.Q);
    pb.o_cw = align16(pb.o_om + sizeof(double) * omega.size());
    pb.o_hw = align16(pb.o_cw + sizeof(double) * std::max<size_t>(cw.size(), 1));
    pb.o_thr = align16(pb.o_hw + sizeof(double) * std::max<size_t>(hh_w.size(), 1));
    pb.o_off = align16(pb.o_thr + sizeof(double) * thr.size());
    pb.o_ch = align16(pb.o_off + sizeof(int) * off.size());
    pb.o_hu = align16(pb.o_ch + sizeof(int) * std::max<size_t>(ch.size(), 1));
    pb.o_hv = align16(pb.o_hu + sizeof(int) * std::max<size_t>(hh_u.size(), 1));
    const size_t total = align16(pb.o_hv + sizeof(int) * std::max<size_t>(hh_v.size(), 1));
    pb.bytes.assign(total, 0);
    auto put = [&](size_t o, const void *src, size_t n) {
        if (n)
            memcpy(pb.bytes.data() + o, src, n);
    };
    put(pb.o_cH, cH.data(), sizeof(double) * g.Q);
    put(pb.o_om, omega.data(), sizeof(double) * omega.size());
    put(pb.o_cw, cw.data(), sizeof(double) * cw.size());
    put(pb.o_hw, hh_w.data(), sizeof(double) * hh_w.size());
    put(pb.o_thr, thr.data(), sizeof(double) * thr.size());
    put(pb.o_off, off.data(), sizeof(int) * off.size());
    put(pb.o_ch, ch.data(), sizeof(int) * ch.size());
    put(pb.o_hu, hh_u.data(), sizeof(int) * hh_u.size());
    put(pb.o_hv, hh_v.data(), sizeof(int) * hh_v.size());
}

// The plan of a blob that lives at base (device memory).
static PlanDev place_plan(const PlanBlob &pb, char *base)
{
    PlanDev dev = pb.dev;
    dev.cH = (const double *)(base + pb.o_cH);
    dev.omega = (const double *)(base + pb.o_om);
    dev.csr_w = (const double *)(base + pb.o_cw);
    dev.hh_w = (const double *)(base + pb.o_hw);
    dev.thr = (const double *)(base + pb.o_thr);
    dev.csr_off = (const int *)(base + pb.o_off);
    dev.csr_h = (const int *)(base + pb.o_ch);
    dev.hh_u = (const int *)(base + pb.o_hu);
    dev.hh_v = (const int *)(base + pb.o_hv);
    return dev;
}

// Device cut plan of graph G for tile geometry g, cached per (geometry, W
// content).  A new plan is uploaded asynchronously on `stream`; a cached plan
// whose upload was queued on another stream makes `stream` wait for it.
// With FQ_PLAN_CACHE plans cached, the least recently used one whose buffers
// are large enough is replaced: `stream` first waits (on the device, through
// an event) for the work queued so far on every stream that plan was handed
// to, so no launch still reading it sees the new contents, and the host
// blocks on nothing but that plan's own earlier upload.
static int get_plan(fq_graph *G, const Geom &g, PlanDev *out, hipStream_t stream)
{
    std::string key = geom_key(g);
    key.append(reinterpret_cast<const char *>(G->W.data()), G->W.size() * sizeof(double));
    auto note_user = [&](PlanEntry &e) {
        e.last_use = ++G->use_clock;
        if (std::find(e.users.begin(), e.users.end(), stream) == e.users.end()) {
            e.users.push_back(stream);
            e.foreign |= !is_own_stream(stream);
        }
    };
    auto it = G->plans.find(key);
    if (it != G->plans.end()) {
        PlanEntry &e = it->second;
        note_user(e);
        if (e.ready && e.stream != stream) {
            const hipError_t q = hipEventQuery(e.ready);
            if (q == hipErrorNotReady) {
                hipGetLastError();
                FQ_HIP(hipStreamWaitEvent(stream, e.ready, 0));
            } else if (q != hipSuccess) {
                return set_err(FQ_ERR_DEVICE, "plan upload failed: %s", hipGetErrorString(q));
            }
        }
        *out = e.dev;
        return FQ_OK;
    }
    PlanBlob pb;
    build_plan(G, g, pb);
    const size_t total = pb.bytes.size();
    PlanEntry pe;
    bool reused = false;
    if (G->plans.size() >= FQ_PLAN_CACHE) {
        auto victim = G->plans.end();
        for (auto v = G->plans.begin(); v != G->plans.end(); ++v)
            if (v->second.cap >= total && (victim == G->plans.end() || v->second.last_use < victim->second.last_use))
                victim = v;
        if (victim == G->plans.end()) {
            // every cached plan is smaller than this one: drop them all (rare;
            // hipFree waits for the device in any case)
            FQ_HIP(hipDeviceSynchronize());
            for (auto &kv : G->plans)
                plan_free(kv.second);
            G->plans.clear();
        } else {
            pe = victim->second;
            G->plans.erase(victim);
            reused = true;
            // the host rewrites staging: its previous upload must be done
            FQ_HIP(hipEventSynchronize(pe.ready));
            if (pe.foreign) {
                // a caller's stream may still read it and cannot be tracked
                FQ_HIP(hipDeviceSynchronize());
            } else {
                for (hipStream_t u : pe.users) {
                    if (u == stream || !is_own_stream(u))
                        continue;
                    FQ_HIP(hipEventRecord(pe.ready, u));
                    FQ_HIP(hipStreamWaitEvent(stream, pe.ready, 0));
                }
            }
            pe.users.clear();
            pe.foreign = false;
        }
        G->plan_epoch++;
    }
    if (!reused) {
        size_t cap = 4096;
        while (cap < total)
            cap *= 2;
        pe.cap = cap;
        if (hipMalloc(&pe.dmem, cap) != hipSuccess ||
            hipHostMalloc(&pe.staging, cap, hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&pe.ready, hipEventDisableTiming) != hipSuccess) {
            hipGetLastError();
            plan_free(pe);
            return set_err(FQ_ERR_OOM, "plan allocation of %zu bytes failed", cap);
        }
    }
    memcpy(pe.staging, pb.bytes.data(), total);
    pe.stream = stream;
    if (hipMemcpyAsync(pe.dmem, pe.staging, total, hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipEventRecord(pe.ready, stream) != hipSuccess) {
        hipGetLastError();
        hipStreamSynchronize(stream);
        plan_free(pe);
        return set_err(FQ_ERR_DEVICE, "plan upload failed");
    }
    const PlanDev dev = place_plan(pb, (char *)pe.dmem);
    pe.dev = dev;
    pe.users.clear();
    pe.foreign = false;
    PlanEntry &ne = G->plans[key];
    ne = pe;
    note_user(ne);
    *out = dev;
    return FQ_OK;
}

// ---- state --------------------------------------------------------------------------

static int check_state(const fq_state *st)
{
    if (!st)
        return set_err(FQ_ERR_ARG, "null state");
    return FQ_OK;
}

// FQ_WIDE_REGS=0 (environment, read per call): the contiguous 14-qubit
// complex64 tile passes always run 1024 threads x 16 amplitudes (tile_a_log).
static bool wide_regs()
{
    const char *e = getenv("FQ_WIDE_REGS");
    return e ? atoi(e) != 0 : true;
}

// Register bits per thread of a packed slab pass.
static int slab_a_log(int dtype, int Q)
{
    const int want = dtype == FQ_C128 ? 3 : 4;
    return std::max(1, std::min(want, Q));
}

// Register bits per thread of a tile pass.  The contiguous full-width
// complex64 tile (qubits 0..13, 128 KiB of LDS) can run 512 threads x 32
// amplitudes instead of 1024 x 16: five register bits per LDS phase, and the
// 256 VGPRs of a 512-thread group hold the 32 amplitudes plus the next tile's
// prefetch.  It is used where it needs fewer LDS phases than four bits would
// (N=34: RX on qubits 4..13 in one phase instead of two, one of them half
// filler), which measured on MI355X: boundary pass 61.6 -> 55.6 ms, plain
// 48.4 -> 47.1 ms, layer-0 pass 37 -> 34 ms, 8.27 -> 8.39 layers/s.  With the
// same number of phases the wider tile only adds filler butterflies and
// halves the waves per CU (N=30: 130.2 vs 127.9 layers/s), and the strided
// tiles measured slower with it (N=34 plain strided pass 55.4 -> 60.5 ms).
// Only the write-only layer-0 pass (first) keeps it where qubits 4 and 5 are
// gated: at 1024 x 16 the layer-boundary pass then gets their RX by lane
// swaps and needs one LDS phase, and the other passes drop the half-filled
// last phase's filler butterflies; N=34 boundary pass 0 57.3-58.3 -> 50.8 ms,
// plain pass 0 46.9 -> 45.6 ms, while the layer-0 pass at 1024 x 16 took
// 35.5 ms against 33.6-34.2.
static int tile_a_log(int dtype, const std::vector<int> &tile, const std::vector<int> &gates, bool first = false)
{
    const int Q = (int)tile.size();
    bool low = true;
    for (int i = 0; i < Q; ++i)
        low = low && tile[i] == i;
    if (dtype == FQ_C64 && Q == 14 && low && wide_regs()) {
        // gated thread bits (tile position = qubit here) and the LDS phases they need
        int g4 = 0, g5 = 0;
        bool l4 = false, l5 = false;
        for (int q : gates) {
            g4 += q < 10;
            g5 += q < 9;
            l4 = l4 || q == 4;
            l5 = l5 || q == 5;
        }
        if ((g5 + 4) / 5 < (g4 + 3) / 4 && (first || !(l4 && l5)))
            return 5;
    }
    // the complex128 write-only layer-0 pass as 512 threads x 16 amplitudes
    // (183 VGPRs, no prefetch to hold): its RX on qubits 5..12 takes one LDS
    // phase of four register bits instead of two of three; N=30 3.55-3.62 ->
    // 2.88 ms (5.96 TB/s written), 68.9-69.1 -> 69.5-69.6 layers/s
    if (dtype == FQ_C128 && Q == 13 && low && first && wide_regs())
        return 4;
    return slab_a_log(dtype, Q);
}

extern "C" int fq_state_create(int N, int D, int B, int C, int dtype, fq_state **out)
{
    if (!out)
        return set_err(FQ_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (dtype != FQ_C128 && dtype != FQ_C64)
        return set_err(FQ_ERR_ARG, "dtype must be FQ_C128 or FQ_C64");
    if (N < 1 || N > 36)
        return set_err(FQ_ERR_ARG, "N=%d out of range [1, 36]", N);
    if (D < 0 || D > N)
        return set_err(FQ_ERR_ARG, "D=%d out of range", D);
    if (D > 0)
        return set_err(FQ_ERR_UNSUPPORTED,
                       "D=%d: multi-device state-vector transposition needs 2^D GPUs; this build "
                       "runs on one GPU (D=0)", D);
    if (B < 0 || B > N - D)
        return set_err(FQ_ERR_ARG, "B=%d must satisfy 0 <= B <= N-D=%d", B, N - D);
    if (C < 1 || C > tile_max_q(dtype))
        return set_err(FQ_ERR_ARG, "C=%d out of range [1, %d] for this dtype (2^C tile in LDS)", C,
                       tile_max_q(dtype));
    fq_state *st = new fq_state();
    st->N = N;
    st->D = D;
    st->B = B;
    st->C = C;
    st->dtype = dtype;
    st->elem = dtype == FQ_C128 ? 16 : 8;
    st->bytes = ((size_t)1 << N) * st->elem;
    for (int p = 0; p < FQ_MAXN; ++p)
        st->qmap[p] = p;
    st->init_amp = pow(1. / sqrt(2.), N);
    if (!budget_take(st->bytes)) {
        delete st;
        return set_err(FQ_ERR_BUDGET, "state of %zu bytes exceeds the memory budget",
                       ((size_t)1 << N) * (dtype == FQ_C128 ? 16 : 8));
    }
    if (hipMalloc(&st->sv, st->bytes) != hipSuccess) {
        hipGetLastError();
        budget_give(((size_t)1 << N) * (dtype == FQ_C128 ? 16 : 8));
        delete st;
        return set_err(FQ_ERR_OOM, "state allocation of %zu bytes failed", ((size_t)1 << N) * (dtype == FQ_C128 ? 16 : 8));
    }
    const int Qr = std::min(12, N);
    st->part_cap = (size_t)1 << (N - Qr);
    st->cand_blocks = (int)std::min<uint64_t>(1024, std::max<uint64_t>(1, ((uint64_t)1 << N) / 4096));
    bool ok = hipMalloc(&st->d_part, sizeof(double) * 2 * st->part_cap) == hipSuccess &&
              hipMalloc(&st->d_out2, sizeof(double) * 2) == hipSuccess &&
              hipMalloc(&st->d_cand_p, sizeof(double) * FQ_TOPK_MAX * st->cand_blocks) == hipSuccess &&
              hipMalloc(&st->d_cand_i, sizeof(uint64_t) * FQ_TOPK_MAX * st->cand_blocks) == hipSuccess &&
              hipStreamCreateWithFlags(&st->own, hipStreamNonBlocking) == hipSuccess;
    if (!ok) {
        hipGetLastError();
        fq_state_destroy(st);
        return set_err(FQ_ERR_OOM, "scratch allocation failed");
    }
    own_stream(st->own, true);
    st->stream = st->own;
    *out = st;
    return FQ_OK;
}

extern "C" int fq_state_destroy(fq_state *st)
{
    if (!st)
        return FQ_OK;
    if (st->stream)
        hipStreamSynchronize(st->stream);
    for (auto &r : st->pending) {
        hipEventDestroy(r.e0);
        hipEventDestroy(r.e1);
    }
    for (auto e : st->pool)
        hipEventDestroy(e);
    if (st->sv)
        budget_give(st->bytes);
    hipFree(st->sv);
    hipFree(st->d_part);
    hipFree(st->d_out2);
    hipFree(st->d_cand_p);
    hipFree(st->d_cand_i);
    hipFree(st->d_phase);
    if (st->aux)
        budget_give(st->bytes);
    hipFree(st->aux);
    hipFree(st->d_grad);
    if (st->own) {
        own_stream(st->own, false);
        hipStreamDestroy(st->own);
    }
    delete st;
    return FQ_OK;
}

extern "C" int fq_state_info(const fq_state *st, int *N, int *C, int *dtype, size_t *bytes)
{
    if (int rc = check_state(st))
        return rc;
    if (N)
        *N = st->N;
    if (C)
        *C = st->C;
    if (dtype)
        *dtype = st->dtype;
    if (bytes)
        *bytes = st->bytes;
    return FQ_OK;
}

extern "C" int fq_state_device_ptr(const fq_state *st, void **dptr)
{
    if (int rc = check_state(st))
        return rc;
    if (!dptr)
        return set_err(FQ_ERR_ARG, "null output pointer");
    *dptr = st->sv;
    return FQ_OK;
}

extern "C" int fq_state_set_stream(fq_state *st, void *hip_stream)
{
    if (int rc = check_state(st))
        return rc;
    FQ_HIP(hipStreamSynchronize(st->stream));
    st->stream = hip_stream ? (hipStream_t)hip_stream : st->own;
    return FQ_OK;
}

extern "C" int fq_state_init_zero(fq_state *st)
{
    if (int rc = check_state(st))
        return rc;
    for (int p = 0; p < FQ_MAXN; ++p)
        st->qmap[p] = p;
    return timed(st, FQ_K_COST, (double)st->bytes,
                 [&] { return launch_fill_zero_one(st->dtype, st->sv, st->N, st->stream); });
}

extern "C" int fq_init_h(fq_state *st)
{
    if (int rc = check_state(st))
        return rc;
    st->init_amp = pow(1. / sqrt(2.), st->N);
    return FQ_OK;
}

// ---- graph ------------------------------------------------------------------------

extern "C" int fq_graph_create(int N, fq_graph **out)
{
    if (!out)
        return set_err(FQ_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (N < 1 || N > FQ_MAXN)
        return set_err(FQ_ERR_ARG, "N=%d out of range", N);
    fq_graph *g = new fq_graph();
    g->N = N;
    g->W.assign((size_t)N * N, 0.0);
    *out = g;
    return FQ_OK;
}

extern "C" int fq_graph_add_edge(fq_graph *g, int u, int v, double w)
{
    if (!g)
        return set_err(FQ_ERR_ARG, "null graph");
    if (u < 0 || v < 0 || u >= g->N || v >= g->N)
        return set_err(FQ_ERR_ARG, "edge (%d,%d) out of range for N=%d", u, v, g->N);
    if (u == v)
        return set_err(FQ_ERR_ARG, "self-loop (%d,%d) is not an edge of the cut", u, v);
    if (!std::isfinite(w))
        return set_err(FQ_ERR_ARG, "edge weight must be finite");
    if (u > v)
        std::swap(u, v);
    g->W[(size_t)u * g->N + v] = w;
    g->version++;
    return FQ_OK;
}

extern "C" int fq_graph_destroy(fq_graph *g)
{
    if (!g)
        return FQ_OK;
    if (!g->plans.empty()) {
        hipDeviceSynchronize();
        for (auto &kv : g->plans)
            plan_free(kv.second);
    }
    delete g;
    return FQ_OK;
}

extern "C" int fq_graph_weights(const fq_graph *g, double *W)
{
    if (!g || !W)
        return set_err(FQ_ERR_ARG, "null argument");
    memcpy(W, g->W.data(), sizeof(double) * g->W.size());
    return FQ_OK;
}

// mapping_after_swap_conti (src/swap.cu:27-50): exchange labels out+t, in+t.
extern "C" int fq_graph_relabel(fq_graph *g, int swap_out, int swap_in, int num_swap)
{
    if (!g)
        return set_err(FQ_ERR_ARG, "null graph");
    const int N = g->N;
    if (num_swap < 0 || swap_out < 0 || swap_out + num_swap > swap_in || swap_in + num_swap > N)
        return set_err(FQ_ERR_ARG, "relabel blocks [%d,+%d) [%d,+%d) invalid for N=%d", swap_out,
                       num_swap, swap_in, num_swap, N);
    double *W = g->W.data();
    for (int t = 0; t < num_swap; ++t) {
        const int a = swap_out + t, b = swap_in + t;
        for (int i = 0; i < N; ++i) {
            if (i == a || i == b)
                continue;
            double &wa = i < a ? W[(size_t)i * N + a] : W[(size_t)a * N + i];
            double &wb = i < b ? W[(size_t)i * N + b] : W[(size_t)b * N + i];
            std::swap(wa, wb);
        }
    }
    g->version++;
    return FQ_OK;
}

// ---- gate ops -------------------------------------------------------------------------

static int graph_matches(const fq_state *st, const fq_graph *g)
{
    if (!g)
        return set_err(FQ_ERR_ARG, "null graph");
    if (g->N != st->N)
        return set_err(FQ_ERR_ARG, "graph has N=%d but the state has N=%d", g->N, st->N);
    return FQ_OK;
}

static std::vector<int> low_tile(int Q)
{
    std::vector<int> t(Q);
    for (int i = 0; i < Q; ++i)
        t[i] = i;
    return t;
}

// Cost-phase launch arguments: angle and launch-control amplitude.
static void set_cost(TileArgs &a, double gamma, double init_amp)
{
    // a.scale (the pass's real RX factor, 1 without gates) is set by the caller
    a.gamma = gamma;
    a.init_amp = init_amp;
    a.has_rr = 0;
    for (int j = 0; j < FQ_MAX_REGS; ++j)
        if (a.plan.rr[j] != 0)
            a.has_rr = 1;
}

// Phase factors for a cost-phase tile launch (grows the state's table; a
// larger table needs the previous launches that read the old one to finish).
// Phase table of one cost launch into d_phase; with cut, the table of cut
// coefficients follows it in the same buffer (*cut points there).
static int prepare_phases(fq_state *st, TileArgs &a, int a_log, double **cut = nullptr)
{
    const size_t n = phase_table_entries(a, a_log);
    const size_t need = n * sizeof(double2) + (cut ? n * sizeof(double) : 0);
    if (need > st->phase_cap) {
        FQ_HIP(hipStreamSynchronize(st->stream));
        hipFree(st->d_phase);
        st->d_phase = nullptr;
        st->phase_cap = 0;
        if (hipMalloc(&st->d_phase, need) != hipSuccess)
            return set_err(FQ_ERR_OOM, "phase table allocation of %zu bytes failed", need);
        st->phase_cap = need;
    }
    double *c = cut ? reinterpret_cast<double *>(st->d_phase + n) : nullptr;
    FQ_HIP(launch_phase_table(a, a_log, st->d_phase, c, st->stream));
    if (cut)
        *cut = c;
    return FQ_OK;
}

// Cost phase alone: 256 threads x 16 amplitudes per work-group, no LDS.
static int cost_pass(fq_state *st, fq_graph *g, double gamma, bool first)
{
    const int Q = std::min(12, st->N);
    const int a_log = std::max(1, std::min(4, Q));
    Geom geom = make_geom(st->N, low_tile(Q), a_log);
    TileArgs a;
    geom_to_args(geom, a);
    if (int rc = get_plan(g, geom, &a.plan, st->stream))
        return rc;
    set_cost(a, gamma, st->init_amp);
    if (int rc = prepare_phases(st, a, a_log))
        return rc;
    const double bytes = (first ? 1.0 : 2.0) * (double)st->bytes;
    return timed(st, FQ_K_COST, bytes, [&] {
        return launch_tile(st->dtype, a_log, true, first, false, st->sv, a, st->d_phase, st->stream);
    });
}

// Half-state (Z2) modes of a tile pass: tiles with z_{N-1} = 0 only, or the
// mirror pass whose top register bit is qubit N-1 (Z2_EXPAND: and that
// writes the full state's upper half, after the last layer).
enum { Z2_OFF = 0, Z2_LOWER = 1, Z2_MIRROR = 2, Z2_EXPAND = 3 };

// Launch arguments of one tile pass of an N-qubit state: tile positions
// (thread bits first), RX gate positions and the half-state mode.
static int pass_args(int N, int dtype, const std::vector<int> &tile, const std::vector<int> &gates, bool cost,
                     int z2, TileArgs &a, Geom &geom, int a_log_override = 0, bool first = false)
{
    const int Q = (int)tile.size();
    const int a_log = a_log_override > 0 ? a_log_override : tile_a_log(dtype, tile, gates, first);
    geom = make_geom(N, tile, a_log);
    if (geom_runs(geom) > FQ_MAX_RUNS)
        return set_err(FQ_ERR_ARG, "internal: tile geometry needs more than %d bit runs", FQ_MAX_RUNS);
    geom_to_args(geom, a);
    uint32_t gm = 0;
    for (int p : gates) {
        int ti = -1;
        for (int i = 0; i < Q; ++i)
            if (tile[i] == p)
                ti = i;
        if (ti < 0)
            return set_err(FQ_ERR_ARG, "internal: gate %d not in tile", p);
        gm |= 1u << ti;
    }
    if (int rc = set_gates(a, geom, gm))
        return rc;
    if (z2 != Z2_OFF) {
        // the lower half is the tiles whose top index bit (qubit N-1) is 0;
        // a mirror pass pairs tile tau with ~tau instead
        const bool ok = z2 == Z2_LOWER ? geom.nh >= 1 && geom.hpos[geom.nh - 1] == N - 1
                                       : geom.nh >= 1 && tile.back() == N - 1 && !cost;
        if (!ok)
            return set_err(FQ_ERR_ARG, "internal: tile does not fit the half-state pass");
        a.ntiles_run = 1ull << (geom.nh - 1);
        if (z2 >= Z2_MIRROR)
            for (int i = 0; i + 1 < Q; ++i)
                a.mirror_pm += 1ull << tile[i];
    }
    return FQ_OK;
}

// One tile pass with RX on gate positions (and optionally the cost phase).
// pre_beta (with cost, not first, not a mirror pass): RX(*pre_beta) on the
// same positions before the cost phase -- the pass that ends one layer and
// opens the next.  The kernel applies each RX as the one-FMA butterfly of
// rx_tform, i.e. sigma^-1 RX per gate; the pass's factor sigma^(#gates) (times
// sigma_pre^(#gates)) and any factor earlier passes left pending go into the
// cost phase of a cost pass; a pass without it multiplies them in before its
// stores, unless defer is set: then they stay pending for the next pass of the
// same fq_run (a layer's RX factors grow the amplitudes by at most 2^(N/2)
// before a cost pass absorbs them).
static int tile_pass(fq_state *st, const std::vector<int> &tile, const std::vector<int> &gates,
                     double beta, fq_graph *g, double gamma, bool cost, bool first, int kind,
                     int z2 = Z2_OFF, const double *pre_beta = nullptr, bool defer = false)
{
    if (pre_beta && (!cost || first || z2 >= Z2_MIRROR))
        return set_err(FQ_ERR_ARG, "internal: a layer-boundary pass needs the cost phase of a later layer");
    TileArgs a;
    Geom geom;
    if (int rc = pass_args(st->N, st->dtype, tile, gates, cost, z2, a, geom, 0, first))
        return rc;
    const int a_log = geom.a_log;
    // complex64 layer-boundary passes on 1024 x 16 tiles whose lane bits 4
    // and 5 are gated: RX on those two by swaps between lanes and registers,
    // one LDS phase fewer (at least one must remain: the pass applies its
    // cost phase in the last phase's layout)
    if (st->dtype == FQ_C64 && a_log == 4 && geom.Q == 14 && z2 < Z2_MIRROR && pre_beta &&
        ((a.gate_mask >> 4) & 3u) == 3u) {
        uint32_t rest = a.gate_mask & ((1u << geom.nthr) - 1u) & ~0x30u;
        if (rest != 0)
            if (int rc = set_gates(a, geom, a.gate_mask, true))
                return rc;
    }
    // the pass's factor: sigma^(#gates) per RX set, times what is pending
    double sr = st->pending_re, si = st->pending_im;
    auto absorb = [&](double b, double *t) {
        double gr, gi;
        rx_tform(b, st->dtype, st->in_run, t, &gr, &gi);
        for (size_t k = 0; k < gates.size(); ++k) {
            const double nr = sr * gr - si * gi;
            si = sr * gi + si * gr;
            sr = nr;
        }
    };
    absorb(beta, &a.rx_t);
    if (pre_beta)
        absorb(*pre_beta, &a.pre_t);
    a.rx_tf = (float)a.rx_t;   // exact for complex64 (rx_tform rounds t to FP32)
    a.pre_tf = (float)a.pre_t;
    if (cost || !defer) {
        a.scale_re = sr;
        a.scale_im = si;
        a.apply_scale = !cost;
        st->pending_re = 1.0;
        st->pending_im = 0.0;
    } else {
        st->pending_re = sr;
        st->pending_im = si;
    }
    if (cost) {
        // the complex64 boundary pass applies the cost phase in its last LDS
        // phase's layout: thread bits = the tile bits outside that phase,
        // register bits = the phase's bits (both in increasing order), so its
        // phase table comes from the plan of that ordering of the same tile
        TileArgs ca = a;
        Geom cgeom = geom;
        if (pre_beta && st->dtype == FQ_C64 && a.ngroups > 0) {
            const unsigned gm = a.grp_mask[a.ngroups - 1];
            std::vector<int> t2;
            for (int i = 0; i < geom.Q; ++i)
                if (!((gm >> i) & 1u))
                    t2.push_back(geom.tpos[i]);
            for (int i = 0; i < geom.Q; ++i)
                if ((gm >> i) & 1u)
                    t2.push_back(geom.tpos[i]);
            cgeom = make_geom(st->N, t2, a_log);
        }
        if (int rc = get_plan(g, cgeom, &ca.plan, st->stream))
            return rc;
        set_cost(ca, gamma, st->init_amp);
        if (int rc = prepare_phases(st, ca, a_log))
            return rc;
        a.plan = ca.plan;
        a.gamma = ca.gamma;
        a.init_amp = ca.init_amp;
        a.has_rr = ca.has_rr;
    }
    // a half-state pass reads and writes half the state; the expanding one
    // writes the whole state
    const double bytes = (double)st->bytes * (z2 == Z2_OFF ? (first ? 1.0 : 2.0) : z2 == Z2_EXPAND ? 1.5 : 1.0);
    return timed(st, kind, bytes, [&] {
        if (z2 >= Z2_MIRROR)
            return launch_tile_mirror(st->dtype, a_log, z2 == Z2_EXPAND, st->sv, a, st->stream);
        if (pre_beta)
            return launch_tile_pre(st->dtype, a_log, st->sv, a, st->d_phase, st->stream);
        return launch_tile(st->dtype, a_log, cost, first, true, st->sv, a, cost ? st->d_phase : nullptr,
                           st->stream);
    });
}

static int carriers_for(int dtype) { return dtype == FQ_C128 ? 3 : 4; }

// Tile width of the fused and strided passes: the LDS limit by default;
// FQ_TILE_Q (environment, read per call) may lower it (smaller tiles: two
// work-groups per CU).
static int pass_tile_q(int dtype)
{
    const char *e = getenv("FQ_TILE_Q");
    const int env_q = e ? atoi(e) : 0;
    const int mx = tile_max_q(dtype);
    return env_q > 0 && env_q < mx ? env_q : mx;
}

// One tile pass of a fused layer: its tile positions and RX targets.
struct FusedPass {
    std::vector<int> tile, gates;
};

// Qubits per block when the strided passes' new qubits are dealt out (see
// cover_passes): 2 by default; FQ_PASS_BLOCK (environment, read per call)
// selects another, 0 for contiguous blocks.
static int pass_block()
{
    const char *e = getenv("FQ_PASS_BLOCK");
    return e ? std::max(0, atoi(e)) : 2;
}

// Whether the carried low qubits' RX moves out of the cost pass (see
// cover_passes): on by default; FQ_CARRIER_GATES=0 (environment, read per
// call) keeps every low qubit's RX in pass 0.
static bool carrier_gates()
{
    const char *e = getenv("FQ_CARRIER_GATES");
    return e ? atoi(e) != 0 : true;
}

// The tile passes of one fused layer, covering every qubit once.  Pass 0 is
// the low Qmax qubits (cost phase and their RX); the other qubits go to the
// fewest passes that fit, each tile being the lowest qubits as carriers
// (runs of at least 128 contiguous bytes) plus the pass's new qubits.  The new qubits are dealt to the
// passes in blocks, from the top qubit down and from the last pass down, so
// the last pass holds qubit N-1 (the half-state mirror pass needs it) and
// every strided pass mixes middle and high index bits.  On MI355X this
// spreads a pass's accesses over the DRAM better than one contiguous block
// of qubits per pass (N=30 complex128: strided passes 8.0 and 7.2 ms -> 6.8
// and 7.0 ms; N=34 complex64: 65 and 98 ms -> 57 and 81 ms).  A pass that
// would need more bit runs than the tile kernel takes gets larger blocks.
static std::vector<FusedPass> cover_passes(int N, int Qmax, int dtype, int boundary = -1)
{
    std::vector<FusedPass> ps;
    Qmax = std::min(Qmax, N);
    FusedPass p1;
    p1.tile = low_tile(Qmax);
    p1.gates = p1.tile;
    ps.push_back(p1);
    const int rest = N - Qmax;
    if (rest <= 0)
        return ps;
    const int cmin = std::min(carriers_for(dtype), Qmax - 1);
    const int per_max = Qmax - cmin;
    const int npass = (rest + per_max - 1) / per_max;
    std::vector<int> want(npass);
    for (int k = 0; k < npass; ++k)
        want[k] = rest / npass + (k < rest % npass ? 1 : 0);
    for (int blk = pass_block() > 0 ? pass_block() : rest;; ++blk) {
        std::vector<std::vector<int>> news(npass);
        int q = N - 1, k = npass - 1;
        while (q >= Qmax) {
            for (int i = 0; i < blk && q >= Qmax && (int)news[k].size() < want[k]; ++i)
                news[k].push_back(q--);
            k = (k + npass - 1) % npass;
        }
        std::vector<FusedPass> more;
        bool fits = true;
        for (int k2 = 0; k2 < npass; ++k2) {
            std::vector<int> &nw = news[k2];
            std::sort(nw.begin(), nw.end());
            const int c = Qmax - want[k2];
            const int a_log = slab_a_log(dtype, Qmax);
            // lane order: the 64 lanes of a wave are the tile's first six
            // positions; past the carriers they take the pass's lowest new
            // qubits whose byte-address bit is outside [19, hi] (one load
            // instruction whose lanes differ only in bits 19..22 measured up
            // to 1.6x slower), else the highest; never N-1, which stays the
            // top register bit for the half-state mirror pass. hi = 22 for
            // states under 32 GiB; from 32 GiB on hi = 28 (lane bits 23..28
            // measured 2-8% slower per layer at N=31..33 complex128 and N=34
            // complex64, while at N=29 complex128 the window to 28 cost 1.5%)
            const int lanes = std::max(0, std::min(6, Qmax - a_log) - c);
            const int elem_log = dtype == FQ_C128 ? 4 : 3;
            const int win_hi = N + elem_log >= 35 ? 28 : 22;
            std::vector<int> lane_q;
            for (int q : nw)
                if ((int)lane_q.size() < lanes && q != N - 1 && (q + elem_log < 19 || q + elem_log > win_hi))
                    lane_q.push_back(q);
            for (int i = (int)nw.size() - 1; i >= 0 && (int)lane_q.size() < lanes; --i)
                if (nw[i] != N - 1 && std::find(lane_q.begin(), lane_q.end(), nw[i]) == lane_q.end())
                    lane_q.push_back(nw[i]);
            std::sort(lane_q.begin(), lane_q.end());
            FusedPass p;
            for (int i = 0; i < c; ++i)
                p.tile.push_back(i);
            for (int q : lane_q)
                p.tile.push_back(q);
            for (int q : nw)
                if (std::find(lane_q.begin(), lane_q.end(), q) == lane_q.end())
                    p.tile.push_back(q);
            p.gates = nw;
            fits = fits && geom_runs(make_geom(N, p.tile, slab_a_log(dtype, Qmax))) <= FQ_MAX_RUNS;
            more.push_back(p);
        }
        if (fits || blk >= rest) {
            if (carrier_gates()) {
                // RX on a carried low qubit may run in any strided pass that
                // carries it (the RX gates of a layer commute and all follow
                // the cost phase of pass 0): give each to the strided pass
                // with the fewest gates, which takes LDS phases and VALU work
                // off the cost pass
                // -- not to the pass given as boundary (pipelined layers run
                // its RX twice per launch), unless no other pass carries q
                for (int q = 0; q < Qmax; ++q) {
                    int best = -1;
                    for (int pass = 0; pass < 2 && best < 0; ++pass)
                        for (int k2 = 0; k2 < npass; ++k2)
                            if (q < Qmax - want[k2] && (pass == 1 || k2 + 1 != boundary) &&
                                (best < 0 || more[k2].gates.size() < more[best].gates.size()))
                                best = k2;
                    if (best < 0)
                        continue;
                    more[best].gates.push_back(q);
                    ps[0].gates.erase(std::find(ps[0].gates.begin(), ps[0].gates.end(), q));
                }
            }
            ps.insert(ps.end(), more.begin(), more.end());
            return ps;
        }
    }
}

// Split arbitrary RX targets into tile passes (targets below C: one chunk pass).
static int rx_targets_once(fq_state *st, std::vector<int> targets, double beta);

static int rx_targets(fq_state *st, std::vector<int> targets, double beta)
{
    if (!st->in_run) {
        // outside fq_run the butterfly keeps its X, so |t| = |tan(beta/2)|:
        // beyond 2.5 (an amplitude growth above 2.7 per gate before the pass's
        // factor is applied) RX(beta) runs as two passes of RX(beta/2)
        double sn, cs;
        sincos(beta / 2, &sn, &cs);
        if (std::fabs(sn) > 2.5 * std::fabs(cs)) {
            if (int rc = rx_targets_once(st, targets, beta / 2))
                return rc;
            return rx_targets_once(st, targets, beta / 2);
        }
    }
    return rx_targets_once(st, targets, beta);
}

static int rx_targets_once(fq_state *st, std::vector<int> targets, double beta)
{
    const int N = st->N;
    const int Ceff = std::min(st->C, N);
    std::sort(targets.begin(), targets.end());
    if (targets.empty())
        return FQ_OK;
    if (targets.back() < Ceff)
        return tile_pass(st, low_tile(Ceff), targets, beta, nullptr, 0, false, false, FQ_K_RX);
    const int Qmax = std::min(pass_tile_q(st->dtype), N);
    const int per = std::max(1, Qmax - std::min(carriers_for(st->dtype), Qmax - 1));
    size_t pos = 0;
    while (pos < targets.size()) {
        std::vector<int> chunk;
        while (pos < targets.size() && (int)chunk.size() < per)
            chunk.push_back(targets[pos++]);
        std::vector<int> tile;
        for (;;) {
            // the chunk's targets plus the lowest other qubits as carriers
            tile = chunk;
            for (int p = 0; (int)tile.size() < Qmax && p < N; ++p)
                if (std::find(chunk.begin(), chunk.end(), p) == chunk.end())
                    tile.push_back(p);
            std::sort(tile.begin(), tile.end());
            const int a_log = tile_a_log(st->dtype, tile, chunk);
            if (chunk.size() == 1 || geom_runs(make_geom(N, tile, a_log)) <= FQ_MAX_RUNS)
                break;
            chunk.pop_back();  // scattered targets: fewer per pass
            --pos;
        }
        if (int rc = tile_pass(st, tile, chunk, beta, nullptr, 0, false, false, FQ_K_RX))
            return rc;
    }
    return FQ_OK;
}

extern "C" int fq_cost_layer(fq_state *st, fq_graph *g, double gamma, int is_first)
{
    if (int rc = check_state(st))
        return rc;
    if (int rc = graph_matches(st, g))
        return rc;
    if (!std::isfinite(gamma))
        return set_err(FQ_ERR_ARG, "gamma must be finite");
    return cost_pass(st, g, gamma, is_first != 0);
}

extern "C" int fq_multi_rx(fq_state *st, const int *targets, int gate_size, double beta)
{
    if (int rc = check_state(st))
        return rc;
    if (gate_size < 0 || (gate_size > 0 && !targets))
        return set_err(FQ_ERR_ARG, "bad target list");
    if (!std::isfinite(beta))
        return set_err(FQ_ERR_ARG, "beta must be finite");
    std::vector<int> t(targets, targets + gate_size);
    std::vector<bool> seen(st->N, false);
    for (int q : t) {
        if (q < 0 || q >= st->N)
            return set_err(FQ_ERR_ARG, "target %d out of range for N=%d", q, st->N);
        if (seen[q])
            return set_err(FQ_ERR_ARG, "target %d listed twice", q);
        seen[q] = true;
    }
    return rx_targets(st, t, beta);
}

extern "C" int fq_swap_conti(fq_state *st, int swap_out, int swap_in, int num_swap)
{
    if (int rc = check_state(st))
        return rc;
    const int N = st->N, k = num_swap;
    if (k < 0 || swap_out < 0 || swap_out + k > swap_in || swap_in + k > N)
        return set_err(FQ_ERR_ARG, "swap blocks [%d,+%d) [%d,+%d) invalid for N=%d", swap_out, k,
                       swap_in, k, N);
    if (k == 0)
        return FQ_OK;
    SwapArgs a;
    memset(&a, 0, sizeof(a));
    a.N = N;
    a.out = swap_out;
    a.in = swap_in;
    a.k = k;
    // out == 0 (the driver's case): pipelined LDS-tiled transpose with tiles
    // of 2^10 (complex128) / 2^12 (complex64) amplitudes read as runs of
    // >= 512 B where the blocks allow; one pair of sub-tiles per work-group
    // for small blocks; otherwise one thread per pair.
    a.t = -1;
    if (swap_out == 0) {
        const int tile_log = st->dtype == FQ_C128 ? 10 : 12;
        const int t = std::min(k, tile_log / 2);
        const int m = t == k ? std::max(0, std::min(tile_log - 2 * k, swap_in - k)) : 0;
        if (2 * t + m == tile_log) {
            a.pipe = 1;
            a.t = t;
            a.m = m;
        } else if (k >= 3) {
            a.t = std::min(k, 5);
        }
        if (a.t >= 0) {
            const uint64_t mb = 1ull << (k - a.t);
            a.npairs = mb * (mb + 1) / 2;
        }
    }
    // rest bits: everything outside both blocks and the tile's rest bits
    bool used[FQ_MAXN] = {false};
    for (int i = 0; i < k; ++i)
        used[swap_out + i] = used[swap_in + i] = true;
    for (int i = 0; i < a.m; ++i)
        used[k + i] = true;
    for (int p = 0; p < N; ++p)
        if (!used[p])
            a.rpos[a.nr++] = (uint8_t)p;
    a.rest_fast = a.pipe && a.nr > 0 && a.rpos[0] < swap_in;
    const double bytes = 2.0 * (double)st->bytes * (1.0 - std::ldexp(1.0, -k));
    int rc = timed(st, FQ_K_SWAP, bytes, [&] { return launch_swap(st->dtype, st->sv, a, st->stream); });
    if (rc)
        return rc;
    for (int i = 0; i < k; ++i)
        std::swap(st->qmap[swap_out + i], st->qmap[swap_in + i]);
    return FQ_OK;
}

// ---- schedules -----------------------------------------------------------------------


// Whether the layers of the fused and half-state schedules are pipelined
// (run_layers): by default (measured on MI355X, N=30 fused: complex128
// 50.0 -> 65.0 layers/s, complex64 94 -> 110); FQ_LAYER_PIPE=0 (environment,
// read per call) runs them in layer order.
static bool layer_pipe()
{
    const char *e = getenv("FQ_LAYER_PIPE");
    return e ? atoi(e) != 0 : true;
}

// The passes of one layer of the fused (z2 false) or half-state schedule and,
// when the layers are pipelined, the pass besides pass 0 that ends every
// other layer (end1 < 0: layer order).  end1 is the last pass of the fused
// schedule and pass 1 of the half-state one (its last pass is the mirror
// pass, which has no cost phase); the carried low qubits' RX goes to the
// other passes.  Measured at N=30 complex128 against ends {0, 1} with the
// carried RX on the fewest-gate pass: fused 62.5 -> 65.0 layers/s,
// half state 120.9 -> 123.8.
struct LayerPlan {
    std::vector<FusedPass> ps;
    int end1;
};

static LayerPlan layer_plan(int N, int dtype, bool z2)
{
    const int Qmax = z2 ? std::min(pass_tile_q(dtype), N - 1) : pass_tile_q(dtype);
    LayerPlan lp{cover_passes(N, Qmax, dtype), -1};
    const int m = (int)lp.ps.size();
    if (layer_pipe() && m >= (z2 ? 3 : 2)) {
        lp.end1 = z2 ? 1 : m - 1;
        if (m >= 3)
            lp.ps = cover_passes(N, Qmax, dtype, lp.end1);
    }
    return lp;
}

// The tile passes of layer p of P (pass 0 the low contiguous tile; z2:
// half-state passes, the last one the mirror pass).  In layer order, layer p
// runs passes 0..m-1 with its cost phase in pass 0.  Pipelined: the cost
// phase is diagonal, so it can run in any pass; the pass that ends layer p
// applies RX(beta_p), the phase of gamma_{p+1} and RX(beta_{p+1}) on its
// qubits, which opens layer p+1.  Layers end alternately on end1 and pass 0,
// so every layer after the first runs m-1 passes: 1 + P(m-1) passes in all
// instead of Pm.  Every RX of a layer still follows its cost phase and
// precedes the next one, and RX gates on distinct qubits commute.  Layer p
// expects layers 0..p-1 to have run with the same plan.  expand: the last
// layer's mirror pass also writes the upper half of the full state.
static int run_layer(fq_state *st, fq_graph *g, const LayerPlan &lp, int p, int P, const double *gammas,
                     const double *betas, bool z2, bool expand)
{
    const std::vector<FusedPass> &ps = lp.ps;
    const int m = (int)ps.size();
    const bool last = p + 1 == P;
    auto mode = [&](int i) {
        if (!z2)
            return (int)Z2_OFF;
        return i + 1 < m ? (int)Z2_LOWER : last && expand ? (int)Z2_EXPAND : (int)Z2_MIRROR;
    };
    // the passes of this layer in launch order: (pass, kind) with kind 0 the
    // layer-0 launch-control pass, 1 a cost pass of layer order, 2 a plain
    // pass, 3 the boundary pass that opens layer p+1
    std::vector<std::pair<int, int>> seq;
    if (lp.end1 < 0) {
        for (int i = 0; i < m; ++i)
            seq.push_back({i, i == 0 ? (p == 0 ? 0 : 1) : 2});
    } else {
        const int e = p % 2 == 0 ? lp.end1 : 0;   // the pass that ends layer p (the other one opened it)
        if (p == 0)
            seq.push_back({0, 0});
        if (last)
            seq.push_back({e, 2});
        for (int i = 1; i < m; ++i)
            if (i != lp.end1)
                seq.push_back({i, 2});
        if (!last)
            seq.push_back({e, 3});
    }
    for (size_t k = 0; k < seq.size(); ++k) {
        const int i = seq[k].first, kind = seq[k].second;
        // plain passes leave their RX factor to the next cost pass, except the
        // run's very last pass, which leaves the state exact
        const bool defer = !(last && k + 1 == seq.size());
        int rc;
        if (kind == 2)
            rc = tile_pass(st, ps[i].tile, ps[i].gates, betas[p], g, 0.0, false, false, FQ_K_FUSED, mode(i),
                           nullptr, defer);
        else if (kind == 3)
            rc = tile_pass(st, ps[i].tile, ps[i].gates, betas[p + 1], g, gammas[p + 1], true, false, FQ_K_FUSED,
                           mode(i), &betas[p]);
        else
            rc = tile_pass(st, ps[i].tile, ps[i].gates, betas[p], g, gammas[p], true, kind == 0, FQ_K_FUSED,
                           mode(i));
        if (rc)
            return rc;
    }
    return FQ_OK;
}

static int run_layers(fq_state *st, fq_graph *g, int P, const double *gammas, const double *betas, bool z2)
{
    const LayerPlan lp = layer_plan(st->N, st->dtype, z2);
    for (int p = 0; p < P; ++p)
        if (int rc = run_layer(st, g, lp, p, P, gammas, betas, z2, true))
            return rc;
    return FQ_OK;
}

extern "C" int fq_run(fq_state *st, fq_graph *g, int P, const double *gammas, const double *betas,
                      int schedule)
{
    if (int rc = check_state(st))
        return rc;
    if (int rc = graph_matches(st, g))
        return rc;
    if (P < 0 || (P > 0 && (!gammas || !betas)))
        return set_err(FQ_ERR_ARG, "bad layer count or angle arrays");
    for (int p = 0; p < P; ++p)
        if (!std::isfinite(gammas[p]) || !std::isfinite(betas[p]))
            return set_err(FQ_ERR_ARG, "angles must be finite");
    const int N = st->N;
    st->pending_re = 1.0;
    st->pending_im = 0.0;
    struct InRun {
        fq_state *st;
        ~InRun() { st->in_run = false; }
    } in_run_guard{st};
    st->in_run = true;
    if (schedule == FQ_SCHED_REFERENCE) {
        // src/Weighted.cpp:97-137
        std::vector<int> sizes;
        for (int i = 0; i < N; i += st->C)
            sizes.push_back(std::min(st->C, N - i));
        for (int p = 0; p < P; ++p) {
            if (int rc = cost_pass(st, g, gammas[p], p == 0))
                return rc;
            int swap_in = 0;
            for (size_t i = 0; i < sizes.size(); ++i) {
                std::vector<int> t(sizes[i]);
                for (int q = 0; q < sizes[i]; ++q)
                    t[q] = q;
                if (int rc = rx_targets(st, t, betas[p]))
                    return rc;
                if (i + 1 < sizes.size()) {
                    swap_in += sizes[i];
                    if (int rc = fq_swap_conti(st, 0, swap_in, sizes[i + 1]))
                        return rc;
                    if (int rc = fq_graph_relabel(g, 0, swap_in, sizes[i + 1]))
                        return rc;
                }
            }
        }
        return FQ_OK;
    }
    if (schedule == FQ_SCHED_Z2) {
        // Cut(b) = Cut(~b), RX^{(x)N} commutes with X^{(x)N} and |+> is
        // X^{(x)N}-invariant, so psi(b) = psi(~b): the layers run on the
        // half z_{N-1} = 0 and the other half is written once at the end.
        if (N < 3)
            return set_err(FQ_ERR_ARG, "the half-state schedule needs N >= 3");
        return run_layers(st, g, P, gammas, betas, true);
    }
    if (schedule != FQ_SCHED_FUSED)
        return set_err(FQ_ERR_ARG, "unknown schedule %d", schedule);
    return run_layers(st, g, P, gammas, betas, false);
}

extern "C" int fq_layer_passes(int N, int dtype, int schedule, int max_pass, int *npass, int *tile_q,
                               int *tiles, uint64_t *gates)
{
    if (N < 1 || N > 36 || (dtype != FQ_C128 && dtype != FQ_C64) || !npass || !tile_q || !tiles || !gates)
        return set_err(FQ_ERR_ARG, "bad N, dtype or null output");
    std::vector<FusedPass> ps;
    if (schedule == FQ_SCHED_FUSED || (schedule == FQ_SCHED_Z2 && N >= 3))
        ps = layer_plan(N, dtype, schedule == FQ_SCHED_Z2).ps;
    else
        return set_err(FQ_ERR_ARG, "schedule must be FQ_SCHED_FUSED or FQ_SCHED_Z2 (N >= 3)");
    if ((int)ps.size() > max_pass)
        return set_err(FQ_ERR_ARG, "%zu passes do not fit max_pass=%d", ps.size(), max_pass);
    *npass = (int)ps.size();
    *tile_q = (int)ps[0].tile.size();
    for (size_t p = 0; p < ps.size(); ++p) {
        for (int i = 0; i < 16; ++i)
            tiles[p * 16 + i] = i < (int)ps[p].tile.size() ? ps[p].tile[i] : -1;
        uint64_t m = 0;
        for (int q : ps[p].gates)
            m |= 1ull << q;
        gates[p] = m;
    }
    return FQ_OK;
}

// ---- observables -------------------------------------------------------------------------

extern "C" int fq_expect_cut(fq_state *st, fq_graph *g, double *expect, double *norm2)
{
    if (int rc = check_state(st))
        return rc;
    if (int rc = graph_matches(st, g))
        return rc;
    if (!expect)
        return set_err(FQ_ERR_ARG, "null output pointer");
    const int Q = std::min(12, st->N);
    const int a_log = std::max(1, std::min(4, Q));
    Geom geom = make_geom(st->N, low_tile(Q), a_log);
    TileArgs a;
    geom_to_args(geom, a);
    if (int rc = get_plan(g, geom, &a.plan, st->stream))
        return rc;
    int rc = timed(st, FQ_K_REDUCE, (double)st->bytes, [&] {
        return launch_expect(st->dtype, a_log, st->sv, a, st->d_part, st->d_out2, st->stream);
    });
    if (rc)
        return rc;
    double h[2];
    FQ_HIP(hipMemcpyAsync(h, st->d_out2, sizeof(h), hipMemcpyDeviceToHost, st->stream));
    FQ_HIP(hipStreamSynchronize(st->stream));
    *expect = h[0];
    if (norm2)
        *norm2 = h[1];
    return FQ_OK;
}

// ---- allocation-free <C> ----------------------------------------------------------------
//
// fq_expect_cut_static runs the kernels of fq_expect_cut with their scratch
// (per-tile partial sums, the cut plan, the two-double result) in `scratch`,
// a fixed device array of this library, instead of memory it allocates: for
// callers that cannot allocate device memory (code under HIP graph capture,
// serving loops).  Each GPU holds its own copy until the process ends; every
// call on one GPU uses that copy, so calls that overlap on one GPU race.

#define FQ_STATIC_MAX_N 30
#define FQ_STATIC_TILE_Q 12
// 2 doubles per 2^12-amplitude tile at the largest N (4 MiB), the cut plan
// of a 12-qubit tile (under 16 KiB at N=30 on a complete graph), the result
#define FQ_STATIC_PART_BYTES ((size_t)16 << (FQ_STATIC_MAX_N - FQ_STATIC_TILE_Q))
#define FQ_STATIC_PLAN_BYTES ((size_t)64 << 10)
#define FQ_STATIC_BYTES (FQ_STATIC_PART_BYTES + FQ_STATIC_PLAN_BYTES + 256)

static __device__ __align__(256) unsigned char scratch[FQ_STATIC_BYTES];

static std::mutex g_scratch_mu;
static std::vector<unsigned char *> g_scratch_ptr;  // device address of scratch, per device

// Device address of this GPU's scratch: hipGetSymbolAddress once per device,
// with that device current.
static int scratch_address(unsigned char **out)
{
    int dev = 0;
    FQ_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= dev)
        g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_err(FQ_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = (unsigned char *)p;
    }
    *out = g_scratch_ptr[dev];
    return FQ_OK;
}

extern "C" int fq_expect_cut_static(fq_state *st, fq_graph *g, double *expect, double *norm2)
{
    if (int rc = check_state(st))
        return rc;
    if (int rc = graph_matches(st, g))
        return rc;
    if (!expect)
        return set_err(FQ_ERR_ARG, "null output pointer");
    if (st->N > FQ_STATIC_MAX_N)
        return set_err(FQ_ERR_ARG, "N=%d exceeds the static scratch of fq_expect_cut_static (N <= %d)", st->N,
                       FQ_STATIC_MAX_N);
    unsigned char *base = nullptr;
    if (int rc = scratch_address(&base))
        return rc;
    double *part = (double *)base;
    char *plan_mem = (char *)(base + FQ_STATIC_PART_BYTES);
    double *out2 = (double *)(base + FQ_STATIC_PART_BYTES + FQ_STATIC_PLAN_BYTES);
    const int Q = std::min(FQ_STATIC_TILE_Q, st->N);
    const int a_log = std::max(1, std::min(4, Q));
    Geom geom = make_geom(st->N, low_tile(Q), a_log);
    TileArgs a;
    geom_to_args(geom, a);
    PlanBlob pb;
    build_plan(g, geom, pb);
    if (pb.bytes.size() > FQ_STATIC_PLAN_BYTES)
        return set_err(FQ_ERR_ARG, "cut plan of %zu bytes exceeds the static scratch (%zu)", pb.bytes.size(),
                       (size_t)FQ_STATIC_PLAN_BYTES);
    a.plan = place_plan(pb, plan_mem);
    // the previous call on this stream may still read the plan: stream order
    // keeps this upload behind it
    FQ_HIP(hipMemcpyAsync(plan_mem, pb.bytes.data(), pb.bytes.size(), hipMemcpyHostToDevice, st->stream));
    FQ_HIP(launch_expect(st->dtype, a_log, st->sv, a, part, out2, st->stream));
    double h[2];
    FQ_HIP(hipMemcpyAsync(h, out2, sizeof(h), hipMemcpyDeviceToHost, st->stream));
    FQ_HIP(hipStreamSynchronize(st->stream));
    *expect = h[0];
    if (norm2)
        *norm2 = h[1];
    return FQ_OK;
}

// ---- adjoint gradient ------------------------------------------------------------------

static int ensure_aux(fq_state *st, size_t grad_slots)
{
    if (!st->aux) {
        if (!budget_take(st->bytes))
            return set_err(FQ_ERR_BUDGET, "adjoint state of %zu bytes exceeds the memory budget", st->bytes);
        if (hipMalloc(&st->aux, st->bytes) != hipSuccess) {
            hipGetLastError();
            st->aux = nullptr;
            budget_give(st->bytes);
            return set_err(FQ_ERR_OOM, "adjoint state allocation of %zu bytes failed", st->bytes);
        }
    }
    if (grad_slots > st->grad_cap) {
        FQ_HIP(hipStreamSynchronize(st->stream));
        hipFree(st->d_grad);
        st->d_grad = nullptr;
        st->grad_cap = 0;
        if (hipMalloc(&st->d_grad, sizeof(double) * 2 * grad_slots) != hipSuccess)
            return set_err(FQ_ERR_OOM, "gradient scratch allocation failed");
        st->grad_cap = grad_slots;
    }
    return FQ_OK;
}

// Reverse-uncompute gradient of <C> (fq_adjoint.hip).  The state must hold
// psi_P of fq_run(FUSED or Z2) with the same graph and angles (identity qubit
// map); it is consumed (for N >= 3 its lower half is walked back to |+>^N and
// its upper half is left as it was).
extern "C" int fq_backward(fq_state *st, fq_graph *g, int P, const double *gammas,
                           const double *betas, double *expect, double *grad_gammas,
                           double *grad_betas)
{
    if (int rc = check_state(st))
        return rc;
    if (int rc = graph_matches(st, g))
        return rc;
    if (P < 1 || !gammas || !betas || !grad_gammas || !grad_betas)
        return set_err(FQ_ERR_ARG, "bad layer count or null arrays");
    for (int p = 0; p < st->N; ++p)
        if (st->qmap[p] != p)
            return set_err(FQ_ERR_ARG, "fq_backward needs the identity qubit map (fused schedule)");
    const int N = st->N;
    // Half states: psi_P and lambda = Cut psi_P satisfy x(b) = x(~b) (Cut(b) =
    // Cut(~b)), and so do both states walked back through U^+ M^+, so for
    // N >= 3 every pass runs on the half z_{N-1} = 0 (the last one as the
    // mirror pass) and every sum counts each element for itself and its
    // complement: the sums are doubled.  The state's upper half is left
    // stale (the state is consumed).
    const bool z2 = N >= 3;
    const double mult = z2 ? 2.0 : 1.0;
    // 2^12-amplitude tiles of psi and lambda share the LDS (2 x 68 KiB);
    // 512 threads x 8 amplitudes of each state
    const int Qb = std::min(12, z2 ? N - 1 : N);
    // The passes of a layer and, as in fq_run's pipelined layers, the pass
    // besides pass 0 that ends every other layer (end1 < 0: layer order)
    std::vector<FusedPass> ps = cover_passes(N, Qb, st->dtype);
    const int m = (int)ps.size();
    const bool pipe = layer_pipe() && m >= (z2 ? 3 : 2);
    const int end1 = pipe ? (z2 ? 1 : m - 1) : -1;
    if (pipe && m >= 3)
        ps = cover_passes(N, Qb, st->dtype, end1);
    // The walk back undoes, in reverse order, the launches of a forward run
    // over this plan: kind 1 a cost pass of layer p (phase, then RX(beta_p)
    // on its qubits: undone as RX then phase), kind 2 RX(beta_p) alone, kind
    // 3 a layer-boundary pass (RX(beta_p), phase gamma_{p+1}, RX(beta_{p+1})).
    // The unitary is the same whichever decomposition the forward used.
    struct Launch {
        int pass, kind, p;
    };
    std::vector<Launch> fwd;
    for (int p = 0; p < P; ++p) {
        if (!pipe) {
            for (int i = 0; i < m; ++i)
                fwd.push_back({i, i == 0 ? 1 : 2, p});
            continue;
        }
        const bool last = p + 1 == P;
        const int e = p % 2 == 0 ? end1 : 0;
        if (p == 0)
            fwd.push_back({0, 1, 0});
        if (last)
            fwd.push_back({e, 2, p});
        for (int i = 1; i < m; ++i)
            if (i != end1)
                fwd.push_back({i, 2, p});
        if (!last)
            fwd.push_back({e, 3, p});
    }
    // gradient slots (two doubles each): slot 0 the <C> sums, then per launch
    // one slot (beta sum, gamma sum), a boundary pass a second (beta_p sum)
    size_t nslot = 1;
    for (const Launch &L : fwd)
        nslot += L.kind == 3 ? 2 : 1;
    if (int rc = ensure_aux(st, nslot))
        return rc;

    // lambda = Cut * psi_P, <C> into slot 0
    {
        const int Q = Qb;
        const int a_log = std::max(1, std::min(4, Q));
        Geom geom = make_geom(N, low_tile(Q), a_log);
        TileArgs a;
        geom_to_args(geom, a);
        if (z2)
            a.ntiles_run = 1ull << (geom.nh - 1);   // qubit N-1 is the top hidden bit
        if (int rc = get_plan(g, geom, &a.plan, st->stream))
            return rc;
        FQ_HIP(launch_cut_mul(st->dtype, a_log, st->sv, st->aux, a, st->d_part, st->stream));
        FQ_HIP(launch_finalize(st->d_part, run_tiles(a), st->d_grad, st->stream));
    }
    // (launch index in fwd, first slot) of every launch, for the sums below
    std::vector<size_t> slot_of(fwd.size());
    {
        size_t s_next = 1;
        for (size_t n = 0; n < fwd.size(); ++n) {
            slot_of[n] = s_next;
            s_next += fwd[n].kind == 3 ? 2 : 1;
        }
    }
    for (size_t n = fwd.size(); n-- > 0;) {
        const Launch &L = fwd[n];
        const int i = L.pass;
        const bool costbw = L.kind != 2;
        const bool boundary = L.kind == 3;
        const int Q = (int)ps[i].tile.size();
        // 512 threads x 8 amplitudes of each state
        const int a_log = std::max(1, std::min(3, Q));
        const bool mirror = z2 && i + 1 == m;
        Geom geom;
        TileArgs a;
        if (int rc = pass_args(N, st->dtype, ps[i].tile, ps[i].gates, costbw,
                               z2 ? (mirror ? Z2_MIRROR : Z2_LOWER) : Z2_OFF, a, geom, a_log))
            return rc;
        // the RX undone first: beta_{p+1} in a boundary pass, else beta_p
        const int pb = boundary ? L.p + 1 : L.p;
        rx_coeffs(-betas[pb], st->dtype, &a.rx_c, &a.rx_s);
        if (boundary)
            rx_coeffs(-betas[L.p], st->dtype, &a.pre_c, &a.pre_s);
        double *ctab = nullptr;
        if (costbw) {
            if (int rc = get_plan(g, geom, &a.plan, st->stream))
                return rc;
            set_cost(a, -gammas[pb], st->init_amp);
            if (int rc = prepare_phases(st, a, a_log, &ctab))
                return rc;
        }
        if (boundary && 2 * run_tiles(a) > st->part_cap)
            return set_err(FQ_ERR_ARG, "internal: partial-sum buffer too small for a boundary pair pass");
        double *slot = st->d_grad + 2 * slot_of[n];
        int rc = timed(st, FQ_K_ADJOINT, 4.0 * (double)st->bytes / mult, [&] {
            unsigned grid = 0;
            hipError_t e = launch_pair(st->dtype, a_log, costbw, mirror, boundary, st->sv, st->aux, a,
                                       costbw ? st->d_phase : nullptr, ctab, st->d_part, &grid, st->stream);
            if (e != hipSuccess)
                return e;
            e = launch_finalize(st->d_part, grid, slot, st->stream);
            if (e != hipSuccess || !boundary)
                return e;
            return launch_finalize(st->d_part + 2 * (size_t)grid, grid, slot + 2, st->stream);
        });
        if (rc)
            return rc;
    }
    std::vector<double> h(2 * nslot);
    FQ_HIP(hipMemcpyAsync(h.data(), st->d_grad, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                          st->stream));
    FQ_HIP(hipStreamSynchronize(st->stream));
    if (expect)
        *expect = mult * h[0];
    for (int k = 0; k < P; ++k) {
        grad_betas[k] = 0.0;
        grad_gammas[k] = 0.0;
    }
    for (size_t n = 0; n < fwd.size(); ++n) {
        const Launch &L = fwd[n];
        const size_t s0 = slot_of[n];
        const int pb = L.kind == 3 ? L.p + 1 : L.p;
        grad_betas[pb] += mult * h[2 * s0];
        if (L.kind != 2)
            grad_gammas[pb] = -mult * h[2 * s0 + 1];
        if (L.kind == 3)
            grad_betas[L.p] += mult * h[2 * (s0 + 1)];
    }
    return FQ_OK;
}

static uint64_t to_logical_index(const fq_state *st, uint64_t b)
{
    uint64_t L = 0;
    for (int p = 0; p < st->N; ++p)
        L |= ((b >> p) & 1ull) << st->qmap[p];
    return L;
}

static uint64_t to_physical_index(const fq_state *st, uint64_t L)
{
    uint64_t b = 0;
    for (int p = 0; p < st->N; ++p)
        b |= ((L >> st->qmap[p]) & 1ull) << p;
    return b;
}

extern "C" int fq_topk(fq_state *st, int k, uint64_t *idx, double *prob)
{
    if (int rc = check_state(st))
        return rc;
    if (k < 1 || k > FQ_TOPK_MAX || !idx || !prob)
        return set_err(FQ_ERR_ARG, "k=%d must be in [1, %d] with non-null outputs", k, FQ_TOPK_MAX);
    if ((uint64_t)k > (1ull << st->N))
        return set_err(FQ_ERR_ARG, "k=%d exceeds the number of basis states", k);
    const int nb = st->cand_blocks;
    int rc = timed(st, FQ_K_TOPK, (double)st->bytes, [&] {
        return launch_topk(st->dtype, st->sv, st->N, nb, st->d_cand_p, st->d_cand_i, st->stream);
    });
    if (rc)
        return rc;
    std::vector<double> cp((size_t)nb * FQ_TOPK_MAX);
    std::vector<uint64_t> ci((size_t)nb * FQ_TOPK_MAX);
    FQ_HIP(hipMemcpyAsync(cp.data(), st->d_cand_p, sizeof(double) * cp.size(),
                          hipMemcpyDeviceToHost, st->stream));
    FQ_HIP(hipMemcpyAsync(ci.data(), st->d_cand_i, sizeof(uint64_t) * ci.size(),
                          hipMemcpyDeviceToHost, st->stream));
    FQ_HIP(hipStreamSynchronize(st->stream));
    std::vector<size_t> order(cp.size());
    for (size_t i = 0; i < order.size(); ++i)
        order[i] = i;
    std::partial_sort(order.begin(), order.begin() + k, order.end(), [&](size_t x, size_t y) {
        return cp[x] > cp[y] || (cp[x] == cp[y] && ci[x] < ci[y]);
    });
    for (int i = 0; i < k; ++i) {
        prob[i] = cp[order[i]];
        idx[i] = to_logical_index(st, ci[order[i]]);
    }
    return FQ_OK;
}

extern "C" int fq_qubit_map(const fq_state *st, int *map)
{
    if (int rc = check_state(st))
        return rc;
    if (!map)
        return set_err(FQ_ERR_ARG, "null output pointer");
    for (int p = 0; p < st->N; ++p)
        map[p] = st->qmap[p];
    return FQ_OK;
}

extern "C" int fq_state_copy_out(fq_state *st, void *host, size_t bytes, int logical)
{
    if (int rc = check_state(st))
        return rc;
    if (!host || bytes != st->bytes)
        return set_err(FQ_ERR_ARG, "copy_out needs exactly %zu bytes", st->bytes);
    bool ident = true;
    for (int p = 0; p < st->N; ++p)
        ident = ident && st->qmap[p] == p;
    if (!logical || ident) {
        FQ_HIP(hipMemcpyAsync(host, st->sv, bytes, hipMemcpyDeviceToHost, st->stream));
        FQ_HIP(hipStreamSynchronize(st->stream));
        return FQ_OK;
    }
    std::vector<char> tmp(bytes);
    FQ_HIP(hipMemcpyAsync(tmp.data(), st->sv, bytes, hipMemcpyDeviceToHost, st->stream));
    FQ_HIP(hipStreamSynchronize(st->stream));
    // physical -> logical index, one table lookup per index byte
    const int nbytes_idx = (st->N + 7) / 8;
    std::vector<uint64_t> tab((size_t)nbytes_idx * 256, 0);
    for (int k = 0; k < nbytes_idx; ++k)
        for (int v = 0; v < 256; ++v) {
            uint64_t L = 0;
            for (int i = 0; i < 8 && 8 * k + i < st->N; ++i)
                if ((v >> i) & 1)
                    L |= 1ull << st->qmap[8 * k + i];
            tab[(size_t)k * 256 + v] = L;
        }
    const uint64_t n = 1ull << st->N;
    char *dst = (char *)host;
    for (uint64_t b = 0; b < n; ++b) {
        uint64_t L = 0;
        for (int k = 0; k < nbytes_idx; ++k)
            L |= tab[(size_t)k * 256 + ((b >> (8 * k)) & 255)];
        memcpy(dst + L * st->elem, tmp.data() + b * st->elem, st->elem);
    }
    return FQ_OK;
}

extern "C" int fq_state_copy_in(fq_state *st, const void *host, size_t bytes)
{
    if (int rc = check_state(st))
        return rc;
    if (!host || bytes != st->bytes)
        return set_err(FQ_ERR_ARG, "copy_in needs exactly %zu bytes", st->bytes);
    FQ_HIP(hipMemcpyAsync(st->sv, host, bytes, hipMemcpyHostToDevice, st->stream));
    FQ_HIP(hipStreamSynchronize(st->stream));
    return FQ_OK;
}

extern "C" int fq_state_gather(fq_state *st, const uint64_t *idx, size_t n, double *out, int logical)
{
    if (int rc = check_state(st))
        return rc;
    if (n == 0)
        return FQ_OK;
    if (!idx || !out)
        return set_err(FQ_ERR_ARG, "null argument");
    const uint64_t lim = 1ull << st->N;
    std::vector<uint64_t> phys(n);
    for (size_t i = 0; i < n; ++i) {
        if (idx[i] >= lim)
            return set_err(FQ_ERR_ARG, "index %llu out of range", (unsigned long long)idx[i]);
        phys[i] = logical ? to_physical_index(st, idx[i]) : idx[i];
    }
    uint64_t *d_idx = nullptr;
    double *d_out = nullptr;
    if (hipMalloc(&d_idx, sizeof(uint64_t) * n) != hipSuccess ||
        hipMalloc(&d_out, sizeof(double) * 2 * n) != hipSuccess) {
        hipFree(d_idx);
        return set_err(FQ_ERR_OOM, "gather scratch allocation failed");
    }
    int rc = FQ_OK;
    if (hipMemcpyAsync(d_idx, phys.data(), sizeof(uint64_t) * n, hipMemcpyHostToDevice, st->stream) != hipSuccess ||
        launch_gather(st->dtype, st->sv, d_idx, n, d_out, st->stream) != hipSuccess ||
        hipMemcpyAsync(out, d_out, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, st->stream) != hipSuccess ||
        hipStreamSynchronize(st->stream) != hipSuccess)
        rc = set_err(FQ_ERR_DEVICE, "gather failed");
    hipFree(d_idx);
    hipFree(d_out);
    return rc;
}

extern "C" int fq_state_clone(fq_state *dst, const fq_state *src)
{
    if (int rc = check_state(dst))
        return rc;
    if (int rc = check_state(src))
        return rc;
    if (dst->N != src->N || dst->dtype != src->dtype)
        return set_err(FQ_ERR_ARG, "clone needs equal N and dtype");
    if (dst == src)
        return FQ_OK;
    // the copy runs on dst's stream after src's queued work; src's later work
    // waits for the copy (it must not overwrite src while it is read)
    hipEvent_t ev = nullptr;
    FQ_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    hipError_t e = hipEventRecord(ev, src->stream);
    if (e == hipSuccess)
        e = hipStreamWaitEvent(dst->stream, ev, 0);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dst->sv, src->sv, src->bytes, hipMemcpyDeviceToDevice, dst->stream);
    if (e == hipSuccess)
        e = hipEventRecord(ev, dst->stream);
    if (e == hipSuccess)
        e = hipStreamWaitEvent(src->stream, ev, 0);
    hipEventDestroy(ev);
    if (e != hipSuccess)
        return set_err(FQ_ERR_DEVICE, "clone: %s", hipGetErrorString(e));
    for (int p = 0; p < FQ_MAXN; ++p)
        dst->qmap[p] = src->qmap[p];
    return FQ_OK;
}

extern "C" int fq_sync(fq_state *st)
{
    if (int rc = check_state(st))
        return rc;
    FQ_HIP(hipStreamSynchronize(st->stream));
    return FQ_OK;
}

extern "C" int fq_timing_enable(fq_state *st, int on)
{
    if (int rc = check_state(st))
        return rc;
    st->timing = on != 0;
    return FQ_OK;
}

extern "C" int fq_timing_get(fq_state *st, int kind, double *ms, int64_t *count, double *bytes)
{
    if (int rc = check_state(st))
        return rc;
    if (kind < 0 || kind >= FQ_K_COUNT)
        return set_err(FQ_ERR_ARG, "unknown kernel class %d", kind);
    if (int rc = timing_collect(st))
        return rc;
    if (ms)
        *ms = st->t_ms[kind];
    if (count)
        *count = st->t_cnt[kind];
    if (bytes)
        *bytes = st->t_bytes[kind];
    return FQ_OK;
}

extern "C" int fq_stream_probe(size_t bytes, int reps, double *inplace_gbs, double *read_gbs,
                               double *write_gbs)
{
    if (bytes < ((size_t)1 << 20) || bytes % 8 || reps < 1 || !inplace_gbs || !read_gbs || !write_gbs)
        return set_err(FQ_ERR_ARG, "stream probe needs bytes >= 1 MiB (a multiple of 8), reps >= 1 and outputs");
    if (!budget_take(bytes))
        return set_err(FQ_ERR_BUDGET, "stream probe of %zu bytes exceeds the memory budget", bytes);
    void *buf = nullptr;
    float *sink = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double out[3] = {0, 0, 0};
    int rc = FQ_OK;
    hipError_t e = hipMalloc(&buf, bytes);
    if (e == hipSuccess)
        e = hipMalloc(&sink, sizeof(float));
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipEventCreate(&e0);
    if (e == hipSuccess)
        e = hipEventCreate(&e1);
    if (e == hipSuccess)
        e = launch_stream(buf, bytes, 2, sink, s);  // first touch
    // modes 0 (in place), 1 (read), 2 (fill): traffic 2, 1, 1 x bytes
    for (int mode = 0; mode < 3 && e == hipSuccess; ++mode) {
        e = hipEventRecord(e0, s);
        for (int r = 0; r < reps && e == hipSuccess; ++r)
            e = launch_stream(buf, bytes, mode, sink, s);
        if (e == hipSuccess)
            e = hipEventRecord(e1, s);
        if (e == hipSuccess)
            e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess)
            e = hipEventElapsedTime(&ms, e0, e1);
        if (e == hipSuccess && ms > 0.0f)
            out[mode] = (mode == 0 ? 2.0 : 1.0) * (double)bytes * reps / (ms * 1e-3) / 1e9;
    }
    if (e != hipSuccess)
        rc = set_err(FQ_ERR_DEVICE, "stream probe: %s", hipGetErrorString(e));
    if (s)
        hipStreamSynchronize(s);
    if (e0)
        hipEventDestroy(e0);
    if (e1)
        hipEventDestroy(e1);
    if (s)
        hipStreamDestroy(s);
    hipFree(sink);
    hipFree(buf);
    budget_give(bytes);
    if (rc)
        return rc;
    *inplace_gbs = out[0];
    *read_gbs = out[1];
    *write_gbs = out[2];
    return FQ_OK;
}

extern "C" int fq_timing_reset(fq_state *st)
{
    if (int rc = check_state(st))
        return rc;
    if (int rc = timing_collect(st))
        return rc;
    for (int k = 0; k < FQ_K_COUNT; ++k) {
        st->t_ms[k] = 0;
        st->t_cnt[k] = 0;
        st->t_bytes[k] = 0;
    }
    return FQ_OK;
}

// ---- packed slabs (continuous batching) ------------------------------------------------

struct SlotHost {
    fq_graph *g = nullptr;
    int P = 0, layer = 0;
    int phase = 0;  // 0 free, 1 active, 2 finished (result pending collection)
    std::vector<double> gammas, betas;
    double amp_scale = 1.0;   // rx_norm_scale of the request
    PlanDev plan{}, eplan{};  // cost and <C> plans of g, fetched at assignment
    uint64_t gversion = 0;    // g->version the plans belong to
    uint64_t gepoch = 0;      // g->plan_epoch the plan pointers belong to
    int has_rr = 0;
};

// Launch arguments of one slab tile pass (fixed for the slab's lifetime).
struct SlabPass {
    TileArgs args;
    int a_log = 1, slot_shift = 0;
    bool mirror = false;
};

// Layers per step of a new slab of large slots (fq_slab_set_layers_per_step):
// FQ_SLAB_LAYERS (environment, read per slab creation), default 1.
static int slab_layers_default()
{
    const char *e = getenv("FQ_SLAB_LAYERS");
    const int v = e ? atoi(e) : 1;
    return v >= 1 ? v : 1;
}

struct fq_slab {
    int N = 0, K = 0, s = 0, dtype = 0;
    bool z2 = false;  // slots hold the half state z_{N-1} = 0
    size_t elem = 8, slot_bytes = 0, bytes = 0;
    void *sv = nullptr;
    hipStream_t stream = nullptr;
    std::vector<SlotHost> slots;
    SlotParams *d_params = nullptr, *h_params = nullptr;
    double2 *d_tabs = nullptr;
    size_t tab_stride = 0;
    double *d_part = nullptr;
    double *d_res = nullptr, *h_res = nullptr;
    hipEvent_t done = nullptr;
    bool pending = false;  // a step was enqueued and not yet observed complete
    double init_amp = 0;
    std::vector<FusedPass> passes;
    // launch arguments built once at creation
    TileArgs cost_args{};      // phase tables: pass-1 geometry of the N-qubit state
    Geom cost_geom{}, exp_geom{};
    std::vector<SlabPass> pass_args;
    TileArgs exp_args{};       // <C> over the virtual state
    int exp_alog = 1, exp_shift = 0;
    // large slots: each slot runs the single-state tile passes of fq_run
    // (pipelined layers) over its own 2^N (or 2^(N-1)) amplitudes instead of
    // one packed launch; views[k] is slot k seen as a state (its amplitudes,
    // the slab's stream, its own phase-table buffer)
    bool per_slot = false;
    LayerPlan slot_plan;
    std::vector<fq_state *> views;
    int layers_per_step = 1;   // per_slot slabs: layers a step runs per slot
};

static void slab_free(fq_slab *sb)
{
    if (sb->stream)
        hipStreamSynchronize(sb->stream);
    for (fq_state *v : sb->views) {
        hipFree(v->d_phase);
        delete v;
    }
    sb->views.clear();
    if (sb->sv)
        budget_give(sb->bytes);
    hipFree(sb->sv);
    hipFree(sb->d_params);
    hipHostFree(sb->h_params);
    hipFree(sb->d_tabs);
    hipFree(sb->d_part);
    hipFree(sb->d_res);
    hipHostFree(sb->h_res);
    if (sb->done)
        hipEventDestroy(sb->done);
    if (sb->stream) {
        own_stream(sb->stream, false);
        hipStreamDestroy(sb->stream);
    }
    delete sb;
}

// Reduction tile width of a slab's <C> (a half-state slab reduces its N-1 stored qubits).
static int slab_qe(const fq_slab *sb) { return std::min(12, sb->z2 ? sb->N - 1 : sb->N); }

// Phase-table arguments of a slab's cost pass: the N-qubit geometry of pass 1
// (a half-state slab uses its lower half of the tiles).
static TileArgs slab_cost_args(const fq_slab *sb, Geom *gout)
{
    const int Q1 = (int)sb->passes[0].tile.size();
    Geom g1 = make_geom(sb->N, sb->passes[0].tile, slab_a_log(sb->dtype, Q1));
    TileArgs a1;
    geom_to_args(g1, a1);
    if (sb->z2)
        a1.ntiles_run = 1ull << (g1.nh - 1);
    if (gout)
        *gout = g1;
    return a1;
}

// Tile-pass and <C> launch arguments of a slab (its virtual state: the Ns
// stored qubits, then the slot bits; the mirror pass of a half-state slab
// keeps qubit N-1 as a position above them).
static int slab_build_args(fq_slab *sb)
{
    const int N = sb->N;
    const int Ns = sb->z2 ? N - 1 : N;
    sb->pass_args.clear();
    for (size_t i = 0; i < sb->passes.size(); ++i) {
        const FusedPass &ps = sb->passes[i];
        const int Q = (int)ps.tile.size();
        SlabPass sp;
        sp.a_log = slab_a_log(sb->dtype, Q);
        sp.mirror = sb->z2 && i + 1 == sb->passes.size();
        std::vector<int> tile = ps.tile;
        if (sp.mirror)
            tile.back() = Ns + sb->s;
        Geom gv = make_geom(Ns + sb->s + (sp.mirror ? 1 : 0), tile, sp.a_log);
        if (geom_runs(gv) > FQ_MAX_RUNS)
            return set_err(FQ_ERR_ARG, "internal: slab tile needs too many bit runs");
        geom_to_args(gv, sp.args);
        uint32_t gm = 0;
        for (int q : ps.gates)
            for (int ti = 0; ti < Q; ++ti)
                if (ps.tile[ti] == q)
                    gm |= 1u << ti;
        if (int rc = set_gates(sp.args, gv, gm))
            return rc;
        if (sp.mirror)
            for (int ti = 0; ti + 1 < Q; ++ti)
                sp.args.mirror_pm += 1ull << tile[ti];
        sp.slot_shift = sp.mirror ? N - Q : Ns - Q;
        sb->pass_args.push_back(sp);
    }
    const int Qe = slab_qe(sb);
    sb->exp_alog = std::max(1, std::min(4, Qe));
    sb->exp_geom = make_geom(N, low_tile(Qe), sb->exp_alog);
    geom_to_args(make_geom(Ns + sb->s, low_tile(Qe), sb->exp_alog), sb->exp_args);
    sb->exp_shift = Ns - Qe;
    // at least 1024 tiles per slot in the first pass: the persistent
    // single-state tile kernel per slot (register prefetch, compile-time tile
    // widths) fills the device on its own
    const int Q1 = (int)sb->passes[0].tile.size();
    sb->per_slot = Ns - Q1 >= 10;
    if (sb->per_slot)
        sb->slot_plan = layer_plan(N, sb->dtype, sb->z2);
    return FQ_OK;
}

// Per-slot state views of a per-slot slab (after its amplitudes and stream
// exist), each with a phase-table buffer for the largest cost pass of the
// layer plan, so no step allocates.
static int slab_make_views(fq_slab *sb)
{
    const LayerPlan &lp = sb->slot_plan;
    const int m = (int)lp.ps.size();
    size_t need = 0;
    for (int i = 0; i < m; ++i) {
        if (i != 0 && i != lp.end1)
            continue;
        // pass 0 runs as the write-only layer-0 pass too, possibly with
        // another register width (tile_a_log)
        for (int first = 0; first <= (i == 0 ? 1 : 0); ++first) {
            TileArgs a;
            Geom geom;
            const int z2 = !sb->z2 ? Z2_OFF : i + 1 < m ? Z2_LOWER : Z2_MIRROR;
            if (int rc = pass_args(sb->N, sb->dtype, lp.ps[i].tile, lp.ps[i].gates, true, z2, a, geom, 0, first != 0))
                return rc;
            need = std::max(need, phase_table_entries(a, geom.a_log) * sizeof(double2));
        }
    }
    for (int k = 0; k < sb->K; ++k) {
        fq_state *v = new fq_state();
        v->N = sb->N;
        v->dtype = sb->dtype;
        v->elem = sb->elem;
        v->bytes = sb->slot_bytes;
        v->sv = (char *)sb->sv + sb->slot_bytes * (size_t)k;
        v->stream = sb->stream;
        for (int q = 0; q < FQ_MAXN; ++q)
            v->qmap[q] = q;
        v->init_amp = sb->init_amp;
        sb->views.push_back(v);
        if (hipMalloc(&v->d_phase, need) != hipSuccess) {
            hipGetLastError();
            v->d_phase = nullptr;
            return set_err(FQ_ERR_OOM, "slab phase-table allocation failed");
        }
        v->phase_cap = need;
    }
    return FQ_OK;
}

// Cost and <C> plans of a slot's graph (again after the graph changed).
static int slab_slot_plans(fq_slab *sb, SlotHost &h)
{
    if (int rc = get_plan(h.g, sb->cost_geom, &h.plan, sb->stream))
        return rc;
    if (int rc = get_plan(h.g, sb->exp_geom, &h.eplan, sb->stream))
        return rc;
    h.has_rr = 0;
    for (int j = 0; j < FQ_MAX_REGS; ++j)
        if (h.plan.rr[j] != 0)
            h.has_rr = 1;
    h.gversion = h.g->version;
    h.gepoch = h.g->plan_epoch;
    return FQ_OK;
}

static bool slab_slot_stale(const SlotHost &h)
{
    return h.g->version != h.gversion || h.g->plan_epoch != h.gepoch;
}

extern "C" int fq_slab_create(int N, int nslots, int dtype, fq_slab **out)
{
    return fq_slab_create_sched(N, nslots, dtype, FQ_SCHED_FUSED, out);
}

extern "C" int fq_slab_create_sched(int N, int nslots, int dtype, int schedule, fq_slab **out)
{
    if (!out)
        return set_err(FQ_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (dtype != FQ_C128 && dtype != FQ_C64)
        return set_err(FQ_ERR_ARG, "dtype must be FQ_C128 or FQ_C64");
    if (schedule != FQ_SCHED_FUSED && schedule != FQ_SCHED_Z2)
        return set_err(FQ_ERR_ARG, "slab schedule must be FQ_SCHED_FUSED or FQ_SCHED_Z2");
    if (nslots < 1 || (nslots & (nslots - 1)))
        return set_err(FQ_ERR_ARG, "nslots=%d must be a power of two", nslots);
    int s = 0;
    while ((1 << s) < nslots)
        ++s;
    if (N < 1 || N + s > 36)
        return set_err(FQ_ERR_ARG, "N=%d with %d slots out of range", N, nslots);
    const bool z2 = schedule == FQ_SCHED_Z2;
    if (z2 && N < 3)
        return set_err(FQ_ERR_ARG, "the half-state schedule needs N >= 3");
    fq_slab *sb = new fq_slab();
    sb->N = N;
    sb->K = nslots;
    sb->s = s;
    sb->dtype = dtype;
    sb->z2 = z2;
    sb->elem = dtype == FQ_C128 ? 16 : 8;
    sb->slot_bytes = ((size_t)1 << (z2 ? N - 1 : N)) * sb->elem;
    sb->bytes = sb->slot_bytes * nslots;
    sb->slots.resize(nslots);
    sb->init_amp = pow(1. / sqrt(2.), N);
    sb->passes = cover_passes(N, std::min(tile_max_q(dtype), z2 ? N - 1 : N), dtype);
    if (!budget_take(sb->bytes)) {
        delete sb;
        return set_err(FQ_ERR_BUDGET, "slab of %d x 2^%d amplitudes exceeds the memory budget", nslots, N);
    }
    // launch arguments and scratch sizes
    sb->cost_args = slab_cost_args(sb, &sb->cost_geom);
    sb->tab_stride = phase_table_entries(sb->cost_args, sb->cost_geom.a_log);
    const int Qe = slab_qe(sb);
    const size_t etiles = ((size_t)1 << ((z2 ? N - 1 : N) - Qe)) * nslots;
    if (int rc = slab_build_args(sb)) {
        budget_give(sb->bytes);
        delete sb;
        return rc;
    }
    bool ok = hipMalloc(&sb->sv, sb->bytes) == hipSuccess;
    if (!ok) {
        hipGetLastError();
        sb->sv = nullptr;
        budget_give(sb->bytes);
        slab_free(sb);
        return set_err(FQ_ERR_OOM, "slab allocation of %zu bytes failed", (size_t)(((size_t)1 << N) * (dtype == FQ_C128 ? 16 : 8) * nslots));
    }
    ok = hipMalloc(&sb->d_params, sizeof(SlotParams) * nslots) == hipSuccess &&
         hipHostMalloc(&sb->h_params, sizeof(SlotParams) * nslots, hipHostMallocDefault) == hipSuccess &&
         hipMalloc(&sb->d_tabs, sizeof(double2) * sb->tab_stride * nslots) == hipSuccess &&
         hipMalloc(&sb->d_part, sizeof(double) * 2 * etiles) == hipSuccess &&
         hipMalloc(&sb->d_res, sizeof(double) * 2 * nslots) == hipSuccess &&
         hipHostMalloc(&sb->h_res, sizeof(double) * 2 * nslots, hipHostMallocDefault) == hipSuccess &&
         hipStreamCreateWithFlags(&sb->stream, hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&sb->done, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        hipGetLastError();
        slab_free(sb);
        return set_err(FQ_ERR_OOM, "slab scratch allocation failed");
    }
    own_stream(sb->stream, true);
    sb->layers_per_step = slab_layers_default();
    if (sb->per_slot)
        if (int rc = slab_make_views(sb)) {
            slab_free(sb);
            return rc;
        }
    memset(sb->h_params, 0, sizeof(SlotParams) * nslots);
    *out = sb;
    return FQ_OK;
}

extern "C" int fq_slab_destroy(fq_slab *sb)
{
    if (sb)
        slab_free(sb);
    return FQ_OK;
}

extern "C" int fq_slab_info(const fq_slab *sb, int *N, int *nslots, int *free_slots, int *active)
{
    if (!sb)
        return set_err(FQ_ERR_ARG, "null slab");
    int fr = 0, ac = 0;
    for (const SlotHost &h : sb->slots) {
        fr += h.phase == 0;
        ac += h.phase == 1;
    }
    if (N)
        *N = sb->N;
    if (nslots)
        *nslots = sb->K;
    if (free_slots)
        *free_slots = fr;
    if (active)
        *active = ac;
    return FQ_OK;
}

extern "C" int fq_slab_assign(fq_slab *sb, int slot, fq_graph *g, int P, const double *gammas,
                              const double *betas)
{
    if (!sb || !g)
        return set_err(FQ_ERR_ARG, "null slab or graph");
    if (slot < 0 || slot >= sb->K)
        return set_err(FQ_ERR_ARG, "slot %d out of range", slot);
    if (g->N != sb->N)
        return set_err(FQ_ERR_ARG, "graph has N=%d, slab has N=%d", g->N, sb->N);
    if (P < 1 || !gammas || !betas)
        return set_err(FQ_ERR_ARG, "bad layer count or angle arrays");
    for (int p = 0; p < P; ++p)
        if (!std::isfinite(gammas[p]) || !std::isfinite(betas[p]))
            return set_err(FQ_ERR_ARG, "angles must be finite");
    SlotHost &h = sb->slots[slot];
    if (h.phase != 0)
        return set_err(FQ_ERR_ARG, "slot %d is not free", slot);
    h.g = g;
    if (int rc = slab_slot_plans(sb, h)) {
        h.g = nullptr;
        return rc;
    }
    h.P = P;
    h.layer = 0;
    h.phase = 1;
    h.gammas.assign(gammas, gammas + P);
    h.betas.assign(betas, betas + P);
    h.amp_scale = rx_norm_scale(sb->N, P, betas, sb->dtype);
    return FQ_OK;
}

extern "C" int fq_slab_poll(fq_slab *sb, int *busy)
{
    if (!sb || !busy)
        return set_err(FQ_ERR_ARG, "null argument");
    if (!sb->pending) {
        *busy = 0;
        return FQ_OK;
    }
    hipError_t e = hipEventQuery(sb->done);
    if (e == hipErrorNotReady) {
        *busy = 1;
        return FQ_OK;
    }
    if (e != hipSuccess)
        return set_err(FQ_ERR_DEVICE, "slab step failed: %s", hipGetErrorString(e));
    sb->pending = false;
    *busy = 0;
    return FQ_OK;
}

// Advance every active slot by one layer: one packed launch per tile pass.
// Slots that complete their last layer also get <C> reduced; their results
// are read with fq_slab_collect once fq_slab_poll reports the step done.
extern "C" int fq_slab_step(fq_slab *sb, int *nfinishing)
{
    if (!sb)
        return set_err(FQ_ERR_ARG, "null slab");
    if (sb->pending) {
        FQ_HIP(hipEventSynchronize(sb->done));
        sb->pending = false;
    }
    const int K = sb->K;
    // refresh plans of changed graphs; a refresh can release a graph's plan
    // cache, so repeat until every active slot's plans are current
    for (bool again = true; again;) {
        again = false;
        for (int k = 0; k < K; ++k) {
            SlotHost &h = sb->slots[k];
            if (h.phase == 1 && slab_slot_stale(h)) {
                if (int rc = slab_slot_plans(sb, h))
                    return rc;
                again = true;
            }
        }
    }
    // layers of this step per slot (per_slot slabs may run several)
    std::vector<int> nl(K, 0);
    int nact = 0, nfin = 0;
    for (int k = 0; k < K; ++k) {
        SlotHost &h = sb->slots[k];
        SlotParams &sp = sb->h_params[k];
        memset(&sp, 0, sizeof(sp));
        if (h.phase != 1)
            continue;
        nl[k] = sb->per_slot ? std::min(sb->layers_per_step, h.P - h.layer) : 1;
        const int L = h.layer;
        rx_coeffs(h.betas[L], sb->dtype, &sp.rx_c, &sp.rx_s);
        sp.gamma = h.gammas[L];
        // the packed kernels apply RX with rx_coeffs (norm factor amp_scale);
        // per-slot slots run fq_run's passes, whose RX factors are exact
        sp.init_amp = sb->init_amp * (sb->per_slot ? 1.0 : h.amp_scale);
        sp.active = 1;
        sp.cost_first = L == 0;
        sp.want_expect = L + nl[k] == h.P;
        sp.ptab = sb->d_tabs + sb->tab_stride * (size_t)k;
        sp.plan = h.plan;
        sp.eplan = h.eplan;
        sp.has_rr = h.has_rr;
        ++nact;
        nfin += sp.want_expect;
    }
    if (nfinishing)
        *nfinishing = nfin;
    if (nact == 0)
        return FQ_OK;
    FQ_HIP(hipMemcpyAsync(sb->d_params, sb->h_params, sizeof(SlotParams) * K, hipMemcpyHostToDevice,
                          sb->stream));
    if (sb->per_slot) {
        // each slot nl[k] layers of fq_run's pipelined schedule on its own
        // amplitudes (the slots' layers interleaved on the slab's stream)
        const int most = *std::max_element(nl.begin(), nl.end());
        for (int rep = 0; rep < most; ++rep) {
            for (int k = 0; k < K; ++k) {
                const SlotParams &sp = sb->h_params[k];
                if (!sp.active || rep >= nl[k])
                    continue;
                SlotHost &h = sb->slots[k];
                fq_state *v = sb->views[k];
                const int L = h.layer + rep;
                v->init_amp = sp.init_amp;
                if (L == 0) {
                    v->pending_re = 1.0;
                    v->pending_im = 0.0;
                }
                v->in_run = true;   // a slot runs fq_run's layers on a QAOA state
                if (int rc = run_layer(v, h.g, sb->slot_plan, L, h.P, h.gammas.data(), h.betas.data(), sb->z2,
                                       false))
                    return rc;
            }
        }
    } else {
        FQ_HIP(launch_slab_phase_table(sb->cost_args, sb->cost_geom.a_log, sb->d_params, K, sb->d_tabs,
                                       sb->tab_stride, sb->stream));
        for (size_t i = 0; i < sb->pass_args.size(); ++i) {
            const SlabPass &sp = sb->pass_args[i];
            FQ_HIP(launch_slab_pass(sb->dtype, sp.a_log, i == 0, sp.mirror, sb->sv, sp.args, sb->d_params,
                                    sp.slot_shift, sb->stream));
        }
    }
    if (nfin > 0) {
        FQ_HIP(launch_slab_expect(sb->dtype, sb->exp_alog, sb->sv, sb->exp_args, sb->d_params, sb->exp_shift, K,
                                  sb->d_part, sb->d_res, sb->stream));
        FQ_HIP(hipMemcpyAsync(sb->h_res, sb->d_res, sizeof(double) * 2 * K, hipMemcpyDeviceToHost, sb->stream));
    }
    FQ_HIP(hipEventRecord(sb->done, sb->stream));
    sb->pending = true;
    for (int k = 0; k < K; ++k) {
        SlotHost &h = sb->slots[k];
        if (h.phase != 1)
            continue;
        h.layer += nl[k];
        if (h.layer == h.P)
            h.phase = 2;
    }
    return FQ_OK;
}

extern "C" int fq_slab_set_layers_per_step(fq_slab *sb, int layers)
{
    if (!sb)
        return set_err(FQ_ERR_ARG, "null slab");
    if (layers < 1)
        return set_err(FQ_ERR_ARG, "layers per step must be >= 1");
    sb->layers_per_step = layers;
    return FQ_OK;
}

// Results of the slots that finished: slot index, <C>, sum |psi|^2.  Waits
// for the enqueued step, then frees those slots.  n receives the count
// (capacity: the slab's slot count).
extern "C" int fq_slab_collect(fq_slab *sb, int *slots, double *expect, double *norm2, int *n)
{
    if (!sb || !n)
        return set_err(FQ_ERR_ARG, "null argument");
    if (sb->pending) {
        FQ_HIP(hipEventSynchronize(sb->done));
        sb->pending = false;
    }
    int c = 0;
    for (int k = 0; k < sb->K; ++k) {
        SlotHost &h = sb->slots[k];
        if (h.phase != 2)
            continue;
        if (slots)
            slots[c] = k;
        // a half-state slab holds one of the two equal halves
        const double f = sb->z2 ? 2.0 : 1.0;
        if (expect)
            expect[c] = f * sb->h_res[2 * k];
        if (norm2)
            norm2[c] = f * sb->h_res[2 * k + 1];
        ++c;
        h.phase = 0;
        h.g = nullptr;
    }
    *n = c;
    return FQ_OK;
}

// Amplitudes of one slot (identity layout) to host memory.  Waits.
extern "C" int fq_slab_copy_out(fq_slab *sb, int slot, void *host, size_t bytes)
{
    if (!sb || !host)
        return set_err(FQ_ERR_ARG, "null argument");
    const size_t full = ((size_t)1 << sb->N) * sb->elem;
    if (slot < 0 || slot >= sb->K || bytes != full)
        return set_err(FQ_ERR_ARG, "bad slot or size (%zu bytes expected)", full);
    FQ_HIP(hipMemcpyAsync(host, (char *)sb->sv + sb->slot_bytes * (size_t)slot, sb->slot_bytes,
                          hipMemcpyDeviceToHost, sb->stream));
    FQ_HIP(hipStreamSynchronize(sb->stream));
    sb->pending = false;
    if (sb->z2) {
        // psi(b) = psi(~b): the upper half is the stored half reversed
        const size_t half = (size_t)1 << (sb->N - 1), e = sb->elem;
        char *h = (char *)host;
        for (size_t i = 0; i < half; ++i)
            memcpy(h + (half + i) * e, h + (half - 1 - i) * e, e);
    }
    return FQ_OK;
}
