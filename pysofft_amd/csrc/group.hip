// Grouped mixed-precision transforms of uneven bandwidths: one persistent DWT launch
// over the ragged (pair, transform, chunk) work of every bandwidth, per-bandwidth FFT
// passes as parallel branches, the whole sequence captured once into a hipGraph per
// (input, output) pair and replayed, and launch shapes (transforms per wave, waves per
// block, grid) picked by timing the candidates on the group's own shapes.
//
// Packed layout (the library's choice, reported by so3ft_group_layout): transforms of
// one bandwidth are contiguous, bandwidths in decreasing order, each bandwidth's
// transforms in the caller's order; complex64 coefficients (NC(bw) each, mnl order)
// and complex64 SO(3) functions ((2bw)^3 each, C-order (gamma, alpha, beta)).
#include <algorithm>
#include <functional>
#include <cstdio>
#include <map>
#include <mutex>
#include <numeric>
#include <vector>

#include "../../include/so3ft.h"
#include "so3ft_internal.h"

namespace {

using namespace so3;

#define GCHK(expr)                                                                                  \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return set_error(SO3FT_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct Seg {
    int64_t bw = 0, N = 0, T = 0, nc = 0, coff = 0, goff = 0;
    so3ft_handle plan = 0;
    DeviceTables tb;
    std::vector<int64_t> members;  // caller indices, in order
};

struct Config {
    int inv_tb = 2, inv_grid = 0;          // grid 0: one block per work item
    int fwd_tb = 4, fwd_nw = 1, fwd_grid = 0;
};

struct Items {
    int4 *d = nullptr;
    int n = 0;
};

struct Group {
    int device = 0;
    int64_t n = 0, ctot = 0, gtot = 0, max_bw = 0;
    std::vector<int64_t> coff, goff;
    std::vector<Seg> segs;
    MixSegment *dsegs = nullptr;
    float2 *F = nullptr;                   // DWT <-> FFT intermediate, gtot entries
    Config cfg;
    bool tuned = false, graphs = true;
    Items inv, fwd;                        // work lists over every segment
    std::vector<Items> inv_seg, fwd_seg;   // one list per segment
    std::vector<int> fwd_seg_tb;           // forward transforms per wave of each segment's list
    // schedule per direction ([0] forward, [1] inverse): 0 = one grouped DWT launch and the
    // segments' FFT passes as parallel branches; 1 = one branch per segment holding its DWT
    // and FFT passes (a segment's FFT overlaps another's DWT); 2 = the segments one after
    // another on one stream through the single-bandwidth kernels; -1 = choose by timing
    int sched[2] = {-1, -1};
    float sched_ms[2][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};   // measured graph time per schedule
    hipStream_t cap = nullptr;
    std::vector<hipStream_t> side;
    hipEvent_t ev_fork = nullptr;
    std::vector<hipEvent_t> ev_join;
    // captured graphs per (input, output) pair, each with an event recorded after its last
    // launch (an evicted graph is updated only once that launch has finished)
    struct GExec {
        hipGraphExec_t ex = nullptr;
        hipEvent_t done = nullptr;
    };
    std::map<std::pair<const void *, void *>, GExec> ginv, gfwd;
};

std::mutex g_mu;
std::map<int64_t, Group *> g_groups;
int64_t g_next = 1;
// tuned configurations by (device, bandwidth multiset)
std::map<std::pair<int, std::vector<std::pair<int64_t, int64_t>>>, Config> g_tuned;

Group *get(int64_t h)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_groups.find(h);
    return it == g_groups.end() ? nullptr : it->second;
}

inline int64_t ncoef(int64_t bw) { return (4 * bw * bw * bw - bw) / 3; }

int check_bws(const int64_t *bws, int64_t n)
{
    if (!bws || n < 1 || n > (1 << 20)) return set_error(SO3FT_EINVAL, "need 1 <= n <= 2^20 bandwidths");
    for (int64_t i = 0; i < n; ++i)
        if (bws[i] < 2 || bws[i] > 256)
            return set_error(SO3FT_EUNSUPPORTED, "grouped transforms need 2 <= bw <= 256, got " +
                                                     std::to_string(bws[i]));
    return SO3FT_OK;
}

// segments (decreasing bandwidth) and per-transform offsets
void layout(const int64_t *bws, int64_t n, std::vector<Seg> &segs, std::vector<int64_t> &coff,
            std::vector<int64_t> &goff, int64_t &ctot, int64_t &gtot)
{
    std::vector<int64_t> distinct(bws, bws + n);
    std::sort(distinct.begin(), distinct.end(), std::greater<int64_t>());
    distinct.erase(std::unique(distinct.begin(), distinct.end()), distinct.end());
    segs.assign(distinct.size(), Seg());
    coff.assign(n, 0);
    goff.assign(n, 0);
    for (size_t s = 0; s < distinct.size(); ++s) {
        segs[s].bw = distinct[s];
        segs[s].N = 2 * distinct[s];
        segs[s].nc = ncoef(distinct[s]);
    }
    for (int64_t i = 0; i < n; ++i) {
        const size_t s = std::find(distinct.begin(), distinct.end(), bws[i]) - distinct.begin();
        segs[s].members.push_back(i);
    }
    ctot = gtot = 0;
    for (auto &sg : segs) {
        sg.T = (int64_t)sg.members.size();
        sg.coff = ctot;
        sg.goff = gtot;
        for (int64_t k = 0; k < sg.T; ++k) {
            coff[sg.members[k]] = ctot + k * sg.nc;
            goff[sg.members[k]] = gtot + k * sg.N * sg.N * sg.N;
        }
        ctot += sg.T * sg.nc;
        gtot += sg.T * sg.N * sg.N * sg.N;
    }
}

// Work list of one direction.  Units (segment, order pair, transform group) of one
// segment are sorted by decreasing degree count and bundled eight at a time; a bundle
// is emitted chunk by chunk, eight items per chunk (padding where a unit has fewer
// chunks), so the chunks of one unit sit at list positions that agree mod 8 and, with a
// grid that is a multiple of 8, run on one XCD (shared L2 for the unit's operands).
// Bundles of all segments are ordered by decreasing cost (largest first for balance).
int build_items(const Group &G, bool inverse, int TB, int NW, std::vector<int4> &out, int only_seg = -1)
{
    struct Unit { int pq, L, t0; };
    struct Bundle { double cost; int seg; std::vector<Unit> u; int nch; };
    std::vector<Bundle> bundles;
    const int per = inverse ? TB : TB * NW;
    for (int s = 0; s < (int)G.segs.size(); ++s) {
        if (only_seg >= 0 && s != only_seg) continue;
        const Seg &sg = G.segs[s];
        const int bw = (int)sg.bw, N = (int)sg.N;
        std::vector<Unit> units;
        for (int q = 0; q < bw; ++q)
            for (int p = 0; p <= q; ++p)
                for (int t0 = 0; t0 < (int)sg.T; t0 += per) units.push_back({p | (q << 16), bw - q, t0});
        std::stable_sort(units.begin(), units.end(), [](const Unit &a, const Unit &b) { return a.L > b.L; });
        for (size_t i = 0; i < units.size(); i += 8) {
            Bundle b;
            b.seg = s;
            b.u.assign(units.begin() + i, units.begin() + std::min(units.size(), i + 8));
            const int L = b.u[0].L;
            if (inverse) {
                b.nch = (N + 63) / 64;
                b.cost = (double)(L + 8);
            } else {   // one item per unit: every beta and degree
                b.nch = 1;
                b.cost = (double)N * (L + 8);
            }
            bundles.push_back(std::move(b));
        }
    }
    std::stable_sort(bundles.begin(), bundles.end(), [](const Bundle &a, const Bundle &b) { return a.cost > b.cost; });
    out.clear();
    for (const Bundle &b : bundles)
        for (int c = 0; c < b.nch; ++c)
            for (int k = 0; k < 8; ++k) {
                if (k < (int)b.u.size())
                    out.push_back(make_int4(b.seg, b.u[k].pq, b.u[k].t0, c));
                else
                    out.push_back(make_int4(-1, 0, 0, 0));
            }
    if (out.size() > (size_t)(1u << 30)) return set_error(SO3FT_EINVAL, "grouped work list too long; split the batch");
    // bounds of every item against what the kernels index
    for (const int4 &w : out) {
        if (w.x < 0) continue;
        if (w.x >= (int)G.segs.size()) return set_error(SO3FT_EINVAL, "internal: grouped work item out of range");
        const Seg &sg = G.segs[w.x];
        const int p = w.y & 0xffff, q = w.y >> 16;
        const bool ok = p <= q && q < sg.bw && w.z >= 0 && w.z < sg.T && w.w >= 0 &&
                        (inverse ? 64 * w.w < sg.N : w.w == 0);
        if (!ok) return set_error(SO3FT_EINVAL, "internal: grouped work item out of range");
    }
    return SO3FT_OK;
}

int upload_items(const std::vector<int4> &v, Items &it)
{
    hipFree(it.d);
    it.d = nullptr;
    it.n = 0;
    if (v.empty()) return SO3FT_OK;
    GCHK(hipMalloc((void **)&it.d, sizeof(int4) * v.size()));
    GCHK(hipMemcpy(it.d, v.data(), sizeof(int4) * v.size(), hipMemcpyHostToDevice));
    it.n = (int)v.size();
    return SO3FT_OK;
}

int set_items(Group &G)
{
    std::vector<int4> v;
    if (int rc = build_items(G, true, G.cfg.inv_tb, 1, v)) return rc;
    if (int rc = upload_items(v, G.inv)) return rc;
    if (int rc = build_items(G, false, G.cfg.fwd_tb, G.cfg.fwd_nw, v)) return rc;
    if (int rc = upload_items(v, G.fwd)) return rc;
    G.inv_seg.resize(G.segs.size());
    G.fwd_seg.resize(G.segs.size());
    G.fwd_seg_tb.assign(G.segs.size(), G.cfg.fwd_tb);
    for (int k = 0; k < (int)G.segs.size(); ++k) {
        if (int rc = build_items(G, true, G.cfg.inv_tb, 1, v, k)) return rc;
        if (int rc = upload_items(v, G.inv_seg[k])) return rc;
        // a segment's own forward launch takes four transforms per wave from bw = 64 (as the
        // single-bandwidth path), else two
        const int tb = G.segs[k].bw >= 64 ? 4 : 2;
        G.fwd_seg_tb[k] = tb;
        if (int rc = build_items(G, false, tb, 1, v, k)) return rc;
        if (int rc = upload_items(v, G.fwd_seg[k])) return rc;
    }
    return SO3FT_OK;
}

void drop_graphs(Group &G)
{
    for (auto *c : {&G.ginv, &G.gfwd})
        for (auto &kv : *c) {
            if (kv.second.ex) hipGraphExecDestroy(kv.second.ex);
            if (kv.second.done) hipEventDestroy(kv.second.done);
        }
    G.ginv.clear();
    G.gfwd.clear();
}

void free_group(Group *G)
{
    drop_graphs(*G);
    for (auto &sg : G->segs)
        if (sg.plan) so3ft_destroy(sg.plan);
    hipFree(G->dsegs);
    hipFree(G->F);
    hipFree(G->inv.d);
    hipFree(G->fwd.d);
    for (auto &it : G->inv_seg) hipFree(it.d);
    for (auto &it : G->fwd_seg) hipFree(it.d);
    for (auto s : G->side) hipStreamDestroy(s);
    if (G->cap) hipStreamDestroy(G->cap);
    if (G->ev_fork) hipEventDestroy(G->ev_fork);
    for (auto e : G->ev_join) hipEventDestroy(e);
    delete G;
}

inline int grid_of(int g, int nitem) { return g > 0 ? g : std::max(1, nitem); }

int launch_inv_dwt(Group &G, const Config &c, const Items &it, const float2 *coeff, hipStream_t s)
{
    GCHK(launch_dwt_inverse_mixed_grouped(c.inv_tb, coeff, G.F, G.dsegs, it.d, it.n, grid_of(c.inv_grid, it.n),
                                          (int)G.max_bw, s));
    return SO3FT_OK;
}

int launch_fwd_dwt(Group &G, const Config &c, const Items &it, float2 *coeff, hipStream_t s)
{
    GCHK(launch_dwt_forward_mixed_grouped(c.fwd_tb, G.F, coeff, G.dsegs, it.d, it.n, grid_of(c.fwd_grid, it.n),
                                          (int)G.max_bw, s));
    return SO3FT_OK;
}

// the two FFT passes of one segment (as in the single-bandwidth complex64 pipeline)
int seg_fft(Group &G, const Seg &sg, bool inverse, const float2 *in, float2 *out, hipStream_t s)
{
    const int64_t N = sg.N, N2 = N * N, N3 = N2 * N;
    const int bw = (int)sg.bw, lmax = bw - 1;
    for (int64_t i0 = 0; i0 < sg.T; i0 += 32768) {
        const int T = (int)std::min<int64_t>(32768, sg.T - i0);
        float2 *Fs = G.F + sg.goff + i0 * N3;
        FftAxis a{};
        a.in_t = N3; a.in_o = N2; a.in_r = N; a.out_t = N3; a.out_o = N2; a.out_r = N;
        FftAxis b{};
        b.in_t = N3; b.in_o = N; b.in_r = N2; b.out_t = N3; b.out_o = N; b.out_r = N2;
        if (inverse) {
            a.mask_rows = 1; a.mask_outer = 1; a.sign = -1; a.scale = 1.0;
            b.mask_rows = 1; b.sign = -1; b.scale = 1.0 / (2.0 * 3.14159265358979323846);
            GCHK(launch_fft_reg((int)N, sg.tb.tw, bw, lmax, Fs, Fs, a, T, 1, 1, 0, (int)N, s));
            GCHK(launch_fft_reg((int)N, sg.tb.tw, bw, lmax, Fs, out + sg.goff + i0 * N3, b, T, 1, 1, 0, (int)N, s));
        } else {
            b.sign = +1; b.scale = 1.0;
            a.mask_outer = 1; a.sign = +1;
            a.scale = 2.0 * 3.14159265358979323846 / ((double)N * (double)N);
            GCHK(launch_fft_reg((int)N, sg.tb.tw, bw, lmax, in + sg.goff + i0 * N3, Fs, b, T, 1, 1, 0, (int)N, s));
            GCHK(launch_fft_reg((int)N, sg.tb.tw, bw, lmax, Fs, Fs, a, T, 1, 1, 0, (int)N, s));
        }
    }
    return SO3FT_OK;
}

int ensure_streams(Group &G)
{
    if (G.cap) return SO3FT_OK;
    GCHK(hipStreamCreateWithFlags(&G.cap, hipStreamNonBlocking));
    GCHK(hipEventCreateWithFlags(&G.ev_fork, hipEventDisableTiming));
    for (size_t i = 1; i < G.segs.size(); ++i) {
        hipStream_t st;
        GCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        G.side.push_back(st);
        hipEvent_t e;
        GCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        G.ev_join.push_back(e);
    }
    return SO3FT_OK;
}

// per-segment FFT passes; with fork, segment i > 0 runs on side stream i-1 (parallel
// branches of a captured graph), joined back into s
int all_fft(Group &G, bool inverse, const float2 *in, float2 *out, hipStream_t s, bool fork)
{
    if (!fork) {
        for (const Seg &sg : G.segs)
            if (int rc = seg_fft(G, sg, inverse, in, out, s)) return rc;
        return SO3FT_OK;
    }
    GCHK(hipEventRecord(G.ev_fork, s));
    for (size_t i = 1; i < G.segs.size(); ++i) GCHK(hipStreamWaitEvent(G.side[i - 1], G.ev_fork, 0));
    if (int rc = seg_fft(G, G.segs[0], inverse, in, out, s)) return rc;
    for (size_t i = 1; i < G.segs.size(); ++i) {
        if (int rc = seg_fft(G, G.segs[i], inverse, in, out, G.side[i - 1])) return rc;
        GCHK(hipEventRecord(G.ev_join[i - 1], G.side[i - 1]));
        GCHK(hipStreamWaitEvent(s, G.ev_join[i - 1], 0));
    }
    return SO3FT_OK;
}

int enqueue(Group &G, bool inverse, const void *in, void *out, hipStream_t s, bool fork)
{
    const float2 *fin = reinterpret_cast<const float2 *>(in);
    float2 *fout = reinterpret_cast<float2 *>(out);
    const int sc = G.sched[inverse ? 1 : 0];
    if (sc == 2) {
        // the segments in sequence on s through the single-bandwidth path's kernels
        // (launch_dwt_*_mixed on each segment's plan tables), as the bandwidths' own
        // isoft_many_c64 / soft_many_c64 calls would run them
        for (const Seg &sg : G.segs) {
            const int64_t N3 = sg.N * sg.N * sg.N;
            const int npair = (int)(sg.bw * (sg.bw + 1) / 2);
            if (!inverse)
                if (int rc = seg_fft(G, sg, false, fin, nullptr, s)) return rc;
            for (int64_t i0 = 0; i0 < sg.T; i0 += 4096) {
                const int T = (int)std::min<int64_t>(4096, sg.T - i0);
                float2 *Fs = G.F + sg.goff + i0 * N3;
                if (inverse)
                    GCHK(launch_dwt_inverse_mixed(sg.tb, (int)sg.bw, (int)sg.N, npair, DWT_COMPLEX,
                                                  fin + sg.coff + i0 * sg.nc, sg.nc, nullptr, Fs, N3, T, s));
                else
                    GCHK(launch_dwt_forward_mixed(sg.tb, (int)sg.bw, (int)sg.N, npair, DWT_COMPLEX, Fs, N3,
                                                  fout + sg.coff + i0 * sg.nc, sg.nc, T, s));
            }
            if (inverse)
                if (int rc = seg_fft(G, sg, true, nullptr, fout, s)) return rc;
        }
        return SO3FT_OK;
    }
    if (sc == 1) {
        // one branch per segment: DWT then FFT (inverse) or FFT then DWT (forward)
        if (fork) {
            GCHK(hipEventRecord(G.ev_fork, s));
            for (size_t i = 1; i < G.segs.size(); ++i) GCHK(hipStreamWaitEvent(G.side[i - 1], G.ev_fork, 0));
        }
        for (size_t i = 0; i < G.segs.size(); ++i) {
            hipStream_t st = (fork && i > 0) ? G.side[i - 1] : s;
            if (inverse) {
                if (int rc = launch_inv_dwt(G, G.cfg, G.inv_seg[i], fin, st)) return rc;
                if (int rc = seg_fft(G, G.segs[i], true, nullptr, fout, st)) return rc;
            } else {
                if (int rc = seg_fft(G, G.segs[i], false, fin, nullptr, st)) return rc;
                Config cs = G.cfg;
                cs.fwd_tb = G.fwd_seg_tb[i];
                if (int rc = launch_fwd_dwt(G, cs, G.fwd_seg[i], fout, st)) return rc;
            }
            if (fork && i > 0) {
                GCHK(hipEventRecord(G.ev_join[i - 1], st));
                GCHK(hipStreamWaitEvent(s, G.ev_join[i - 1], 0));
            }
        }
        return SO3FT_OK;
    }
    if (inverse) {
        if (int rc = launch_inv_dwt(G, G.cfg, G.inv, fin, s)) return rc;
        return all_fft(G, true, nullptr, fout, s, fork);
    }
    if (int rc = all_fft(G, false, fin, nullptr, s, fork)) return rc;
    return launch_fwd_dwt(G, G.cfg, G.fwd, fout, s);
}

// capture the launch sequence for (in, out) and make it executable: into *exec by
// instantiation, or, given a graph already instantiated for other buffers (reuse), by
// updating that one's node arguments in place (the same launch sequence with other pointers
// has the same topology), which avoids a new instantiation; reuse is consumed either way
int capture(Group &G, bool inverse, const void *in, void *out, hipGraphExec_t *exec,
            hipGraphExec_t reuse = nullptr)
{
    if (int rc = ensure_streams(G)) {
        if (reuse) hipGraphExecDestroy(reuse);
        return rc;
    }
    const hipError_t eb = hipStreamBeginCapture(G.cap, hipStreamCaptureModeThreadLocal);
    if (eb != hipSuccess) {
        if (reuse) hipGraphExecDestroy(reuse);
        return set_error(SO3FT_EHIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(eb));
    }
    const int rc = enqueue(G, inverse, in, out, G.cap, true);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(G.cap, &graph);
    if (rc || e != hipSuccess) {
        if (e == hipSuccess && graph) hipGraphDestroy(graph);
        if (reuse) hipGraphExecDestroy(reuse);
        if (rc) return rc;
        return set_error(SO3FT_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    }
    if (reuse) {
        hipGraphNode_t err_node = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateSuccess;
        if (hipGraphExecUpdate(reuse, graph, &err_node, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
            hipGraphDestroy(graph);
            *exec = reuse;
            return SO3FT_OK;
        }
        (void)hipGetLastError();   // not updatable: instantiate a new one
        hipGraphExecDestroy(reuse);
    }
    const hipError_t ei = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ei != hipSuccess) return set_error(SO3FT_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    return SO3FT_OK;
}

// ---- tuner: time every candidate launch shape of each grouped DWT on the group's
// own work (inputs zero-filled scratch; MFMA and recurrence time do not depend on the
// values) and keep the fastest
float time_launches(hipStream_t s, hipEvent_t e0, hipEvent_t e1, const std::function<int()> &launch, int *rc)
{
    *rc = launch();  // warm-up
    if (*rc) return 0.f;
    if (hipEventRecord(e0, s) != hipSuccess) { *rc = set_error(SO3FT_EHIP, "hipEventRecord"); return 0.f; }
    for (int r = 0; r < 2 && !*rc; ++r) *rc = launch();
    if (*rc) return 0.f;
    float ms = 0.f;
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        *rc = set_error(SO3FT_EHIP, "tuner timing failed");
        return 0.f;
    }
    return ms / 2.f;
}

std::vector<int> grid_candidates(int occ)
{
    int cus = 0, dev = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    std::vector<int> g = {0};
    if (occ > 0 && cus > 0) {
        const int full = occ * cus / 8 * 8;
        if (full >= 8) g.push_back(full);
        if (full / 2 >= 8) g.push_back(full / 2 / 8 * 8);
    }
    return g;
}

int tune(Group &G, hipStream_t s)
{
    float2 *ctmp = nullptr;
    GCHK(hipMalloc((void **)&ctmp, sizeof(float2) * (size_t)G.ctot));
    hipEvent_t e0, e1;
    GCHK(hipEventCreate(&e0));
    GCHK(hipEventCreate(&e1));
    int rc = SO3FT_OK;
    Config best = G.cfg;
    if (hipMemsetAsync(ctmp, 0, sizeof(float2) * (size_t)G.ctot, s) != hipSuccess ||
        hipMemsetAsync(G.F, 0, sizeof(float2) * (size_t)G.gtot, s) != hipSuccess)
        rc = set_error(SO3FT_EHIP, "tuner: clearing scratch failed");
    float tbest = 1e30f;
    for (int TB : {1, 2}) {
        if (rc) break;
        std::vector<int4> v;
        Items it;
        if ((rc = build_items(G, true, TB, 1, v)) || (rc = upload_items(v, it))) break;
        for (int g : grid_candidates(mixed_grouped_occupancy(1, TB, (int)G.max_bw))) {
            Config c = G.cfg;
            c.inv_tb = TB;
            c.inv_grid = g;
            const float ms = time_launches(s, e0, e1, [&] { return launch_inv_dwt(G, c, it, ctmp, s); }, &rc);
            if (rc) break;
            if (ms < tbest) { tbest = ms; best.inv_tb = TB; best.inv_grid = g; }
        }
        hipFree(it.d);
    }
    tbest = 1e30f;
    for (int TB : {4, 2, 1}) {
        if (rc) break;
        std::vector<int4> v;
        Items it;
        if ((rc = build_items(G, false, TB, 1, v)) || (rc = upload_items(v, it))) break;
        for (int g : grid_candidates(mixed_grouped_occupancy(0, TB, (int)G.max_bw))) {
            Config c = G.cfg;
            c.fwd_tb = TB;
            c.fwd_nw = 1;
            c.fwd_grid = g;
            const float ms = time_launches(s, e0, e1, [&] { return launch_fwd_dwt(G, c, it, ctmp, s); }, &rc);
            if (rc) break;
            if (ms < tbest) { tbest = ms; best.fwd_tb = TB; best.fwd_nw = 1; best.fwd_grid = g; }
        }
        hipFree(it.d);
    }
    hipStreamSynchronize(s);
    hipFree(ctmp);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    G.cfg = best;
    if ((rc = set_items(G))) return rc;
    drop_graphs(G);
    G.tuned = true;
    std::vector<std::pair<int64_t, int64_t>> key;
    for (auto &sg : G.segs) key.push_back({sg.bw, sg.T});
    std::lock_guard<std::mutex> lk(g_mu);
    g_tuned[{G.device, key}] = best;
    return SO3FT_OK;
}

int run(int64_t h, bool inverse, const void *in, void *out, void *stream)
{
    Group *G = get(h);
    if (!G) return set_error(SO3FT_EHANDLE, "unknown group handle");
    if (!in || !out) return set_error(SO3FT_EINVAL, "null array pointer");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    GCHK(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone) {  // caller's own capture: launches go into the caller's graph
        if (G->sched[inverse ? 1 : 0] < 0) G->sched[inverse ? 1 : 0] = 0;
        return enqueue(*G, inverse, in, out, s, false);
    }
    if (!G->tuned)
        if (int rc = tune(*G, s)) return rc;
    if (!G->graphs) {
        const int keep = G->sched[inverse ? 1 : 0];
        if (keep < 0) G->sched[inverse ? 1 : 0] = 0;
        return enqueue(*G, inverse, in, out, s, false);
    }
    auto &cache = inverse ? G->ginv : G->gfwd;
    const auto key = std::make_pair(in, out);
    auto itg = cache.find(key);
    if (itg == cache.end()) {
        // at most 8 graphs per direction; past that, the evicted graph is updated for the new
        // buffers instead of instantiating another (callers passing fresh buffers every call
        // keep the replay without an instantiation per call)
        hipGraphExec_t evicted = nullptr;
        hipEvent_t done = nullptr;
        if (cache.size() >= 8) {
            evicted = cache.begin()->second.ex;
            done = cache.begin()->second.done;
            cache.erase(cache.begin());
            const hipError_t ew = hipEventSynchronize(done);   // its last launch has finished
            if (ew != hipSuccess) {
                hipGraphExecDestroy(evicted);
                hipEventDestroy(done);
                return set_error(SO3FT_EHIP, std::string("hipEventSynchronize: ") + hipGetErrorString(ew));
            }
        } else {
            GCHK(hipEventCreateWithFlags(&done, hipEventDisableTiming));
        }
        hipGraphExec_t ex = nullptr;
        int &sc = G->sched[inverse ? 1 : 0];
        if (sc < 0 && evicted) {   // not reached: a full cache implies a chosen schedule
            hipGraphExecDestroy(evicted);
            evicted = nullptr;
        }
        if (sc < 0) {
            // time every schedule's graph on the caller's buffers (the output is rewritten by
            // the launch below) and keep the fastest
            hipEvent_t e0, e1;
            GCHK(hipEventCreate(&e0));
            GCHK(hipEventCreate(&e1));
            float best = 1e30f;
            int rc = SO3FT_OK;
            for (int cand = 0; cand < 3 && !rc; ++cand) {
                sc = cand;
                hipGraphExec_t cx = nullptr;
                if ((rc = capture(*G, inverse, in, out, &cx))) break;
                float ms = 0.f;
                hipError_t e = hipGraphLaunch(cx, s);
                if (e == hipSuccess) e = hipEventRecord(e0, s);
                for (int r = 0; r < 2 && e == hipSuccess; ++r) e = hipGraphLaunch(cx, s);
                if (e == hipSuccess) e = hipEventRecord(e1, s);
                if (e == hipSuccess) e = hipEventSynchronize(e1);
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                if (e != hipSuccess) {
                    hipGraphExecDestroy(cx);
                    rc = set_error(SO3FT_EHIP, std::string("schedule timing: ") + hipGetErrorString(e));
                    break;
                }
                if (ms < best) {
                    best = ms;
                    if (ex) hipGraphExecDestroy(ex);
                    ex = cx;
                    G->sched_ms[inverse ? 1 : 0][cand] = ms / 2;
                } else {
                    G->sched_ms[inverse ? 1 : 0][cand] = ms / 2;
                    hipGraphExecDestroy(cx);
                }
            }
            hipEventDestroy(e0);
            hipEventDestroy(e1);
            if (rc) {
                if (ex) hipGraphExecDestroy(ex);
                hipEventDestroy(done);
                sc = -1;
                return rc;
            }
            const float *ms3 = G->sched_ms[inverse ? 1 : 0];
            sc = 0;
            for (int c = 1; c < 3; ++c)
                if (ms3[c] < ms3[sc]) sc = c;
        } else if (int rc = capture(*G, inverse, in, out, &ex, evicted)) {
            hipEventDestroy(done);
            return rc;
        }
        Group::GExec ge;
        ge.ex = ex;
        ge.done = done;
        itg = cache.emplace(key, ge).first;
    }
    GCHK(hipGraphLaunch(itg->second.ex, s));
    GCHK(hipEventRecord(itg->second.done, s));
    return SO3FT_OK;
}

}  // namespace

extern "C" {

int so3ft_group_layout(const int64_t *bws, int64_t n, int64_t *coeff_off, int64_t *grid_off, int64_t *coeff_total,
                       int64_t *grid_total)
{
    if (int rc = check_bws(bws, n)) return rc;
    std::vector<Seg> segs;
    std::vector<int64_t> co, go;
    int64_t ct = 0, gt = 0;
    layout(bws, n, segs, co, go, ct, gt);
    if (coeff_off) std::copy(co.begin(), co.end(), coeff_off);
    if (grid_off) std::copy(go.begin(), go.end(), grid_off);
    if (coeff_total) *coeff_total = ct;
    if (grid_total) *grid_total = gt;
    return SO3FT_OK;
}

int so3ft_group_create(const int64_t *bws, int64_t n, int64_t *out)
{
    if (!out) return set_error(SO3FT_EINVAL, "null output handle");
    if (int rc = check_bws(bws, n)) return rc;
    Group *G = new Group();
    G->n = n;
    layout(bws, n, G->segs, G->coff, G->goff, G->ctot, G->gtot);
    G->max_bw = G->segs.front().bw;   // segments are in decreasing bandwidth
    int rc = SO3FT_OK;
    if (hipGetDevice(&G->device) != hipSuccess) rc = set_error(SO3FT_EHIP, "hipGetDevice");
    std::vector<MixSegment> ms;
    for (auto &sg : G->segs) {
        if (rc) break;
        if ((rc = so3ft_init(sg.bw, sg.bw - 1, 0, 1, 0, 0, &sg.plan))) break;
        int64_t bw = 0, lmax = 0;
        if ((rc = plan_device_tables(sg.plan, &sg.tb, &bw, &lmax))) break;
        MixSegment m{};
        m.cosb = sg.tb.cosb; m.w = sg.tb.w; m.seeds = sg.tb.seeds; m.slice = sg.tb.slice;
        m.rc = sg.tb.rc; m.rc_off = sg.tb.rc_off;
        m.bw = (int)sg.bw; m.N = (int)sg.N; m.T = (int)sg.T; m.pad = 0;
        m.coff = sg.coff; m.goff = sg.goff; m.nc = sg.nc;
        ms.push_back(m);
    }
    if (!rc && hipMalloc((void **)&G->dsegs, sizeof(MixSegment) * ms.size()) != hipSuccess)
        rc = set_error(SO3FT_ENOMEM, "segment table allocation failed");
    if (!rc && hipMemcpy(G->dsegs, ms.data(), sizeof(MixSegment) * ms.size(), hipMemcpyHostToDevice) != hipSuccess)
        rc = set_error(SO3FT_EHIP, "segment table upload failed");
    if (!rc && hipMalloc((void **)&G->F, sizeof(float2) * (size_t)G->gtot) != hipSuccess) {
        hipGetLastError();
        rc = set_error(SO3FT_ENOMEM, "grouped scratch allocation of " + std::to_string(8 * G->gtot) + " bytes failed");
    }
    if (!rc) {  // a tuned configuration of the same bandwidth multiset is reused
        std::vector<std::pair<int64_t, int64_t>> key;
        for (auto &sg : G->segs) key.push_back({sg.bw, sg.T});
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_tuned.find({G->device, key});
        if (it != g_tuned.end()) {
            G->cfg = it->second;
            G->tuned = true;
        }
    }
    if (!rc) rc = set_items(*G);
    if (rc) {
        free_group(G);
        return rc;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    *out = g_next++;
    g_groups[*out] = G;
    return SO3FT_OK;
}

int so3ft_group_destroy(int64_t h)
{
    Group *G = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_groups.find(h);
        if (it == g_groups.end()) return set_error(SO3FT_EHANDLE, "unknown group handle");
        G = it->second;
        g_groups.erase(it);
    }
    free_group(G);
    return SO3FT_OK;
}

int so3ft_group_isoft_c64(int64_t h, const void *coeffs, void *so3funcs, void *stream)
{
    return run(h, true, coeffs, so3funcs, stream);
}

int so3ft_group_soft_c64(int64_t h, const void *so3funcs, void *coeffs, void *stream)
{
    return run(h, false, so3funcs, coeffs, stream);
}

int so3ft_group_autotune(int64_t h, void *stream)
{
    Group *G = get(h);
    if (!G) return set_error(SO3FT_EHANDLE, "unknown group handle");
    return tune(*G, reinterpret_cast<hipStream_t>(stream));
}

int so3ft_group_get_config(int64_t h, int32_t *cfg)
{
    Group *G = get(h);
    if (!G) return set_error(SO3FT_EHANDLE, "unknown group handle");
    if (!cfg) return set_error(SO3FT_EINVAL, "null config array");
    const Config &c = G->cfg;
    const int32_t v[8] = {c.inv_tb, c.inv_grid, c.fwd_tb, c.fwd_nw, c.fwd_grid, G->tuned ? 1 : 0, G->sched[1],
                          G->sched[0]};
    std::copy(v, v + 8, cfg);
    return SO3FT_OK;
}

int so3ft_group_set_config(int64_t h, const int32_t *cfg)
{
    Group *G = get(h);
    if (!G) return set_error(SO3FT_EHANDLE, "unknown group handle");
    if (!cfg) return set_error(SO3FT_EINVAL, "null config array");
    Config c;
    c.inv_tb = cfg[0]; c.inv_grid = cfg[1]; c.fwd_tb = cfg[2]; c.fwd_nw = cfg[3]; c.fwd_grid = cfg[4];
    const bool inv_ok = c.inv_tb == 1 || c.inv_tb == 2;
    const bool fwd_ok = (c.fwd_tb == 1 || c.fwd_tb == 2 || c.fwd_tb == 4) && c.fwd_nw == 1;
    if (!inv_ok || !fwd_ok || c.inv_grid < 0 || c.fwd_grid < 0 || (c.inv_grid % 8) || (c.fwd_grid % 8))
        return set_error(SO3FT_EINVAL, "config: inverse TB in {1,2}; forward TB in {1,2,4} with NW = 1; "
                                       "grids 0 (one block per item) or a positive multiple of 8");
    if (cfg[6] < -1 || cfg[6] > 2 || cfg[7] < -1 || cfg[7] > 2)
        return set_error(SO3FT_EINVAL, "config: schedules -1 (by timing), 0 (grouped launch), 1 (segment branches) "
                                       "or 2 (segments in sequence)");
    G->cfg = c;
    if (int rc = set_items(*G)) return rc;
    drop_graphs(*G);
    G->tuned = true;
    G->sched[1] = cfg[6];
    G->sched[0] = cfg[7];
    return SO3FT_OK;
}

int so3ft_group_set_graphs(int64_t h, int enable)
{
    Group *G = get(h);
    if (!G) return set_error(SO3FT_EHANDLE, "unknown group handle");
    G->graphs = enable != 0;
    return SO3FT_OK;
}

}  // extern "C"
