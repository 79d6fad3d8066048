// engine.hip -- device-side engine: HBM layout upload, lowering of the planner's
// work list to kernel descriptors, step runs (direct or hipGraph replays),
// queries and batches of instances.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "kernels.hpp"
#include "recorder.hpp"

namespace sc {

// Bit ranges one task reads (per lane slice, or per incoming projection of a team
// neuron) and the 64-byte segments its gathers touch: the input of the staging
// plan.  lo > hi: the slice / projection reads nothing.
// Events a function creates for its own timing, destroyed on every path out of it.
struct LocalEvents {
    std::vector<hipEvent_t> ev;
    LocalEvents() = default;
    LocalEvents(const LocalEvents &) = delete;
    LocalEvents &operator=(const LocalEvents &) = delete;
    ~LocalEvents()
    {
        for (hipEvent_t x : ev) (void)hipEventDestroy(x);
    }
    int make(hipEvent_t *out)
    {
        *out = nullptr;
        HIP_TRY(hipEventCreate(out));
        ev.push_back(*out);
        return SCAFFOLD_OK;
    }
};

struct TaskWin {
    int n = 0;
    int32_t delay[SC_MAX_IN] = {}, pre[SC_MAX_IN] = {};
    uint32_t lo[SC_MAX_IN] = {}, hi[SC_MAX_IN] = {};
    int64_t segs = 0;
    int64_t qsegs[SC_MAX_IN] = {};  // segments of each slice / projection
    bool partial = false;           // a team in the 16-bit form: its projections can be staged one by one
};

// Optional timing of the step kernels inside the runs: a pair of events recorded
// on the engine's stream around the step kernels of every run piece (direct
// launches, or one replay of a graph that holds the step kernels only, the
// statistics and counter advance of the piece following it as direct launches).
struct KTimer {
    bool on = false;
    int every = 1;                 // time one run in `every` (the runs counted in `seq`)
    int64_t seq = 0;
    std::vector<hipEvent_t> pool;  // pairs
    size_t used = 0;
    std::vector<int> piece_k;      // step launches of each measured piece
    double ms = 0.0;
    int64_t launches = 0;
};

// The captured graphs of an engine or a batch, one per use: a whole run (step
// kernels, then the run's end), a whole run whose end also gathers the output
// train, and the step kernels alone (timed runs).  Each is captured on first use
// and again when the run length or the output train changes.
struct RunGraphs {
    enum { RUN, RUN_OUT, STEPS, N };
    struct Slot {
        hipGraphExec_t exec = nullptr;
        int chunk = 0;
        ScOut out;
    } slot[N];
    void destroy()
    {
        for (Slot &x : slot) {
            if (x.exec) (void)hipGraphExecDestroy(x.exec);
            x.exec = nullptr;
        }
    }
};

struct Engine {
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    int64_t bytes = 0;
    ScInstance hi{};
    ScInstance *di = nullptr;
    DevWork w;
    // host copies of the lowered work (batches concatenate them)
    std::vector<ScLaneTask> h_lanes;
    std::vector<ScStimTask> h_stims;
    std::vector<ScTeamTask> h_teams;
    std::vector<int> h_team_t;  // team width (wavefronts) of each team task
    std::vector<TaskWin> h_lane_win, h_team_win;
    ScStim *stims = nullptr;
    int n_stims = 0;
    int n_pops = 0;
    int32_t *cnt_buf = nullptr;  // compaction count
    int32_t *scratch = nullptr;  // the static scratch on this engine's device (no-allocation queries)
    std::vector<int32_t> dev_to_ref;
    int Ndev = 0, total_ref = 0;
    RunGraphs graphs;
    KTimer kt;
    // output spike train: device bits of the OUTPUT neurons, the T_out x n_out
    // byte buffer on the device and its pinned host copy
    const int32_t *out_bits = nullptr;
    uint8_t *out_d = nullptr;
    uint8_t *out_h = nullptr;
    uint8_t *out_hd = nullptr;  // the device's address of out_h (mapped pinned memory), nullptr: copy instead
    hipEvent_t train_ev = nullptr;  // recorded once a call's output train is in out_h (with out_hd)
    int n_out = 0, T_out = 0;
    // plastic projections: multiplier array and per-synapse index, for read-back
    struct PlasticView {
        int proj;
        const float *val;
        int64_t n_val;
        std::vector<uint32_t> vidx;
    };
    std::vector<PlasticView> plastic_views;
};

// Static scratch of scaffold_get_fired_noalloc: word 0 receives the fired count,
// words SC_SCRATCH_IDS .. the ids (256-byte aligned).  One copy per GPU, held
// until the process ends; its device address is looked up once per device with
// hipGetSymbolAddress (in host code the array's name is not a device address).
static __device__ __align__(256) int32_t scratch[SC_SCRATCH_IDS + SC_FIRED_NOALLOC_MAX];
static constexpr int kMaxDevices = 64;
static int32_t *g_scratch[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// The scratch address on the current device (looked up on first use per device).
static int scratch_on_current_device(int32_t **out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(SCAFFOLD_ERR_CAPACITY, "device %d beyond the scratch table", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = static_cast<int32_t *>(p);
    }
    *out = g_scratch[dev];
    return SCAFFOLD_OK;
}

template <typename T>
static int dalloc(std::vector<void *> &allocs, int64_t *bytes, T **p, size_t count)
{
    size_t b = sizeof(T) * (count ? count : 1);
    void *q = nullptr;
    HIP_TRY(hipMalloc(&q, b));
    allocs.push_back(q);
    HIP_TRY(hipMemset(q, 0, b));
    if (bytes) *bytes += (int64_t)b;
    *p = static_cast<T *>(q);
    return SCAFFOLD_OK;
}

template <typename T>
static int dupload(std::vector<void *> &allocs, int64_t *bytes, T **p, const std::vector<T> &v)
{
    int rc = dalloc(allocs, bytes, p, v.size());
    if (rc) return rc;
    if (!v.empty()) HIP_TRY(hipMemcpy(*p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return SCAFFOLD_OK;
}

static int check_device()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return fail(SCAFFOLD_ERR_NO_DEVICE, "no HIP device visible");
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, 0));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(SCAFFOLD_ERR_NO_DEVICE, "device 0 is %s; this library is built for gfx950", prop.gcnArchName);
    return SCAFFOLD_OK;
}

static int wbytes_of(int wprec) { return wprec == SC_W_F32 ? 4 : wprec == SC_W_BF16 ? 2 : wprec == SC_W_FP8 ? 1 : 0; }

// Block layout of the step kernel: team neurons by team width (widest first, so
// the longest rows start earliest), then lane groups, then stimulus groups, then
// the plasticity rules' window-count groups, 16 wavefronts per block.  `team_t` lists the team width of every team descriptor,
// widest first.
static ScLayout make_layout(const std::vector<int> &team_t, int n_lanes, int n_stims, int n_hist, int *n_blocks)
{
    ScLayout L{};
    int blocks = 0;
    size_t i = 0;
    for (int c = 0; c < SC_CLASS_LANE; ++c) {
        const int t = SC_MAX_TEAM >> c, per = 1 << c;
        size_t j = i;
        while (j < team_t.size() && team_t[j] == t) ++j;
        L.cls_task0[c] = (int)i;
        L.cls_ntask[c] = (int)(j - i);
        blocks += (L.cls_ntask[c] + per - 1) / per;
        L.cls_end[c] = blocks;
        i = j;
    }
    L.cls_ntask[SC_CLASS_LANE] = n_lanes;
    blocks += (n_lanes + SC_LANE_PER_BLOCK - 1) / SC_LANE_PER_BLOCK;
    L.cls_end[SC_CLASS_LANE] = blocks;
    L.cls_ntask[SC_CLASS_STIM] = n_stims;
    blocks += (n_stims + SC_STIM_PER_BLOCK - 1) / SC_STIM_PER_BLOCK;
    L.cls_end[SC_CLASS_STIM] = blocks;
    L.cls_ntask[SC_CLASS_HIST] = n_hist;
    blocks += (n_hist + SC_BLOCK / SC_WAVE - 1) / (SC_BLOCK / SC_WAVE);
    L.cls_end[SC_CLASS_HIST] = blocks;
    *n_blocks = blocks;
    return L;
}

// Staging plan: for every block of lane groups or team neurons, the windows of
// ring-row words its tasks read, one per (instance, delay, presynaptic
// population) spanning the union of their bit ranges.  A block stages all its
// windows when they fit the LDS budget and hold at most `ratio` words per
// 64-byte segment its gathers would touch (ratio 8: the copy touches at most
// half the segments the gathers would, words / 16 <= segments / 2).  When they
// do not, a block of team neurons in the 16-bit form (whose chunks name their
// projection) stages the windows that pay most -- most segments per word, each
// within the ratio -- while they fit, and its other projections read the ring
// in memory (their ScTeamIn.lds is SC_LDS_GLOBAL).  The slices and projections
// of a staged window carry the LDS word that stands for word 0 of its ring row.
static std::vector<ScStage> plan_stages(const ScLayout &L, int n_blocks, std::vector<ScLaneTask> &lanes,
                                        const std::vector<TaskWin> &lane_win, std::vector<ScTeamTask> &teams,
                                        const std::vector<TaskWin> &team_win, int ratio, int *n_staged)
{
    constexpr int MAXK = 4 * SC_MAX_WIN;  // candidate windows per block
    std::vector<ScStage> st((size_t)n_blocks);
    *n_staged = 0;
    for (int c = 0; c <= SC_CLASS_LANE; ++c) {
        const bool lane = c == SC_CLASS_LANE;
        const int b0 = c ? L.cls_end[c - 1] : 0;
        for (int b = b0; b < L.cls_end[c]; ++b) {
            const int bi = b - b0, per = lane ? SC_LANE_PER_BLOCK : 1 << c;
            const int first = (lane ? 0 : L.cls_task0[c]) + bi * per;
            const int count = std::min(per, L.cls_ntask[c] - bi * per);
            ScWin wins[MAXK];
            int32_t pre_of[MAXK];
            uint32_t lo_of[MAXK], hi_of[MAXK];
            int64_t seg_of[MAXK];
            int nw = 0;
            bool ok = true, partial = !lane;
            int64_t segs = 0;
            auto key_of = [&](int inst, int delay, int pre) {
                for (int i = 0; i < nw; ++i)
                    if (wins[i].inst == inst && wins[i].delay == delay && pre_of[i] == pre) return i;
                return -1;
            };
            for (int i = 0; i < count && ok; ++i) {
                const TaskWin &W = lane ? lane_win[first + i] : team_win[first + i];
                const int inst = lane ? lanes[first + i].inst : teams[first + i].inst;
                segs += W.segs;
                partial = partial && W.partial;
                for (int r = 0; r < W.n && ok; ++r) {
                    if (W.lo[r] > W.hi[r]) continue;
                    int kx = key_of(inst, W.delay[r], W.pre[r]);
                    if (kx < 0) {
                        if (nw == MAXK) { ok = false; break; }
                        kx = nw++;
                        wins[kx] = ScWin{inst, W.delay[r], 0, 0, 0, 0};
                        pre_of[kx] = W.pre[r];
                        lo_of[kx] = W.lo[r];
                        hi_of[kx] = W.hi[r];
                        seg_of[kx] = 0;
                    }
                    lo_of[kx] = std::min(lo_of[kx], W.lo[r]);
                    hi_of[kx] = std::max(hi_of[kx], W.hi[r]);
                    seg_of[kx] += W.qsegs[r];
                }
            }
            if (!ok || nw == 0) continue;
            int64_t total = 4;  // LDS words 0..3: zero (padding entries read word 0)
            for (int i = 0; i < nw; ++i) {  // whole 16-byte vectors: rows are 16-byte aligned
                wins[i].w0 = (int32_t)((lo_of[i] >> 5) & ~3u);
                wins[i].nw = (int32_t)((((hi_of[i] >> 5) + 4) & ~3u) - (uint32_t)wins[i].w0);
                total += wins[i].nw;
            }
            bool chosen[MAXK] = {};
            int n_chosen = 0;
            if (nw <= SC_MAX_WIN && total <= SC_WIN_WORDS && total <= (int64_t)ratio * segs) {
                for (int i = 0; i < nw; ++i) chosen[i] = true;
                n_chosen = nw;
            } else if (partial) {
                int order[MAXK];
                for (int i = 0; i < nw; ++i) order[i] = i;
                std::sort(order, order + nw, [&](int x, int y) {  // most segments per word first
                    return seg_of[x] * wins[y].nw > seg_of[y] * wins[x].nw;
                });
                int64_t words = 4;
                for (int k = 0; k < nw && n_chosen < SC_MAX_WIN; ++k) {
                    const int i = order[k];
                    if (wins[i].nw > (int64_t)ratio * seg_of[i] || words + wins[i].nw > SC_WIN_WORDS) continue;
                    chosen[i] = true;
                    words += wins[i].nw;
                    ++n_chosen;
                }
            }
            if (n_chosen == 0) continue;
            ScStage &S = st[(size_t)b];
            int64_t off = 4;
            for (int i = 0; i < nw; ++i) {
                if (!chosen[i]) continue;
                wins[i].off = (int32_t)off;
                off += wins[i].nw;
                S.w[S.n++] = wins[i];
            }
            for (int i = 0; i < count; ++i) {
                const TaskWin &W = lane ? lane_win[first + i] : team_win[first + i];
                const int inst = lane ? lanes[first + i].inst : teams[first + i].inst;
                for (int r = 0; r < W.n; ++r) {
                    const int kx = W.lo[r] > W.hi[r] ? -1 : key_of(inst, W.delay[r], W.pre[r]);
                    const int32_t lds = kx < 0 ? 0 : chosen[kx] ? wins[kx].off - wins[kx].w0 : SC_LDS_GLOBAL;
                    if (lane) lanes[first + i].sl[r].lds = lds;
                    else teams[first + i].in[r].lds = lds;
                }
            }
            ++*n_staged;
        }
    }
    return st;
}

static int upload_work(std::vector<void *> &allocs, int64_t *bytes, DevWork &w, const std::vector<int> &team_t,
                       std::vector<ScLaneTask> lanes, const std::vector<ScStimTask> &stims,
                       std::vector<ScTeamTask> teams, const std::vector<TaskWin> &lane_win,
                       const std::vector<TaskWin> &team_win, bool stage, int ratio, int n_hist)
{
    for (size_t i = 1; i < team_t.size(); ++i)
        if (team_t[i] > team_t[i - 1]) return fail(SCAFFOLD_ERR_STATE, "team descriptors not ordered by width");
    if (lane_win.size() != lanes.size() || team_win.size() != teams.size())
        return fail(SCAFFOLD_ERR_STATE, "staging ranges do not match the descriptors");
    int rc;
    ScLaneTask *l;
    ScStimTask *s;
    ScTeamTask *t;
    w.layout = make_layout(team_t, (int)lanes.size(), (int)stims.size(), n_hist, &w.n_blocks);
    int n_staged = 0;
    std::vector<ScStage> st;
    if (stage) st = plan_stages(w.layout, w.n_blocks, lanes, lane_win, teams, team_win, ratio, &n_staged);
    w.stages = nullptr;
    w.n_staged = n_staged;
    if (n_staged > 0) {
        ScStage *ds;
        if ((rc = dupload(allocs, bytes, &ds, st))) return rc;
        w.stages = ds;
    }
    if ((rc = dupload(allocs, bytes, &l, lanes))) return rc;
    if ((rc = dupload(allocs, bytes, &s, stims))) return rc;
    if ((rc = dupload(allocs, bytes, &t, teams))) return rc;
    w.lanes = l;
    w.stims = s;
    w.teams = t;
    if ((rc = dalloc(allocs, bytes, &w.end_done, 1))) return rc;  // zeroed
    w.idx16 = false;
    for (const ScLaneTask &x : lanes) w.idx16 = w.idx16 || x.idx16;
    for (const ScTeamTask &x : teams) w.idx16 = w.idx16 || x.hdr;
    return SCAFFOLD_OK;
}

int engine_create(Engine **out, Plan &P)
{
    int rc = check_device();
    if (rc) return rc;
    Engine *e = new Engine();
    *out = nullptr;
    auto bail = [&](int code) { engine_destroy(e); return code; };
    auto &A = e->allocs;
    int64_t *B = &e->bytes;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(SCAFFOLD_ERR_HIP, "hipStreamCreate failed"));
    ScInstance &I = e->hi;
    I.R = P.R;
    I.Wr = P.Wr;
    I.n_groups = P.n_groups;
    I.n_pops = (int)P.pops.size();
    I.n_stims = (int)P.stims.size();
    I.Ndev = P.Ndev;
    e->n_pops = I.n_pops;
    e->Ndev = P.Ndev;
    e->total_ref = P.total_ref;
    e->dev_to_ref = P.dev_to_ref;
    if ((rc = dalloc(A, B, &I.st, P.Ndev))) return bail(rc);
    // 16 bytes of slack: a staged window's last 16-byte load may run past the last row
    if ((rc = dalloc(A, B, &I.ring, (size_t)P.R * P.Wr + 2))) return bail(rc);
    if ((rc = dalloc(A, B, &I.counts, (size_t)P.R * I.n_pops))) return bail(rc);
    if ((rc = dalloc(A, B, &I.events, (size_t)P.R))) return bail(rc);
    if ((rc = dalloc(A, B, &I.totals, (size_t)I.n_pops + 1))) return bail(rc);
    ScGroup *groups;
    if ((rc = dupload(A, B, &groups, P.groups))) return bail(rc);
    I.groups = groups;
    ScPop *pops;
    if ((rc = dupload(A, B, &pops, P.pops))) return bail(rc);
    I.pops = pops;
    if ((rc = dupload(A, B, &e->stims, P.stims))) return bail(rc);
    I.stims = e->stims;
    e->n_stims = (int)P.stims.size();
    int32_t *ss, *se;
    if ((rc = dupload(A, B, &ss, P.stim_s))) return bail(rc);
    if ((rc = dupload(A, B, &se, P.stim_e))) return bail(rc);
    I.stim_s = ss;
    I.stim_e = se;
    uint32_t *od;
    if ((rc = dupload(A, B, &od, P.outdeg))) return bail(rc);
    I.outdeg = od;
    ScCtl *ctl;
    if ((rc = dalloc(A, B, &ctl, 1))) return bail(rc);
    I.ctl = ctl;

    // synapse streams: sliced-ELL slices of lane-mode projections
    std::vector<const uint8_t *> pcol(P.projs.size(), nullptr);  // 32-bit words or 16-bit offsets
    std::vector<const uint32_t *> phdr(P.projs.size(), nullptr);
    std::vector<const uint8_t *> pval(P.projs.size(), nullptr);
    for (size_t pi = 0; pi < P.projs.size(); ++pi) {
        if (P.projs[pi].mode != SC_MODE_LANE) continue;
        HostProjArrays &H = P.parr[pi];
        if (H.idx16) {
            uint16_t *c16;
            uint32_t *h;
            if ((rc = dupload(A, B, &c16, H.col16))) return bail(rc);
            if ((rc = dupload(A, B, &h, H.hdr))) return bail(rc);
            pcol[pi] = reinterpret_cast<const uint8_t *>(c16);
            phdr[pi] = h;
        } else {
            uint32_t *col;
            if ((rc = dupload(A, B, &col, H.col))) return bail(rc);
            pcol[pi] = reinterpret_cast<const uint8_t *>(col);
        }
        if (!H.val.empty()) {
            uint8_t *val;
            if ((rc = dupload(A, B, &val, H.val))) return bail(rc);
            pval[pi] = val;
        }
        std::vector<uint32_t>().swap(H.col);  // now on the device only
        std::vector<uint16_t>().swap(H.col16);
        std::vector<uint32_t>().swap(H.hdr);
        std::vector<uint8_t>().swap(H.val);
    }
    // merged rows of team-mode populations
    std::vector<const uint8_t *> tcol(P.teams.size(), nullptr);
    std::vector<const uint32_t *> thdr(P.teams.size(), nullptr);
    std::vector<const uint8_t *> tval(P.teams.size(), nullptr);
    std::vector<const uint32_t *> tgrp(P.teams.size(), nullptr);
    for (size_t ti = 0; ti < P.teams.size(); ++ti) {
        TeamPop &T = P.teams[ti];
        if (!T.grp.empty()) {
            uint32_t *g;
            if ((rc = dupload(A, B, &g, T.grp))) return bail(rc);
            tgrp[ti] = g;
        }
        if (T.idx16) {
            uint16_t *c16;
            uint32_t *h;
            if ((rc = dupload(A, B, &c16, T.col16))) return bail(rc);
            if ((rc = dupload(A, B, &h, T.hdr))) return bail(rc);
            tcol[ti] = reinterpret_cast<const uint8_t *>(c16);
            thdr[ti] = h;
        } else {
            uint32_t *col;
            if ((rc = dupload(A, B, &col, T.col))) return bail(rc);
            tcol[ti] = reinterpret_cast<const uint8_t *>(col);
        }
        if (!T.val.empty()) {
            uint8_t *val;
            if ((rc = dupload(A, B, &val, T.val))) return bail(rc);
            tval[ti] = val;
        }
        std::vector<uint32_t>().swap(T.col);
        std::vector<uint16_t>().swap(T.col16);
        std::vector<uint32_t>().swap(T.hdr);
        std::vector<uint8_t>().swap(T.val);
        std::vector<uint32_t>().swap(T.grp);
    }

    // lower the planner's work list to descriptors with absolute pointers
    for (const ScTask &t : P.tasks) {
        if (t.kind == SC_TASK_ROW) {
            const int pop = t.proj, j = t.post;
            const TeamPop &T = P.teams[P.team_of[pop]];
            const ScPop &pp = P.pops[pop];
            ScTeamTask d{};
            const int ti = P.team_of[pop];
            // 32-bit form: row_ptr counts entries; 16-bit form: 8-entry chunks (16 bytes
            // of offsets, one header word and 8 multipliers each)
            const int64_t e0 = T.idx16 ? T.row_ptr[j] * 8 : T.row_ptr[j];
            d.col = tcol[ti] + e0 * (T.idx16 ? 2 : 4);
            d.hdr = thdr[ti] ? thdr[ti] + T.row_ptr[j] : nullptr;
            d.val = tval[ti] ? tval[ti] + e0 * wbytes_of(T.wprec) : nullptr;
            d.group = t.group;
            d.bit = j % SC_WAVE;
            d.pop = pop;
            d.n_in = pp.n_in;
            d.wprec = T.wprec;
            // 16-bit form: the padded group length; 32-bit: the synapse count
            d.len = T.idx16 ? (int)(T.row_ptr[j + 1] - T.row_ptr[j])  // chunks
                            : (pp.n_in ? T.seg_end[(size_t)j * pp.n_in + pp.n_in - 1] : 0);
            d.n_grp = (int)(T.grp_ptr[j + 1] - T.grp_ptr[j]);
            d.grp = d.n_grp ? tgrp[ti] + 4 * T.grp_ptr[j] : nullptr;
            TaskWin tw;
            tw.n = pp.n_in;
            tw.segs = T.row_seg[j];
            tw.partial = T.idx16;
            for (int q = 0; q < pp.n_in; ++q) {
                const ScProj &sp = P.projs[pp.in_proj[q]];
                tw.delay[q] = sp.delay;
                tw.pre[q] = sp.pre_pop;
                tw.qsegs[q] = T.q_seg[(size_t)j * pp.n_in + q];
                tw.lo[q] = T.q_lo[(size_t)j * pp.n_in + q];
                tw.hi[q] = T.q_hi[(size_t)j * pp.n_in + q];
                d.in[q].end = T.seg_end[(size_t)j * pp.n_in + q];
                d.in[q].delay = sp.delay;
                d.in[q].exc = sp.exc;
                d.in[q].W = sp.W;
                d.in[q].scale = sp.scale * sp.sum_scale;  // exact: scale is a power of two
            }
            e->h_teams.push_back(d);
            e->h_team_t.push_back(T.team_waves);
            e->h_team_win.push_back(tw);
        } else if (t.kind == SC_TASK_LANE) {
            const ScGroup &G = P.groups[t.group];
            const ScPop &pp = P.pops[G.pop];
            const int gi = t.group - pp.group_base;
            ScLaneTask l{};
            TaskWin tw;
            l.wp = SC_W_UNIT;
            l.group = t.group;
            l.pop = G.pop;
            l.nvalid = G.nvalid;
            for (int q = 0; q < pp.n_in; ++q) {
                const int pi = pp.in_proj[q];
                const ScProj &sp = P.projs[pi];
                const HostProjArrays &H = P.parr[pi];
                tw.delay[l.n_sl] = sp.delay;
                tw.pre[l.n_sl] = sp.pre_pop;
                tw.lo[l.n_sl] = H.sl_lo[gi];
                tw.hi[l.n_sl] = H.sl_hi[gi];
                tw.segs += H.sl_seg[gi];
                tw.qsegs[l.n_sl] = H.sl_seg[gi];
                tw.n = l.n_sl + 1;
                ScSlice &s = l.sl[l.n_sl++];
                const int64_t off = H.sell_off[gi];
                s.col = H.idx16 ? pcol[pi] + H.sell_off16[gi] * 4 : pcol[pi] + off * 4;
                s.hdr = H.idx16 ? phdr[pi] + H.sell_off16[gi] / SC_WAVE : nullptr;
                l.idx16 = H.idx16 ? 1 : 0;
                s.val = pval[pi] ? pval[pi] + off * wbytes_of(sp.wprec) : nullptr;
                s.width = H.sell_w[gi];
                s.delay = sp.delay;
                s.wprec = sp.wprec;
                s.exc = sp.exc;
                s.W = sp.W;
                s.scale = sp.scale * sp.sum_scale;  // exact: scale is a power of two
                if (s.val) l.wp = sp.wprec;
            }
            e->h_lanes.push_back(l);
            e->h_lane_win.push_back(tw);
        } else {
            const ScGroup &G = P.groups[t.group];
            ScStimTask s{};
            s.group = t.group;
            s.pop = G.pop;
            s.nvalid = G.nvalid;
            e->h_stims.push_back(s);
        }
    }
    int n_hist = 0;  // window-count groups of the plasticity rules
    for (const PlasticPlan &X : P.plastic) {
        const int groups = (P.pops[P.projs[X.rule.target].pre_pop].num + SC_WAVE - 1) / SC_WAVE;
        n_hist += (groups + SC_HIST_GROUPS - 1) / SC_HIST_GROUPS;  // wavefronts of the window-count class
    }
    if ((rc = upload_work(A, B, e->w, e->h_team_t, e->h_lanes, e->h_stims, e->h_teams, e->h_lane_win, e->h_team_win,
                          P.stage != 0, P.stage_ratio, n_hist)))
        return bail(rc);
    // plasticity rules: per-synapse arrays; the multipliers they update are the
    // f32 arrays of the target's lane slices or merged rows
    e->w.plastic.clear();
    for (const PlasticPlan &X : P.plastic) {
        ScPlastic d{};
        uint32_t *pre, *vidx, *cb;
        int32_t *ct;
        if ((rc = dupload(A, B, &pre, X.pre))) return bail(rc);
        if ((rc = dupload(A, B, &vidx, X.vidx))) return bail(rc);
        if ((rc = dupload(A, B, &cb, X.ch_begin))) return bail(rc);
        if ((rc = dupload(A, B, &ct, X.ch_trig))) return bail(rc);
        const uint8_t *v = X.team ? tval[X.owner] : pval[X.owner];
        if (!v && !X.pre.empty()) return bail(fail(SCAFFOLD_ERR_STATE, "plastic projection without multipliers"));
        const ScPop &pp = P.pops[P.projs[X.rule.target].pre_pop];
        uint8_t *hist;
        if ((rc = dalloc(A, B, &hist, (size_t)(pp.num + SC_WAVE - 1) / SC_WAVE * SC_WAVE))) return bail(rc);
        d.hist = hist;
        d.pre_group0 = pp.group_base;
        d.pre_groups = (pp.num + SC_WAVE - 1) / SC_WAVE;
        d.pre = pre;
        d.vidx = vidx;
        d.ch_begin = cb;
        d.ch_trig = ct;
        d.val = const_cast<float *>(reinterpret_cast<const float *>(v));
        d.n_chunks = (int32_t)X.ch_trig.size();
        d.skip_quiet = X.skip_quiet ? 1 : 0;
        d.rule = X.rule.rule;
        d.delay = X.delay;
        d.t_delay = X.t_delay;
        d.window = X.rule.window;
        d.present = (int32_t)(X.rule.present & 15u);
        d.grid_exp = X.grid_exp;
        for (int c = 0; c < 4; ++c) {
            d.coef[c] = X.rule.coef[c];
            d.ac[c] = X.rule.ac[c];
            d.aa[c] = X.rule.aa[c];
            d.ab[c] = X.rule.ab[c];
        }
        d.affine = (int32_t)(X.rule.affine & 15u);
        e->w.plastic.push_back(d);
        e->plastic_views.push_back(Engine::PlasticView{X.rule.target, d.val, X.n_val, X.vidx});
    }
    if (!e->w.plastic.empty()) {
        ScPlastic *dp;
        if ((rc = dupload(A, B, &dp, e->w.plastic))) return bail(rc);
        e->w.d_plastic = dp;
    }
    e->w.max_groups = P.n_groups;
    e->w.nontemporal = P.nontemporal;
    if ((rc = dalloc(A, B, &e->cnt_buf, 1))) return bail(rc);
    if ((rc = scratch_on_current_device(&e->scratch))) return bail(rc);
    // output spike train (T_out steps x n_out neurons), copied to pinned host memory
    e->n_out = P.n_output;
    e->T_out = std::max(1, P.T_out);
    {
        int32_t *ob;
        if ((rc = dupload(A, B, &ob, P.out_bits))) return bail(rc);
        e->out_bits = ob;
        const size_t nb = (size_t)e->T_out * (size_t)std::max(1, e->n_out);
        if ((rc = dalloc(A, B, &e->out_d, nb))) return bail(rc);
        if (hipHostMalloc(reinterpret_cast<void **>(&e->out_h), nb, hipHostMallocMapped) != hipSuccess)
            return bail(fail(SCAFFOLD_ERR_HIP, "hipHostMalloc of the output train failed"));
        memset(e->out_h, 0, nb);
        // the gather writes the train straight into the mapped pinned buffer (no
        // separate device-to-host copy); where the runtime gives no device address
        // for it, the train is gathered on the device and copied
        void *dp = nullptr;
        if (hipHostGetDevicePointer(&dp, e->out_h, 0) == hipSuccess && dp) e->out_hd = static_cast<uint8_t *>(dp);
        else (void)hipGetLastError();
        if (e->out_hd && hipEventCreateWithFlags(&e->train_ev, hipEventDisableTiming) != hipSuccess)
            return bail(fail(SCAFFOLD_ERR_HIP, "hipEventCreate failed"));
    }
    std::vector<ScInstance> one{I};
    if ((rc = dupload(A, B, &e->di, one))) return bail(rc);
    *out = e;
    return SCAFFOLD_OK;
}

void engine_destroy(Engine *e)
{
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    e->graphs.destroy();
    for (void *p : e->allocs) (void)hipFree(p);
    if (e->out_h) (void)hipHostFree(e->out_h);
    if (e->train_ev) (void)hipEventDestroy(e->train_ev);
    for (hipEvent_t x : e->kt.pool) (void)hipEventDestroy(x);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

int engine_read_plastic(Engine *e, int proj, float *out)
{
    for (const Engine::PlasticView &v : e->plastic_views) {
        if (v.proj != proj) continue;
        std::vector<float> buf((size_t)v.n_val);
        HIP_TRY(hipStreamSynchronize(e->stream));
        if (v.n_val > 0) HIP_TRY(hipMemcpy(buf.data(), v.val, sizeof(float) * (size_t)v.n_val, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < v.vidx.size(); ++i) out[i] = buf[v.vidx[i]];
        return SCAFFOLD_OK;
    }
    return -1;
}

int64_t engine_device_bytes(Engine *e) { return e ? e->bytes : 0; }

void engine_blocks(Engine *e, int *n_blocks, int *n_staged)
{
    *n_blocks = e ? e->w.n_blocks : 0;
    *n_staged = e ? e->w.n_staged : 0;
}

int engine_sync(Engine *e)
{
    HIP_TRY(hipStreamSynchronize(e->stream));
    return SCAFFOLD_OK;
}

int engine_init_params(Engine *e, uint64_t seed, int64_t call_index)
{
    launch_init_params(e->stream, e->di, std::max(e->Ndev, e->hi.Wr), seed, (uint64_t)call_index);
    HIP_TRY(hipGetLastError());
    return engine_sync(e);
}

int engine_zero_g(Engine *e)
{
    launch_zero_g(e->stream, e->hi.st, e->Ndev);
    HIP_TRY(hipGetLastError());
    return engine_sync(e);
}

// Start the next run at step n.  The ring row of step n is cleared: every word of
// it is rewritten during step n, and team-mode words are assembled by atomicOr.
int engine_set_ctl(Engine *e, int64_t n, int64_t m)
{
    const int row = (int)(((n % e->hi.R) + e->hi.R) % e->hi.R);
    HIP_TRY(hipMemsetAsync(e->hi.ring + (size_t)row * e->hi.Wr, 0, sizeof(uint64_t) * e->hi.Wr, e->stream));
    launch_set_ctl(e->stream, e->hi.ctl, n, m, row);
    HIP_TRY(hipGetLastError());
    return engine_sync(e);
}

int engine_upload_stims(Engine *e, const std::vector<ScStim> &stims)
{
    if ((int)stims.size() != e->n_stims) return fail(SCAFFOLD_ERR_STATE, "stimulus table size changed");
    if (stims.empty()) return SCAFFOLD_OK;
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(hipMemcpy(e->stims, stims.data(), sizeof(ScStim) * stims.size(), hipMemcpyHostToDevice));
    return SCAFFOLD_OK;
}

// Steps in runs of at most `chunk` (< ring rows, so the end of a run still finds
// every row of the run): the step kernels (and the plasticity kernels after
// each), then the run's end (statistics, the output train when the run is the
// last of a hand-off, the counter advance).  Runs are launched directly; with the
// `graph` option full runs replay a captured hipGraph (batches always do).
static void launch_step_kernels(hipStream_t s, const ScInstance *single, const ScInstance *insts, const DevWork &w,
                                int K)
{
    for (int k = 0; k < K; ++k) {
        launch_step(s, single, insts, w, k);
        if (single) launch_plastic(s, *single, w, k);
    }
}

static void launch_end(hipStream_t s, const ScInstance *single, const ScInstance *insts, int n_inst, const DevWork &w,
                       int K, const ScOut *out)
{
    launch_run_end(s, single, insts, n_inst, K, w.max_groups, out, w.end_done);
}

static bool same_out(const ScOut &a, const ScOut &b)
{
    return a.bits == b.bits && a.out == b.out && a.T == b.T && a.n_out == b.n_out;
}

// Captures `chunk` steps as a graph: the step kernels, and with `end` the run's
// end as well (with the output train when out is set).
static hipError_t capture_run(hipStream_t s, hipGraphExec_t *gexec, const ScInstance *single, const ScInstance *insts,
                              int n_inst, const DevWork &w, int chunk, bool end, const ScOut *out)
{
    hipGraph_t g = nullptr;
    hipError_t err = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (err != hipSuccess) return err;
    launch_step_kernels(s, single, insts, w, chunk);
    if (end) launch_end(s, single, insts, n_inst, w, chunk, out);
    err = hipStreamEndCapture(s, &g);
    if (err == hipSuccess) err = hipGraphInstantiate(gexec, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (err != hipSuccess) *gexec = nullptr;
    return err;
}

// The graph `which` of G for runs of `chunk` steps, captured on first use.
static int graph_for(hipStream_t s, RunGraphs &G, int which, const ScInstance *single, const ScInstance *insts,
                     int n_inst, const DevWork &w, int chunk, const ScOut *out, hipGraphExec_t *ge)
{
    RunGraphs::Slot &x = G.slot[which];
    const ScOut o = out ? *out : ScOut{};
    if (!x.exec || x.chunk != chunk || !same_out(x.out, o)) {
        if (x.exec) (void)hipGraphExecDestroy(x.exec);
        x.exec = nullptr;
        hipError_t err = capture_run(s, &x.exec, single, insts, n_inst, w, chunk, which != RunGraphs::STEPS, out);
        if (err != hipSuccess) return fail(SCAFFOLD_ERR_HIP, "graph capture: %s", hipGetErrorString(err));
        x.chunk = chunk;
        x.out = o;
    }
    *ge = x.exec;
    return SCAFFOLD_OK;
}

// The next pair of timer events (created on first use).
static int timer_pair(KTimer *kt, int K, hipEvent_t *a, hipEvent_t *b)
{
    while (kt->used + 2 > kt->pool.size()) {
        hipEvent_t x;
        HIP_TRY(hipEventCreate(&x));
        kt->pool.push_back(x);
    }
    *a = kt->pool[kt->used];
    *b = kt->pool[kt->used + 1];
    kt->used += 2;
    kt->piece_k.push_back(K);
    return SCAFFOLD_OK;
}

// n_steps steps in runs: graph replays of `chunk` steps, then direct launches for
// the rest (every run direct without graphs).  With the timer on, each run's step
// kernels lie between two events and its end is launched after them.  `out`: the
// output train of the last out->T steps, gathered by the last run's end when that
// run holds them all (*out_taken then set; otherwise the caller gathers it).
// With `train_ev` (one instance), a last run launched directly gathers the train
// by a launch of its own (its first ring row out_row0), records train_ev after it
// and only then launches its end, so a caller that waits for the train alone
// does not wait for the run's statistics (*train_ev_set then set).
static int run_steps(hipStream_t s, RunGraphs &G, const ScInstance *single, const ScInstance *insts, int n_inst,
                     const DevWork &w, int n_steps, int use_graph, int chunk, KTimer *kt = nullptr,
                     const ScOut *out = nullptr, bool *out_taken = nullptr, hipEvent_t train_ev = nullptr,
                     int out_row0 = 0, bool *train_ev_set = nullptr)
{
    if (out_taken) *out_taken = false;
    if (train_ev_set) *train_ev_set = false;
    if (n_steps <= 0) return SCAFFOLD_OK;
    const bool timer = kt && kt->on;
    const int last_k = n_steps % chunk ? n_steps % chunk : chunk;  // steps of the last run
    const ScOut *last_out = out && out->out && out->T > 0 && out->T <= last_k ? out : nullptr;
    int done = 0;
    while (done < n_steps) {
        const int K = std::min(chunk, n_steps - done);
        const bool last = done + K == n_steps;
        const ScOut *o = last ? last_out : nullptr;
        const bool graph = use_graph && K == chunk;
        const bool timed = timer && kt->seq++ % kt->every == 0;  // one run in kt->every
        hipEvent_t a = nullptr, b = nullptr;
        if (timed) {
            int rc = timer_pair(kt, K, &a, &b);
            if (rc) return rc;
            HIP_TRY(hipEventRecord(a, s));
        }
        if (graph) {
            hipGraphExec_t ge = nullptr;
            const int which = timed ? RunGraphs::STEPS : o ? RunGraphs::RUN_OUT : RunGraphs::RUN;
            int rc = graph_for(s, G, which, single, insts, n_inst, w, chunk, which == RunGraphs::RUN_OUT ? o : nullptr,
                               &ge);
            if (rc) return rc;
            HIP_TRY(hipGraphLaunch(ge, s));
        } else {
            launch_step_kernels(s, single, insts, w, K);
        }
        if (timed) HIP_TRY(hipEventRecord(b, s));
        if (!graph && o && train_ev && single) {
            launch_out_train(s, single->ring, single->Wr, single->R, out_row0, o->T, o->bits, o->n_out, o->out);
            HIP_TRY(hipEventRecord(train_ev, s));
            if (train_ev_set) *train_ev_set = true;
            launch_end(s, single, insts, n_inst, w, K, nullptr);
        } else if (timed || !graph) {
            launch_end(s, single, insts, n_inst, w, K, o);
        }
        HIP_TRY(hipGetLastError());
        done += K;
    }
    if (out_taken) *out_taken = last_out != nullptr;
    return SCAFFOLD_OK;
}

static int chunk_for(int graph_steps, int R) { return std::max(1, std::min(graph_steps, R - 1)); }

// Waits for the engine's stream (or only for `upto`, an event recorded on it after
// every timer event), then adds what the step-kernel timer measured.
static int engine_finish(Engine *e, hipEvent_t upto = nullptr)
{
    if (upto) HIP_TRY(hipEventSynchronize(upto));
    else HIP_TRY(hipStreamSynchronize(e->stream));
    KTimer &kt = e->kt;
    float x = 0.f;
    for (size_t i = 0; i < kt.piece_k.size(); ++i) {
        HIP_TRY(hipEventElapsedTime(&x, kt.pool[2 * i], kt.pool[2 * i + 1]));
        kt.ms += x;
        kt.launches += kt.piece_k[i];
    }
    kt.piece_k.clear();
    kt.used = 0;
    return SCAFFOLD_OK;
}

int engine_run(Engine *e, int n_steps, int use_graph, int graph_steps)
{
    int rc = run_steps(e->stream, e->graphs, &e->hi, e->di, 1, e->w, n_steps, use_graph,
                       chunk_for(graph_steps, e->hi.R), &e->kt);
    if (rc) return rc;
    return engine_finish(e);
}

int engine_run_output(Engine *e, int n_steps, int use_graph, int graph_steps, int64_t first_out_step,
                      const uint8_t **train, Recorder *rec, int rec_T, int64_t step0)
{
    const int chunk = chunk_for(graph_steps, e->hi.R);
    int rc = SCAFFOLD_OK;
    // the train of steps [first_out_step, first_out_step + T_out) is gathered by
    // the end of the call's last run when the window closes the call and that run
    // holds all of it; otherwise by a launch of its own after the runs
    ScOut out;
    out.bits = e->out_bits;
    out.out = e->out_hd ? e->out_hd : e->out_d;
    out.T = e->T_out;
    out.n_out = e->n_out;
    const bool closes = train && out.T * out.n_out > 0 && first_out_step + out.T == step0 + n_steps;
    const int R = e->hi.R;
    const int row0 = (int)(((first_out_step % R) + R) % R);
    // the call returns once the train is in host memory (mapped pinned memory:
    // written by the gather itself), not after the last run's statistics
    hipEvent_t early = e->out_hd ? e->train_ev : nullptr;
    bool taken = false, early_set = false;
    if (!rec) {
        rc = run_steps(e->stream, e->graphs, &e->hi, e->di, 1, e->w, n_steps, use_graph, chunk, &e->kt,
                       closes ? &out : nullptr, &taken, early, row0, &early_set);
        if (rc) return rc;
    }
    for (int done = 0; rec && done < n_steps;) {
        const int64_t s = step0 + done;
        const int K = (int)std::min<int64_t>(n_steps - done, (s / rec_T + 1) * rec_T - s);
        if ((rc = recorder_before(rec, e->stream, s + K))) return rc;
        rc = run_steps(e->stream, e->graphs, &e->hi, e->di, 1, e->w, K, use_graph, chunk, &e->kt,
                       closes && done + K == n_steps ? &out : nullptr, &taken, early, row0, &early_set);
        if (rc) return rc;
        if ((s + K) % rec_T == 0 && (rc = recorder_window(rec, e->stream, e->hi.ring, s + K))) return rc;
        done += K;
    }
    if (train) {
        const size_t nb = (size_t)e->T_out * (size_t)e->n_out;
        if (!taken) {
            launch_out_train(e->stream, e->hi.ring, e->hi.Wr, R, row0, e->T_out, e->out_bits, e->n_out, out.out);
            HIP_TRY(hipGetLastError());
        }
        if (!e->out_hd && nb) HIP_TRY(hipMemcpyAsync(e->out_h, e->out_d, nb, hipMemcpyDeviceToHost, e->stream));
        *train = e->out_h;  // complete once engine_finish has synchronized the stream
    }
    return engine_finish(e, early_set ? e->train_ev : nullptr);
}

int engine_record_sync(Engine *e, Recorder *rec)
{
    return rec ? recorder_wait_all(rec, e->stream) : SCAFFOLD_OK;
}

int engine_timer(Engine *e, int on, double *ms, int64_t *launches)
{
    if (on >= 0) {
        e->kt.on = on != 0;
        if (on) { e->kt.ms = 0.0; e->kt.launches = 0; e->kt.every = on; e->kt.seq = 0; }
    }
    if (ms) *ms = e->kt.ms;
    if (launches) *launches = e->kt.launches;
    return SCAFFOLD_OK;
}

int engine_time(Engine *e, int n_steps, int use_graph, int graph_steps, double *ms_total, double *ms_kernel_mean)
{
    LocalEvents L;
    hipEvent_t a, b;
    int rc = L.make(&a);
    if (!rc) rc = L.make(&b);
    if (rc) return rc;
    const int chunk = chunk_for(graph_steps, e->hi.R);
    if (!ms_kernel_mean) {
        HIP_TRY(hipEventRecord(a, e->stream));
        rc = run_steps(e->stream, e->graphs, &e->hi, e->di, 1, e->w, n_steps, use_graph, chunk, &e->kt);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(b, e->stream));
    } else {
        // direct launches; one event pair brackets each run of back-to-back step
        // kernels, so the mean is per step kernel without per-launch event overhead
        std::vector<hipEvent_t> ev;
        HIP_TRY(hipEventRecord(a, e->stream));
        int done = 0;
        while (done < n_steps) {
            const int K = std::min(chunk, n_steps - done);
            hipEvent_t s0, s1;
            rc = L.make(&s0);
            if (!rc) rc = L.make(&s1);
            if (rc) return rc;
            ev.push_back(s0);
            ev.push_back(s1);
            HIP_TRY(hipEventRecord(s0, e->stream));
            for (int k = 0; k < K; ++k) {
                launch_step(e->stream, &e->hi, e->di, e->w, k);
                launch_plastic(e->stream, e->hi, e->w, k);
            }
            HIP_TRY(hipEventRecord(s1, e->stream));
            launch_end(e->stream, &e->hi, e->di, 1, e->w, K, nullptr);
            done += K;
        }
        HIP_TRY(hipEventRecord(b, e->stream));
        HIP_TRY(hipEventSynchronize(b));
        double sum = 0.0;
        for (size_t i = 0; i < ev.size(); i += 2) {
            float x = 0.f;
            HIP_TRY(hipEventElapsedTime(&x, ev[i], ev[i + 1]));
            sum += x;
        }
        *ms_kernel_mean = sum / n_steps;
    }
    HIP_TRY(hipEventSynchronize(b));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, a, b));
    if (ms_total) *ms_total = ms;
    return engine_finish(e);  // the step-kernel timer's events of this run
}

int engine_get_state(Engine *e, float *V, float *g, int32_t *refr)
{
    HIP_TRY(hipStreamSynchronize(e->stream));
    const size_t N = (size_t)e->Ndev;
    std::vector<ScState> hs(N);
    HIP_TRY(hipMemcpy(hs.data(), e->hi.st, sizeof(ScState) * N, hipMemcpyDeviceToHost));
    for (size_t d = 0; d < N; ++d) {
        const int r = e->dev_to_ref[d];
        if (r < 0) continue;
        if (V) V[r] = hs[d].V;
        if (g) { g[2 * (size_t)r] = hs[d].g_exc; g[2 * (size_t)r + 1] = hs[d].g_inh; }
        if (refr) refr[r] = hs[d].refr;
    }
    return SCAFFOLD_OK;
}

int engine_set_state(Engine *e, const float *V, const float *g, const int32_t *refr)
{
    HIP_TRY(hipStreamSynchronize(e->stream));
    const size_t N = (size_t)e->Ndev;
    std::vector<ScState> hs(N);
    HIP_TRY(hipMemcpy(hs.data(), e->hi.st, sizeof(ScState) * N, hipMemcpyDeviceToHost));
    for (size_t d = 0; d < N; ++d) {
        const int r = e->dev_to_ref[d];
        if (r < 0) continue;
        if (V) hs[d].V = V[r];
        if (g) { hs[d].g_exc = g[2 * (size_t)r]; hs[d].g_inh = g[2 * (size_t)r + 1]; }
        if (refr) hs[d].refr = refr[r];
    }
    HIP_TRY(hipMemcpy(e->hi.st, hs.data(), sizeof(ScState) * N, hipMemcpyHostToDevice));
    return SCAFFOLD_OK;
}

int engine_get_ring_row(Engine *e, int row, uint64_t *words)
{
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(hipMemcpy(words, e->hi.ring + (size_t)row * e->hi.Wr, sizeof(uint64_t) * e->hi.Wr, hipMemcpyDeviceToHost));
    return SCAFFOLD_OK;
}

int engine_get_counts(Engine *e, int row, int32_t *counts, unsigned long long *events)
{
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (counts)
        HIP_TRY(hipMemcpy(counts, e->hi.counts + (size_t)row * e->n_pops, sizeof(int32_t) * e->n_pops,
                          hipMemcpyDeviceToHost));
    if (events) HIP_TRY(hipMemcpy(events, e->hi.events + row, sizeof(unsigned long long), hipMemcpyDeviceToHost));
    return SCAFFOLD_OK;
}

int engine_get_totals(Engine *e, unsigned long long *totals)
{
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(hipMemcpy(totals, e->hi.totals, sizeof(unsigned long long) * (e->n_pops + 1), hipMemcpyDeviceToHost));
    return SCAFFOLD_OK;
}

int engine_get_fired(Engine *e, int row, int32_t *ids, int32_t cap, int32_t *n)
{
    int32_t *dout = nullptr;
    const int c = std::max(cap, 1);
    HIP_TRY(hipMalloc(&dout, sizeof(int32_t) * c));
    launch_compact(e->stream, e->hi.ring + (size_t)row * e->hi.Wr, e->hi.n_groups, e->hi.groups, e->hi.pops, dout,
                   cap, e->cnt_buf);
    hipError_t err = hipGetLastError();
    if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
    if (err == hipSuccess) err = hipMemcpy(n, e->cnt_buf, sizeof(int32_t), hipMemcpyDeviceToHost);
    if (err == hipSuccess && cap > 0)
        err = hipMemcpy(ids, dout, sizeof(int32_t) * std::min(cap, *n), hipMemcpyDeviceToHost);
    (void)hipFree(dout);
    if (err != hipSuccess) return fail(SCAFFOLD_ERR_HIP, "fired-list compaction: %s", hipGetErrorString(err));
    return SCAFFOLD_OK;
}

// engine_get_fired without allocating: the compaction writes its count and ids
// into the static scratch of the engine's device, then both are copied out.
int engine_get_fired_noalloc(Engine *e, int row, int32_t *ids, int32_t cap, int32_t *n)
{
    if (cap < 0 || cap > SC_FIRED_NOALLOC_MAX)
        return fail(SCAFFOLD_ERR_CAPACITY, "cap %d: the static scratch holds at most %d ids", cap, SC_FIRED_NOALLOC_MAX);
    int32_t *count = e->scratch, *out = e->scratch + SC_SCRATCH_IDS;
    launch_compact(e->stream, e->hi.ring + (size_t)row * e->hi.Wr, e->hi.n_groups, e->hi.groups, e->hi.pops, out, cap,
                   count);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(hipMemcpy(n, count, sizeof(int32_t), hipMemcpyDeviceToHost));
    const int32_t k = std::min(cap, *n);
    if (k > 0) HIP_TRY(hipMemcpy(ids, out, sizeof(int32_t) * (size_t)k, hipMemcpyDeviceToHost));
    return SCAFFOLD_OK;
}

int engine_copy_probe(int64_t bytes, int iters, double *ms)
{
    int rc = check_device();
    if (rc) return rc;
    const int64_t n = std::max<int64_t>(1, bytes / 16);
    struct Buffers {
        void *a = nullptr, *b = nullptr;
        ~Buffers()
        {
            (void)hipFree(a);
            (void)hipFree(b);
        }
    } B;
    void *&a = B.a, *&b = B.b;
    HIP_TRY(hipMalloc(&a, n * 16));
    HIP_TRY(hipMalloc(&b, n * 16));
    HIP_TRY(hipMemset(a, 1, n * 16));
    LocalEvents L;
    hipEvent_t e0, e1;
    rc = L.make(&e0);
    if (!rc) rc = L.make(&e1);
    if (rc) return rc;
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 256 * 64);
    launch_copy(nullptr, a, b, n, grid);
    HIP_TRY(hipEventRecord(e0));
    for (int i = 0; i < iters; ++i) launch_copy(nullptr, a, b, n, grid);
    HIP_TRY(hipEventRecord(e1));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    *ms = t / iters;
    return SCAFFOLD_OK;
}

// ---------------------------------------------------------------------- batching
// Independent instances of uneven size advance together: the descriptors of all
// instances (each naming its instance) form one step kernel per exchange step.
struct Batch {
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    ScInstance *di = nullptr;
    int n_inst = 0;
    int min_R = 0;
    DevWork w;
    RunGraphs graphs;
};

void batch_destroy(Batch *b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    b->graphs.destroy();
    for (void *p : b->allocs) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int batch_build(Batch **out, std::vector<Engine *> &engines, const std::vector<const Plan *> &plans)
{
    int stage = 1;  // staged only if every instance allows it
    for (const Plan *p : plans) stage = stage && p->stage;
    // nontemporal streams when the instances' streams together exceed 64 MB, or
    // when every instance's own plan loads them so
    int64_t algo = 0;
    bool nt_all = true;
    for (const Plan *p : plans) {
        algo += p->algo_bytes;
        nt_all = nt_all && p->nontemporal;
    }
    *out = nullptr;
    Batch *b = new Batch();
    auto bail = [&](int code) { batch_destroy(b); return code; };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(SCAFFOLD_ERR_HIP, "hipStreamCreate failed"));
    std::vector<ScInstance> insts;
    std::vector<ScLaneTask> lanes;
    std::vector<ScStimTask> stims;
    struct TeamEntry {
        int width;  // wavefronts
        size_t win; // index of its ranges in team_win_u
        ScTeamTask t;
    };
    std::vector<TeamEntry> teams;
    std::vector<TaskWin> lane_win, team_win_u;
    b->min_R = 1 << 30;
    for (size_t i = 0; i < engines.size(); ++i) {
        const Engine *e = engines[i];
        insts.push_back(e->hi);
        b->min_R = std::min(b->min_R, e->hi.R);
        for (ScLaneTask t : e->h_lanes) { t.inst = (int)i; lanes.push_back(t); }
        lane_win.insert(lane_win.end(), e->h_lane_win.begin(), e->h_lane_win.end());
        for (ScStimTask t : e->h_stims) { t.inst = (int)i; stims.push_back(t); }
        for (size_t j = 0; j < e->h_teams.size(); ++j) {
            ScTeamTask t = e->h_teams[j];
            t.inst = (int)i;
            teams.push_back({e->h_team_t[j], team_win_u.size(), t});
            team_win_u.push_back(e->h_team_win[j]);
        }
    }
    std::stable_sort(teams.begin(), teams.end(), [](const TeamEntry &x, const TeamEntry &y) { return x.width > y.width; });
    std::vector<ScTeamTask> tt;
    std::vector<int> tw;
    std::vector<TaskWin> team_win;
    for (const TeamEntry &p : teams) {
        tw.push_back(p.width);
        team_win.push_back(team_win_u[p.win]);
        tt.push_back(p.t);
    }
    int rc;
    if ((rc = dupload(b->allocs, nullptr, &b->di, insts))) return bail(rc);
    if ((rc = upload_work(b->allocs, nullptr, b->w, tw, lanes, stims, tt, lane_win, team_win, stage != 0,
                          plans[0]->stage_ratio, 0)))
        return bail(rc);
    for (const ScInstance &x : insts) b->w.max_groups = std::max(b->w.max_groups, x.n_groups);
    b->w.nontemporal = nt_all || algo > ((int64_t)64 << 20);
    b->n_inst = (int)insts.size();
    *out = b;
    return SCAFFOLD_OK;
}

int batch_run(Batch *b, int n_steps, int graph_steps, double *ms)
{
    LocalEvents L;
    hipEvent_t e0, e1;
    int rc = L.make(&e0);
    if (!rc) rc = L.make(&e1);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(e0, b->stream));
    rc = run_steps(b->stream, b->graphs, nullptr, b->di, b->n_inst, b->w, n_steps, 1,
                   chunk_for(graph_steps, b->min_R));
    if (rc) return rc;
    HIP_TRY(hipEventRecord(e1, b->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t;
    return SCAFFOLD_OK;
}

}  // namespace sc
