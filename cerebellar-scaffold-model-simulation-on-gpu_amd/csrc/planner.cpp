// planner.cpp -- .dat loader and the host planner that turns a network description
// into the device layout (device_types.hpp): device order, sliced-ELL slices and
// merged per-neuron CSR rows, weight storage, the work list, stimulus ownership and
// out-degrees.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "model.hpp"

namespace sc {

// ------------------------------------------------------------------ .dat loader
// GetMaxConv + LoadConnectivityFile_ELL (init_params.cu:13-67,96-128): text lines
// "pre post" (population-local ids).  Rows keep the file order of their entries
// (h_col[post]++ at :43-48).  Unlike the reference, which writes out of bounds
// for such lines, ids outside [0,preNum) x [0,postNum) are rejected; pairs whose
// post id lies in [post population size, postNum) are dropped, as the reference
// allocates those ELL rows but never reads them.
int load_dat(const char *file, int preNum, int postNum, int postPopNum, Projection &p)
{
    FILE *fp = fopen(file, "r");
    if (!fp) return fail(SCAFFOLD_ERR_IO, "can't open file :  %s", file);
    std::vector<int32_t> pre, post;
    char line[256];
    int lineno = 0;
    while (fgets(line, sizeof(line), fp)) {
        ++lineno;
        char *s = line;
        while (*s == ' ' || *s == '\t') ++s;
        if (*s == '\n' || *s == '\r' || *s == 0) continue;
        char *end1, *end2;
        long a = strtol(s, &end1, 10);
        long b = strtol(end1, &end2, 10);
        if (end1 == s || end2 == end1) {
            fclose(fp);
            return fail(SCAFFOLD_ERR_FORMAT, "%s:%d: expected 'pre<TAB>post'", file, lineno);
        }
        if (a < 0 || a >= preNum || b < 0 || b >= postNum) {
            fclose(fp);
            return fail(SCAFFOLD_ERR_FORMAT, "%s:%d: pair (%ld, %ld) outside [0,%d) x [0,%d)", file, lineno, a, b,
                        preNum, postNum);
        }
        if (b >= postPopNum) continue;  // an ELL row the kernels never read (init_params.cu:743)
        pre.push_back((int32_t)a);
        post.push_back((int32_t)b);
    }
    fclose(fp);
    p.row_ptr.assign(postPopNum + 1, 0);
    for (int32_t b : post) p.row_ptr[b + 1]++;
    for (int j = 0; j < postPopNum; ++j) p.row_ptr[j + 1] += p.row_ptr[j];
    p.col.assign(pre.size(), 0);
    std::vector<int64_t> fillp(p.row_ptr.begin(), p.row_ptr.end() - 1);
    for (size_t i = 0; i < pre.size(); ++i) p.col[fillp[post[i]]++] = pre[i];
    p.val.clear();  // every stored ELL value is 1 (init_params.cu:46)
    p.max_conv = 0;
    for (int j = 0; j < postPopNum; ++j)
        p.max_conv = std::max(p.max_conv, (int)(p.row_ptr[j + 1] - p.row_ptr[j]));
    return SCAFFOLD_OK;
}

// ------------------------------------------------------------ weight storage
static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

static uint16_t f32_to_bf16(float f)  // round to nearest even
{
    uint32_t u = f2u(f);
    uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(r >> 16);
}
static float bf16_to_f32(uint16_t h) { return u2f((uint32_t)h << 16); }

// OCP FP8 e4m3: bias 7, no infinities, 0x7F/0xFF NaN, largest finite 448.
static float e4m3_to_f32(uint8_t c)
{
    int s = c >> 7, e = (c >> 3) & 15, m = c & 7;
    float v = e == 0 ? std::ldexp((float)m, -9) : std::ldexp(1.f + (float)m / 8.f, e - 7);
    return s ? -v : v;
}

static uint8_t f32_to_e4m3(float x)  // nearest, ties to even code, saturating at 448
{
    struct Table {
        float v[127];
        Table() { for (int c = 0; c < 127; ++c) v[c] = e4m3_to_f32((uint8_t)c); }
    };
    static const Table tab;  // thread-safe one-time initialization
    const float *table = tab.v;
    uint8_t sign = x < 0 ? 0x80 : 0;
    float a = std::fabs(x);
    if (a >= 448.f) return sign | 126;
    int hi = (int)(std::upper_bound(table, table + 127, a) - table);  // first > a
    int lo = hi - 1;
    if (hi >= 127) return sign | 126;
    float dl = a - table[lo], dh = table[hi] - a;
    int c = (dl < dh) ? lo : (dh < dl) ? hi : ((lo & 1) ? hi : lo);
    return sign | (uint8_t)c;
}

// Effective multiplier of synapse e of projection `proj` as the engine applies it.
float plan_multiplier(const scaffold_env *env, int proj, int64_t e)
{
    const Projection &p = env->projs[proj];
    if (p.val.empty()) return 1.f;
    const ScProj &sp = env->plan.projs[proj];
    switch (sp.wprec) {
    case SC_W_BF16: return bf16_to_f32(f32_to_bf16(p.val[e]));
    case SC_W_FP8: return e4m3_to_f32(f32_to_e4m3(p.val[e] / sp.scale)) * sp.scale;
    default: return p.val[e];
    }
}

// Multipliers as storage `wprec` holds them, in place (dequantized): BF16 rounds
// to nearest even; FP8-e4m3 stores x / scale with the projection's power-of-two
// scale 2^ceil(log2(max|x| / 448)), so every stored value lies within the
// format's 448.  *sum_scale (BF16 / FP8) is the sum of the given values over the
// sum of the stored ones: round to nearest biases the mean of a distribution
// that straddles a binade (U(0.8, 1.2) loses 0.3 % in FP8), and the engine
// multiplies each row sum of the projection by this factor.
void encode_values(int wprec, float *v, int64_t n, float *scale, float *sum_scale)
{
    *scale = 1.f;
    *sum_scale = 1.f;
    double gen_sum = 0.0, q_sum = 0.0;
    for (int64_t e = 0; e < n; ++e) gen_sum += v[e];
    if (wprec == SC_W_FP8) {
        float mx = 0.f;
        for (int64_t e = 0; e < n; ++e) mx = std::max(mx, std::fabs(v[e]));
        *scale = mx > 0.f ? std::ldexp(1.f, (int)std::ceil(std::log2((double)mx / 448.0))) : 1.f;
        for (int64_t e = 0; e < n; ++e) v[e] = e4m3_to_f32(f32_to_e4m3(v[e] / *scale)) * *scale;
    } else if (wprec == SC_W_BF16) {
        for (int64_t e = 0; e < n; ++e) v[e] = bf16_to_f32(f32_to_bf16(v[e]));
    }
    if (wprec == SC_W_FP8 || wprec == SC_W_BF16) {
        for (int64_t e = 0; e < n; ++e) q_sum += v[e];
        if (q_sum > 0.0 && gen_sum > 0.0) *sum_scale = (float)(gen_sum / q_sum);
    }
}

// Multipliers are summed in f64.  The sum of a row is exact when every value is a
// multiple of 2^q (q from the smallest magnitude) and the row's absolute sum stays
// below 2^(q+52); then the order of accumulation cannot change the result.
static bool row_sums_exact(const std::vector<float> &v, const std::vector<int64_t> &row_ptr)
{
    int q = 1 << 30;
    for (float x : v) {
        if (x == 0.f) continue;
        if (!std::isfinite(x)) return false;
        int e;
        std::frexp(x, &e);  // |x| = f * 2^e, f in [0.5, 1)
        int qx = std::max(e - 24, -149);
        q = std::min(q, qx);
    }
    if (q == (1 << 30)) return true;
    const double bound = std::ldexp(1.0, q + 52);
    for (size_t j = 0; j + 1 < row_ptr.size(); ++j) {
        double s = 0.0;
        for (int64_t e = row_ptr[j]; e < row_ptr[j + 1]; ++e) s += std::fabs((double)v[e]);
        if (s >= bound) return false;
    }
    return true;
}

// ------------------------------------------------------------------- planner
int plan_build(scaffold_env *env, Plan &P)
{
    const Options &opt = env->opt;
    const bool want16 = opt.index_bits == 16;
    const int npop = env->n_types;
    P.pops.assign(npop, ScPop{});
    int dev = 0, ref = 0;
    for (int p = 0; p < npop; ++p) {
        const Population &src = env->pops[p];
        ScPop &d = P.pops[p];
        const int num = src.set ? src.num : 0;
        d.dev_base = dev;
        d.group_base = dev / SC_WAVE;
        d.ref_base = ref;
        d.num = num;
        d.kind = src.set ? src.kind : SCAFFOLD_INPUT;
        d.c_num = src.c_num;
        d.dtc = src.Cm > 0 ? kDT / (double)src.Cm : 0.0;
        d.El = src.El; d.Vr = src.Vr; d.Vth = src.Vth; d.Ie = src.Ie; d.gL = src.gL;
        d.Vr_span = src.Vth - src.Vr;
        d.ref_step = (int)((double)src.dt_ref / kDT);
        for (int c = 0; c < 2; ++c) {
            d.dtt[c] = (c < src.c_num) ? kDT / (double)src.tau[c] : 0.0;
            d.E[c] = src.E[c];
            d.gw[c] = src.g_bar[c] * src.w[c];
            d.exc_ch[c] = src.E[c] > src.El ? 1 : 0;
        }
        dev += (num + SC_WAVE - 1) / SC_WAVE * SC_WAVE;
        ref += num;
        if (src.set && src.kind == SCAFFOLD_OUTPUT) P.out_pops.push_back(p);
    }
    P.Ndev = dev;
    P.n_groups = dev / SC_WAVE;
    P.Wr = (P.n_groups + 2) & ~1;  // at least one always-zero word; even, so rows are 16-byte aligned
    P.total_ref = ref;
    P.n_output = 0;
    P.out_bits.clear();
    for (int p : P.out_pops) {
        P.n_output += P.pops[p].num;
        for (int i = 0; i < P.pops[p].num; ++i) P.out_bits.push_back(P.pops[p].dev_base + i);
    }
    P.T_out = env->T_cpu_sim > 0 ? env->T_cpu_sim : 1;
    if ((int64_t)dev + SC_WAVE > (int64_t)SC_BIT_MASK)
        return fail(SCAFFOLD_ERR_CAPACITY, "more than 2^29 neurons in one instance");
    const uint32_t sentinel = (uint32_t)(P.Wr - 1) * SC_WAVE;  // a bit of the row's last, always-zero word

    P.dev_to_ref.assign(dev, -1);
    P.ref_to_dev.assign(ref, 0);
    P.groups.clear();
    for (int p = 0; p < npop; ++p) {
        const ScPop &d = P.pops[p];
        for (int i = 0; i < d.num; ++i) {
            P.dev_to_ref[d.dev_base + i] = d.ref_base + i;
            P.ref_to_dev[d.ref_base + i] = d.dev_base + i;
        }
        const int ng = (d.num + SC_WAVE - 1) / SC_WAVE;
        for (int gi = 0; gi < ng; ++gi) {
            ScGroup g{};
            g.pop = p;
            g.first_local = gi * SC_WAVE;
            g.nvalid = std::min(SC_WAVE, d.num - gi * SC_WAVE);
            g.self_final = 1;
            P.groups.push_back(g);
        }
    }

    // projections (index order fixes the accumulation order per post population)
    const int nproj = env->n_conn;
    // Spike ring rows.  Step n writes row n and clears row n + 1 (its team-mode
    // words are assembled by atomicOr during step n + 1), so every row a step reads
    // -- n - delay, and for plasticity n - delay - s (s < window) and n - t_delay --
    // must differ from both: R >= delay + 2 and R >= delay + window + 2.  The
    // reference's 2 x T_print rows (GPUScaffoldSimulation.cu:361) are grown where
    // they fall short (T_print = 1 with one-step delays gives 2 rows, which the
    // reference can use because it never clears a row ahead).
    for (int pi = 0; pi < nproj; ++pi)
        if (env->projs[pi].set) P.R = std::max(P.R, env->projs[pi].delay() + 2);
    for (const Plasticity &r : env->plast) {
        if (r.window > SC_PLASTIC_WINDOW_MAX)
            return fail(SCAFFOLD_ERR_CAPACITY, "plasticity window %d above %d steps", r.window, SC_PLASTIC_WINDOW_MAX);
        if (r.window < 1) return fail(SCAFFOLD_ERR_INVALID_ARG, "plasticity window must be at least one step");
        P.R = std::max(P.R, env->projs[r.target].delay() + r.window + 2);
        if (r.rule == SC_STDP_TEACHER) P.R = std::max(P.R, env->projs[r.teacher].delay() + 2);
    }
    P.projs.assign(nproj, ScProj{});
    P.parr.assign(nproj, HostProjArrays{});
    P.outdeg.assign(dev, 0);
    P.nnz = 0;
    P.max_delay = 0;
    int64_t weight_bytes = 0;
    // Gather mode per post population:
    //   lane mode: one lane per post neuron over sliced-ELL slices, for populations
    //              whose mean total fan-in is at most lane_max_fanin (and at most
    //              SC_MAX_SLICES incoming projections);
    //   team mode: a team of 1..16 wavefronts per post neuron over one merged CSR
    //              row holding the neuron's synapses of every incoming projection,
    //              projection after projection in index order.
    std::vector<char> team_pop(npop, 0);
    std::vector<int64_t> fanin(npop, 0);
    std::vector<char> weighted(npop, 0);
    for (int pi = 0; pi < nproj; ++pi) {
        const Projection &src = env->projs[pi];
        if (!src.set) continue;
        fanin[src.post] += (int64_t)src.col.size();
        if (!src.val.empty()) weighted[src.post] = 1;
        ScPop &post = P.pops[src.post];
        if (post.n_in >= SC_MAX_IN)
            return fail(SCAFFOLD_ERR_CAPACITY, "population %d receives more than %d projections", src.post, SC_MAX_IN);
        post.in_proj[post.n_in++] = pi;
    }
    for (int p = 0; p < npop; ++p) {
        const ScPop &d = P.pops[p];
        if (d.kind == SCAFFOLD_INPUT || d.num == 0) continue;
        const double mean = (double)fanin[p] / d.num;
        team_pop[p] = (mean > (double)opt.lane_max_fanin || d.n_in > SC_MAX_SLICES) ? 1 : 0;
    }
    // plasticity: the target projections hold f32 multipliers (1.0 where none were
    // given), and so do the other weighted projections onto their post
    // populations (a lane group and a merged row have one multiplier storage)
    std::vector<char> plastic_proj(nproj, 0), plastic_post(npop, 0);
    for (const Plasticity &r : env->plast) {
        plastic_proj[r.target] = 1;
        plastic_post[env->projs[r.target].post] = 1;
        weighted[env->projs[r.target].post] = 1;
    }
    auto storage_for = [&](bool has_val, bool f32) {
        if (!has_val) return SC_W_UNIT;
        if (f32) return SC_W_F32;
        return opt.weight_precision == SCAFFOLD_WEIGHTS_BF16 ? SC_W_BF16
             : opt.weight_precision == SCAFFOLD_WEIGHTS_FP8 ? SC_W_FP8 : SC_W_F32;
    };
    auto wbytes_of = [](int wp) { return wp == SC_W_F32 ? 4 : wp == SC_W_BF16 ? 2 : wp == SC_W_FP8 ? 1 : 0; };
    // encoded multipliers of one projection in storage `wp` (scale: FP8 power of two)
    std::vector<std::vector<float>> eff(nproj);
    for (int pi = 0; pi < nproj; ++pi) {
        const Projection &src = env->projs[pi];
        ScProj &sp = P.projs[pi];
        sp.scale = 1.f;
        sp.sum_scale = 1.f;
        if (!src.set) { sp.mode = -1; continue; }
        const ScPop &pre = P.pops[src.pre];
        sp.delay = src.delay();
        sp.exc = src.weight > 0 ? 1 : 0;                       // GPUScaffoldSimulation.cu:732
        sp.W = src.weight > 0 ? src.weight : src.weight * (-1.f); // :733
        sp.pre_pop = src.pre;
        sp.post_pop = src.post;
        if (sp.delay + 2 > P.R) return fail(SCAFFOLD_ERR_STATE, "delay %d beyond the spike ring", sp.delay);
        P.max_delay = std::max(P.max_delay, sp.delay);
        const int64_t nnz = (int64_t)src.col.size();
        P.nnz += nnz;
        sp.mode = team_pop[src.post] ? SC_MODE_ROW : SC_MODE_LANE;
        sp.wprec = storage_for(sp.mode == SC_MODE_ROW ? weighted[src.post] : !src.val.empty() || plastic_proj[pi],
                               plastic_post[src.post]);
        std::vector<float> &ev = eff[pi];
        if (sp.wprec != SC_W_UNIT) {
            ev.resize(nnz);
            for (int64_t e = 0; e < nnz; ++e) ev[e] = src.val.empty() ? 1.f : src.val[e];
            encode_values(sp.wprec, ev.data(), nnz, &sp.scale, &sp.sum_scale);
            if (!row_sums_exact(ev, src.row_ptr))
                return fail(SCAFFOLD_ERR_WEIGHT_RANGE,
                            "projection %d: multipliers span too wide a range for exact f64 row sums", pi);
        }
        weight_bytes += nnz * wbytes_of(sp.wprec);
        for (int64_t e = 0; e < nnz; ++e) P.outdeg[pre.dev_base + src.col[e]]++;
    }
    // Layout defaults by size.  A step whose streams fit the L2 twice over (at
    // most 64 MB) is bound by the latency of its slowest block: wide teams (8
    // synapses per lane) shorten the long rows' gathers.  Above that the blocks
    // queue for the CUs and the step is bound by their throughput: every block
    // costs about the same chain of dependent round trips whatever its work, so
    // narrow teams (64 synapses per lane: fewer, fuller blocks) carry more
    // synapses per block, and blocks stage windows up to 32 words per 64-byte
    // segment their gathers would touch.
    const bool large = P.nnz * (want16 ? 2 : 4) + weight_bytes + (int64_t)ref * 32 > ((int64_t)64 << 20);
    P.lane_synapses = opt.lane_synapses > 0 ? opt.lane_synapses : large ? 64 : 8;
    P.stage_ratio = opt.stage_ratio > 0 ? opt.stage_ratio : large ? 32 : 8;
    // word groups for team-mode projections with at least group_min synapses per
    // presynaptic spike word: 2 within 64 MB (at 96.7k cells every parallel-fibre
    // projection and the stellate and basket cells' own recurrent inputs
    // qualify, so their rows hold groups only), 8 above (the parallel fibres onto
    // Purkinje cells at 4x and 16x; at 100x they hold 5 per word and are slower as
    // groups)
    const int group_min = opt.word_groups >= 0 ? opt.word_groups : large ? 8 : 2;
    P.group_min = group_min;
    // write one encoded multiplier
    auto put_w = [&](std::vector<uint8_t> &out, int64_t pos, const ScProj &sp, float x) {
        if (sp.wprec == SC_W_F32) memcpy(&out[pos * 4], &x, 4);
        else if (sp.wprec == SC_W_BF16) { uint16_t h = f32_to_bf16(x); memcpy(&out[pos * 2], &h, 2); }
        else if (sp.wprec == SC_W_FP8) out[pos] = f32_to_e4m3(x / sp.scale);
    };
    // multiplier position of every synapse of a plastic projection in the array
    // that holds it (lane slices or merged rows)
    std::vector<std::vector<int64_t>> vpos(nproj);
    for (int pi = 0; pi < nproj; ++pi)
        if (plastic_proj[pi]) vpos[pi].assign(env->projs[pi].col.size(), -1);
    // lane-mode slices
    for (int pi = 0; pi < nproj; ++pi) {
        const Projection &src = env->projs[pi];
        const ScProj &sp = P.projs[pi];
        if (!src.set || sp.mode != SC_MODE_LANE) continue;
        const ScPop &pre = P.pops[src.pre];
        const int rows = P.pops[src.post].num;
        const int wb = wbytes_of(sp.wprec);
        HostProjArrays &A = P.parr[pi];
        const int ng = (rows + SC_WAVE - 1) / SC_WAVE;
        A.sell_off.assign(ng, 0);
        A.sell_w.assign(ng, 0);
        int64_t total = 0;
        for (int gi = 0; gi < ng; ++gi) {
            int w = 0;
            for (int l = 0; l < SC_WAVE; ++l) {
                int j = gi * SC_WAVE + l;
                if (j < rows) w = std::max(w, (int)(src.row_ptr[j + 1] - src.row_ptr[j]));
            }
            A.sell_off[gi] = total;
            A.sell_w[gi] = w;
            total += (int64_t)(w + 3) / 4 * 4 * SC_WAVE;
        }
        A.col.assign(total, sentinel);
        if (wb) A.val.assign(total * wb, 0);
        for (int gi = 0; gi < ng; ++gi)
            for (int l = 0; l < SC_WAVE; ++l) {
                int j = gi * SC_WAVE + l;
                if (j >= rows) continue;
                int k = 0;
                for (int64_t e = src.row_ptr[j]; e < src.row_ptr[j + 1]; ++e, ++k) {
                    const int64_t pos = sell_pos(A.sell_off[gi], k, l);
                    A.col[pos] = (uint32_t)(pre.dev_base + src.col[e]);
                    if (wb) put_w(A.val, pos, sp, eff[pi][e]);
                    if (plastic_proj[pi]) vpos[pi][e] = pos;
                }
            }
    }
    // bit ranges of the lane slices and the segments their gathers touch (a column
    // of 64 entries spanning b bits touches about min(64, b / 512 + 1) segments)
    for (int pi = 0; pi < nproj; ++pi) {
        const ScProj &sp = P.projs[pi];
        if (sp.mode != SC_MODE_LANE) continue;
        HostProjArrays &A = P.parr[pi];
        const size_t ng = A.sell_off.size();
        A.sl_lo.assign(ng, UINT32_MAX);
        A.sl_hi.assign(ng, 0);
        A.sl_seg.assign(ng, 0);
        for (size_t gi = 0; gi < ng; ++gi)
            for (int k = 0; k < A.sell_w[gi]; ++k) {
                uint32_t lo = UINT32_MAX, hi = 0;
                for (int l = 0; l < SC_WAVE; ++l) {
                    const uint32_t b = A.col[sell_pos(A.sell_off[gi], k, l)];
                    if (b == sentinel) continue;
                    lo = std::min(lo, b);
                    hi = std::max(hi, b);
                }
                if (lo > hi) continue;
                A.sl_lo[gi] = std::min(A.sl_lo[gi], lo);
                A.sl_hi[gi] = std::max(A.sl_hi[gi], hi);
                A.sl_seg[gi] += std::min<int64_t>(SC_WAVE, (hi - lo) / 512 + 1);
            }
    }
    // 16-bit form of lane-mode slices, per post population: columns are taken in
    // pairs (2p, 2p+1), whose 128 entries must span fewer than 65,535 device bits.
    // A lane's offsets of pairs 2g and 2g+1 (columns 4g .. 4g+3) are adjacent 32-bit
    // words (one 8-byte load); the pair count is padded to even.
    if (want16) {
        struct Conv {
            std::vector<uint16_t> c16;
            std::vector<uint32_t> h;
            std::vector<int64_t> off16;
        };
        for (int p = 0; p < npop; ++p) {
            const ScPop &d = P.pops[p];
            if (team_pop[p] || d.kind == SCAFFOLD_INPUT || d.num == 0) continue;
            std::vector<Conv> conv;
            bool ok = true;
            for (int q = 0; q < d.n_in && ok; ++q) {
                const HostProjArrays &A = P.parr[d.in_proj[q]];
                const size_t ng = A.sell_off.size();
                Conv C;
                C.off16.assign(ng, 0);
                int64_t words = 0;
                for (size_t gi = 0; gi < ng; ++gi) {
                    C.off16[gi] = words;
                    words += (int64_t)(A.sell_w[gi] + 3) / 4 * 2 * SC_WAVE;
                }
                C.c16.assign(words * 2, (uint16_t)SC_PAD16);
                C.h.assign(words / SC_WAVE, 0);
                for (size_t gi = 0; gi < ng && ok; ++gi) {
                    const int w = A.sell_w[gi];
                    for (int pp = 0; 2 * pp < w && ok; ++pp) {
                        uint32_t lo = UINT32_MAX, hi = 0;
                        for (int k = 2 * pp; k < std::min(w, 2 * pp + 2); ++k)
                            for (int l = 0; l < SC_WAVE; ++l) {
                                const uint32_t b = A.col[sell_pos(A.sell_off[gi], k, l)];
                                if (b == sentinel) continue;
                                lo = std::min(lo, b);
                                hi = std::max(hi, b);
                            }
                        if (lo == UINT32_MAX) continue;
                        if (hi - lo >= SC_PAD16) { ok = false; break; }
                        C.h[C.off16[gi] / SC_WAVE + pp] = lo;
                        for (int k = 2 * pp; k < std::min(w, 2 * pp + 2); ++k)
                            for (int l = 0; l < SC_WAVE; ++l) {
                                const uint32_t b = A.col[sell_pos(A.sell_off[gi], k, l)];
                                const int64_t word = C.off16[gi] + ((int64_t)(pp >> 1) * SC_WAVE + l) * 2 + (pp & 1);
                                if (b != sentinel) C.c16[word * 2 + (k & 1)] = (uint16_t)(b - lo);
                            }
                    }
                }
                conv.push_back(std::move(C));
            }
            if (!ok) continue;  // this population keeps 32-bit columns
            for (int q = 0; q < d.n_in; ++q) {
                HostProjArrays &A = P.parr[d.in_proj[q]];
                A.col16 = std::move(conv[q].c16);
                A.hdr = std::move(conv[q].h);
                A.sell_off16 = std::move(conv[q].off16);
                A.idx16 = true;
                std::vector<uint32_t>().swap(A.col);
            }
        }
    }

    // team-mode merged rows
    P.teams.clear();
    P.n_grp = P.grp_syn = 0;
    P.team_of.assign(npop, -1);
    std::vector<std::pair<uint32_t, int64_t>> srow;  // one row of a dense projection: (column, entry), sorted
    for (int p = 0; p < npop; ++p) {
        if (!team_pop[p]) continue;
        const ScPop &d = P.pops[p];
        TeamPop T;
        T.pop = p;
        T.wprec = storage_for(weighted[p], plastic_post[p]);
        const int wb = wbytes_of(T.wprec);
        // Dense incoming projections -- at least group_min synapses per
        // 64-bit spike word of their presynaptic neurons, over all rows -- are held
        // as word groups: per post neuron and presynaptic spike word one 16-byte
        // record (projection slot << 29 | word index, index of its first multiplier,
        // 64-bit mask of the connected neurons of the word) and the multipliers of
        // its set bits, ascending.  A column repeated within a row stays in the
        // merged row (a mask holds each presynaptic neuron once).
        bool dense[SC_MAX_IN] = {};
        auto sort_row = [&](const Projection &src, int j) {
            srow.clear();
            for (int64_t e = src.row_ptr[j]; e < src.row_ptr[j + 1]; ++e) srow.push_back({src.col[e], e});
            std::sort(srow.begin(), srow.end());
        };
        for (int q = 0; q < d.n_in && group_min > 0; ++q) {
            const Projection &src = env->projs[d.in_proj[q]];
            int64_t words = 0;
            for (int j = 0; j < d.num; ++j) {
                sort_row(src, j);
                for (size_t i = 0; i < srow.size(); ++i)
                    words += i == 0 || (srow[i].first >> 6) != (srow[i - 1].first >> 6);
            }
            dense[q] = !src.col.empty() && (int64_t)src.col.size() >= (int64_t)group_min * words;
        }
        T.row_ptr.assign(d.num + 1, 0);
        T.seg_end.assign((size_t)d.num * d.n_in, 0);
        for (int j = 0; j < d.num; ++j) {
            int64_t len = 0;
            for (int q = 0; q < d.n_in; ++q) {
                const Projection &src = env->projs[d.in_proj[q]];
                if (dense[q]) {  // repeated columns only
                    sort_row(src, j);
                    for (size_t i = 1; i < srow.size(); ++i) len += srow[i].first == srow[i - 1].first;
                } else {
                    len += src.row_ptr[j + 1] - src.row_ptr[j];
                }
                T.seg_end[(size_t)j * d.n_in + q] = (int32_t)len;
            }
            // 32-bit form: rows padded to whole 4-entry vectors (16-byte loads)
            T.row_ptr[j + 1] = T.row_ptr[j] + (want16 ? len : (len + 3) / 4 * 4);
        }
        const int64_t tot = T.row_ptr[d.num];
        T.col.resize(tot);
        if (wb) T.val.assign(tot * wb, 0);
        T.grp_ptr.assign(d.num + 1, 0);
        std::vector<uint8_t> gval;  // multipliers of the word groups, group after group
        int64_t n_gval = 0;
        for (int j = 0; j < d.num; ++j) {
            int64_t pos = T.row_ptr[j];
            for (int q = 0; q < d.n_in; ++q) {
                const int pi = d.in_proj[q];
                const Projection &src = env->projs[pi];
                const ScPop &pre = P.pops[src.pre];
                if (dense[q]) {
                    sort_row(src, j);
                    for (size_t i = 0; i < srow.size(); ++i) {
                        const uint32_t bit = (uint32_t)(pre.dev_base + srow[i].first);
                        const int64_t e = srow[i].second;
                        if (i > 0 && srow[i].first == srow[i - 1].first) {  // repeated: merged row
                            T.col[pos] = ((uint32_t)q << SC_Q_SHIFT) | bit;
                            if (wb) put_w(T.val, pos, P.projs[pi], eff[pi][e]);
                            if (plastic_proj[pi]) vpos[pi][e] = pos;
                            ++pos;
                            continue;
                        }
                        const uint32_t word = bit >> 6;
                        const size_t g = T.grp.size() / 4;
                        if (g == 0 || (int64_t)g == T.grp_ptr[j] ||
                            T.grp[4 * (g - 1)] != (((uint32_t)q << SC_Q_SHIFT) | word)) {
                            T.grp.push_back(((uint32_t)q << SC_Q_SHIFT) | word);
                            T.grp.push_back((uint32_t)n_gval);
                            T.grp.push_back(0u);
                            T.grp.push_back(0u);
                        }
                        const size_t gl = T.grp.size() - 4;
                        T.grp[gl + 2 + ((bit >> 5) & 1)] |= 1u << (bit & 31);
                        if (wb) {
                            gval.resize((size_t)(n_gval + 1) * wb);
                            put_w(gval, n_gval, P.projs[pi], eff[pi][e]);
                        }
                        if (plastic_proj[pi]) vpos[pi][e] = -1 - n_gval;  // placed after the merged rows below
                        ++n_gval;
                    }
                } else {
                    for (int64_t e = src.row_ptr[j]; e < src.row_ptr[j + 1]; ++e, ++pos) {
                        T.col[pos] = ((uint32_t)q << SC_Q_SHIFT) | (uint32_t)(pre.dev_base + src.col[e]);
                        if (wb) put_w(T.val, pos, P.projs[pi], eff[pi][e]);
                        if (plastic_proj[pi]) vpos[pi][e] = pos;
                    }
                }
            }
            // padding repeats the last entry (a bit the row reads anyway) with a
            // zero multiplier; the kernel masks it by the row length as well
            for (; pos < T.row_ptr[j + 1]; ++pos) T.col[pos] = T.col[pos - 1];
            T.grp_ptr[j + 1] = (int64_t)T.grp.size() / 4;
        }
        // bit ranges per neuron and projection, segments per neuron (per 64 entries
        // of a projection segment spanning b bits: min(64, b / 512 + 1))
        T.q_lo.assign((size_t)d.num * d.n_in, UINT32_MAX);
        T.q_hi.assign((size_t)d.num * d.n_in, 0);
        T.row_seg.assign(d.num, 0);
        T.q_seg.assign((size_t)d.num * d.n_in, 0);
        for (int j = 0; j < d.num; ++j)
            for (int q = 0; q < d.n_in; ++q) {
                const int64_t s0 = T.row_ptr[j] + (q ? T.seg_end[(size_t)j * d.n_in + q - 1] : 0);
                const int64_t s1 = T.row_ptr[j] + T.seg_end[(size_t)j * d.n_in + q];
                uint32_t &qlo = T.q_lo[(size_t)j * d.n_in + q], &qhi = T.q_hi[(size_t)j * d.n_in + q];
                for (int64_t a = s0; a < s1; a += SC_WAVE) {
                    uint32_t lo = UINT32_MAX, hi = 0;
                    for (int64_t e = a; e < std::min(s1, a + SC_WAVE); ++e) {
                        lo = std::min(lo, T.col[e] & SC_BIT_MASK);
                        hi = std::max(hi, T.col[e] & SC_BIT_MASK);
                    }
                    qlo = std::min(qlo, lo);
                    qhi = std::max(qhi, hi);
                    const int64_t sg = std::min<int64_t>(SC_WAVE, (hi - lo) / 512 + 1);
                    T.row_seg[j] += sg;
                    T.q_seg[(size_t)j * d.n_in + q] += sg;
                }
            }
        // word groups: the bits they read and the 64-byte segments of their words
        for (int j = 0; j < d.num; ++j)
            for (int64_t g = T.grp_ptr[j]; g < T.grp_ptr[j + 1]; ++g) {
                const uint32_t *x = &T.grp[4 * g];
                const int q = (int)(x[0] >> SC_Q_SHIFT);
                const uint32_t word = x[0] & SC_BIT_MASK;
                const uint64_t m = (uint64_t)x[2] | ((uint64_t)x[3] << 32);
                const size_t k = (size_t)j * d.n_in + q;
                T.q_lo[k] = std::min(T.q_lo[k], 64 * word + (uint32_t)__builtin_ctzll(m));
                T.q_hi[k] = std::max(T.q_hi[k], 64 * word + 63 - (uint32_t)__builtin_clzll(m));
                if (g == T.grp_ptr[j] || (T.grp[4 * (g - 1)] & SC_BIT_MASK) >> 3 != word >> 3 ||
                    T.grp[4 * (g - 1)] >> SC_Q_SHIFT != (uint32_t)q) {
                    T.row_seg[j] += 1;
                    T.q_seg[k] += 1;
                }
            }
        if (want16) {
            // 16-bit form: each row is a sequence of 8-entry chunks -- one 16-byte
            // load of offsets per lane -- with one 32-bit header per chunk
            // (q << 29 | base device bit).  A chunk holds entries of one projection,
            // ascending, spanning fewer than 65,535 bits; an entry is its offset
            // from the base; 0xFFFF is padding (multiplier 0).  row_ptr counts
            // chunks.
            constexpr int CH = 8;
            std::vector<int64_t> rp(d.num + 1, 0);
            std::vector<uint16_t> c16;
            std::vector<uint32_t> h;
            std::vector<uint8_t> v16;
            c16.reserve(tot + tot / 4);
            std::vector<int64_t> perm;
            std::vector<int64_t> newpos;  // plastic post population: merged-row position -> chunk entry
            if (plastic_post[p]) newpos.assign(tot, -1);
            for (int j = 0; j < d.num; ++j) {
                // entries of each projection segment in ascending bit order (the
                // row sums are exact, so the order cannot change them)
                perm.resize(T.row_ptr[j + 1] - T.row_ptr[j]);
                for (size_t i = 0; i < perm.size(); ++i) perm[i] = T.row_ptr[j] + (int64_t)i;
                std::stable_sort(perm.begin(), perm.end(), [&](int64_t x, int64_t y) { return T.col[x] < T.col[y]; });
                size_t e = 0;
                while (e < perm.size()) {
                    const int64_t c0 = (int64_t)c16.size();
                    c16.resize(c0 + CH, (uint16_t)SC_PAD16);
                    if (wb) v16.resize((c0 + CH) * wb, 0);
                    const uint32_t first = T.col[perm[e]];
                    const uint32_t q = first >> SC_Q_SHIFT, lo = first & SC_BIT_MASK;
                    h.push_back(first);
                    for (int n = 0; n < CH && e < perm.size(); ++n, ++e) {
                        const uint32_t c = T.col[perm[e]];
                        if ((c >> SC_Q_SHIFT) != q || (c & SC_BIT_MASK) - lo >= SC_PAD16) break;
                        c16[c0 + n] = (uint16_t)((c & SC_BIT_MASK) - lo);
                        if (wb) memcpy(&v16[(c0 + n) * wb], &T.val[perm[e] * wb], wb);
                        if (!newpos.empty()) newpos[perm[e]] = c0 + n;
                    }
                }
                rp[j + 1] = (int64_t)h.size();
            }
            for (int q = 0; q < d.n_in && !newpos.empty(); ++q)
                for (int64_t &x : vpos[d.in_proj[q]])
                    if (x >= 0) x = newpos[x];  // (word-group entries are placed below)
            T.idx16 = true;
            T.row_ptr = std::move(rp);
            T.col16 = std::move(c16);
            T.hdr = std::move(h);
            T.val = std::move(v16);
            std::vector<uint32_t>().swap(T.col);
        }
        // the word groups' multipliers follow the merged rows' in the same array; a
        // group's first-multiplier index is taken relative to its neuron's row
        // start (entry row_ptr[j], or 8 x row_ptr[j] in the 16-bit form)
        if (n_gval > 0) {
            const int64_t base = wb ? (int64_t)T.val.size() / wb : 0;
            if (base + n_gval >= (int64_t)UINT32_MAX)
                return fail(SCAFFOLD_ERR_CAPACITY, "population %d: multipliers beyond 2^32 entries", p);
            for (int j = 0; j < d.num; ++j) {
                const int64_t e0 = T.idx16 ? T.row_ptr[j] * 8 : T.row_ptr[j];
                for (int64_t g = T.grp_ptr[j]; g < T.grp_ptr[j + 1]; ++g)
                    T.grp[4 * g + 1] = (uint32_t)(T.grp[4 * g + 1] + base - e0);
            }
            T.val.insert(T.val.end(), gval.begin(), gval.end());
            for (int q = 0; q < d.n_in; ++q)
                for (int64_t &x : vpos[d.in_proj[q]])
                    if (x < 0) x = base + (-1 - x);
        }
        T.n_grp = (int64_t)T.grp.size() / 4;
        T.n_gval = n_gval;
        P.n_grp += T.n_grp;
        P.grp_syn += n_gval;
        // team width: about lane_synapses synapses per lane for the mean row,
        // counting the synapses of the word groups (narrower teams for rows of
        // groups ran faster at low firing rates but up to 2x slower at high ones:
        // a lane's spiking bits are added one after another)
        const double work = (double)tot + (double)n_gval;
        const double per_lane = work / std::max(1, d.num) / SC_WAVE / std::max(1, P.lane_synapses);
        int t = 1;
        while (t < SC_MAX_TEAM && t < per_lane) t <<= 1;
        T.team_waves = t;
        P.team_of[p] = (int)P.teams.size();
        P.teams.push_back(std::move(T));
    }
    // plasticity rules lowered to the layout (rules in table order)
    P.plastic.clear();
    for (const Plasticity &r : env->plast) {
        const Projection &src = env->projs[r.target];
        const ScProj &sp = P.projs[r.target];
        PlasticPlan X;
        X.rule = r;
        X.delay = sp.delay;
        // the A2_JI window reads rows n - delay - s, s < window; the row of step
        // n + 1 is cleared during step n, so it lies outside the window (the ring
        // was grown above to delay + window + 2 rows)
        if (sp.delay + r.window + 2 > P.R)
            return fail(SCAFFOLD_ERR_STATE, "plasticity on projection %d: window beyond the spike ring", r.target);
        const Projection *tp = r.rule == SC_STDP_TEACHER ? &env->projs[r.teacher] : nullptr;
        if (tp) {
            X.t_delay = tp->delay();
            if (X.t_delay + 2 > P.R) return fail(SCAFFOLD_ERR_STATE, "teacher delay beyond the spike ring");
        }
        // Row sums of the projection stay exact in f64 while every multiplier is a
        // multiple of 2^grid_exp and every row sums to less than 2^(grid_exp + 53):
        // grid_exp is the finest grid for the largest row and multiplier bound.
        // Hebb clamps to [0, 1].  Teacher clamps at 0 only: with constant terms it
        // must not potentiate, so its initial maximum bounds it; a weight-dependent
        // term p - a*w (p = c + a*b) with a > 0 potentiates only below p/a, and
        // while the potentiating terms' a, times the times each applies in one
        // step (1; the window for A2_JI), sum to A <= 1, an update from w <= B
        // gives at most w (1 - A) + A B <= B for B = max(initial maximum, p/a).
        // The bound is widened by 2^-16 for the float rounding of the update.
        const std::vector<float> &ev = eff[r.target];
        float wmax = 0.f;
        for (float x : ev) wmax = std::max(wmax, std::fabs(x));
        double wbound = wmax;
        if (r.rule == SC_STDP_HEBB) {
            wbound = std::max(wbound, 1.0);
        } else {
            double pot_a = 0.0, fix = 0.0;
            for (int c = 0; c < 4; ++c) {
                if (!((r.present >> c) & 1)) continue;
                if (!((r.affine >> c) & 1)) {
                    if (r.coef[c] > 0.f)
                        return fail(SCAFFOLD_ERR_INVALID_ARG, "teacher rule with a potentiating coefficient has no "
                                    "upper bound for its multipliers");
                    continue;
                }
                const double a = r.aa[c], p = r.ac[c] + r.aa[c] * r.ab[c];
                if (a < 0.0)
                    return fail(SCAFFOLD_ERR_INVALID_ARG, "teacher rule: coefficient %d grows with the multiplier "
                                "(a < 0), so its multipliers have no upper bound", c);
                if (p > 0.0) {
                    pot_a += (c == 3 ? r.window : 1) * a;
                    fix = std::max(fix, p / a);
                }
            }
            if (pot_a > 1.0)
                return fail(SCAFFOLD_ERR_INVALID_ARG, "teacher rule: the potentiating terms' slopes sum to %g per step "
                            "(at most 1 keeps the multipliers bounded)", pot_a);
            wbound = std::max(wbound, fix) * (1.0 + std::ldexp(1.0, -16));
        }
        int64_t lmax = 0;
        for (size_t j = 0; j + 1 < src.row_ptr.size(); ++j) lmax = std::max(lmax, src.row_ptr[j + 1] - src.row_ptr[j]);
        const double bound = (double)lmax * wbound;
        X.grid_exp = bound > 0.0 ? std::max(-126, (int)std::ceil(std::log2(bound)) - 52) : -126;
        for (float x : ev) {
            const float y = std::ldexp(x, -X.grid_exp);
            if (y != std::rint(y))
                return fail(SCAFFOLD_ERR_WEIGHT_RANGE, "plasticity on projection %d: multiplier %g is finer than the "
                            "exact-sum grid 2^%d", r.target, (double)x, X.grid_exp);
        }
        if (sp.mode == SC_MODE_LANE) {
            X.team = false;
            X.owner = r.target;
            X.n_val = (int64_t)P.parr[r.target].val.size() / 4;
        } else {
            X.team = true;
            X.owner = P.team_of[src.post];
            X.n_val = (int64_t)P.teams[X.owner].val.size() / 4;
        }
        const int64_t nnz = (int64_t)src.col.size();
        if (nnz >= (int64_t)UINT32_MAX) return fail(SCAFFOLD_ERR_CAPACITY, "plastic projection above 2^32 synapses");
        X.pre.resize(nnz);
        X.vidx.resize(nnz);
        X.ch_begin.assign(1, 0);
        // quiet synapses keep their multiplier when dw = +0 leaves w as it is:
        // no A0 term, A1_J absent or zero, every multiplier inside the clamp
        // range and none of them -0
        bool quiet_keeps = !(r.present & 1u) && (!(r.present & 4u) || (!(r.affine & 4u) && r.coef[2] == 0.f));
        for (float x : ev)
            quiet_keeps = quiet_keeps && !std::signbit(x) && (r.rule != SC_STDP_HEBB || x <= 1.f);
        X.skip_quiet = quiet_keeps;
        const ScPop &post = P.pops[src.post];
        for (size_t j = 0; j + 1 < src.row_ptr.size(); ++j) {
            int32_t pb = (int32_t)(post.dev_base + (int)j);  // the row's trigger
            if (tp) {  // the teacher: the first entry of the teacher projection's row (SynapticPlasticity.cu:93)
                pb = -1;
                if (j + 1 < tp->row_ptr.size() && tp->row_ptr[j] < tp->row_ptr[j + 1])
                    pb = (int32_t)(P.pops[tp->pre].dev_base + tp->col[tp->row_ptr[j]]);
            }
            for (int64_t e = src.row_ptr[j]; e < src.row_ptr[j + 1]; ++e) {
                const int64_t v = vpos[r.target][e];
                if (v < 0 || v >= X.n_val)  // every synapse has one place in the layout
                    return fail(SCAFFOLD_ERR_STATE, "plasticity on projection %d: synapse %lld has no multiplier slot",
                                r.target, (long long)e);
                X.pre[e] = (uint32_t)src.col[e];
                X.vidx[e] = (uint32_t)v;
            }
            for (int64_t b = src.row_ptr[j]; b < src.row_ptr[j + 1]; b += SC_PLASTIC_CHUNK) {
                X.ch_trig.push_back(pb);
                X.ch_begin.push_back((uint32_t)std::min<int64_t>(b + SC_PLASTIC_CHUNK, src.row_ptr[j + 1]));
            }
        }
        P.plastic.push_back(std::move(X));
    }
    // rules sharing a target share its grid: the coarsest of theirs, which bounds
    // every value any of them can produce
    for (PlasticPlan &X : P.plastic)
        for (const PlasticPlan &Y : P.plastic)
            if (Y.rule.target == X.rule.target) X.grid_exp = std::max(X.grid_exp, Y.grid_exp);
    int64_t index_bytes = 0;  // 2 B per synapse in the 16-bit form, else 4 B
    P.syn16 = 0;
    for (int pi = 0; pi < nproj; ++pi) {
        const Projection &src = env->projs[pi];
        if (!src.set) continue;
        const ScProj &sp = P.projs[pi];
        const bool is16 = sp.mode == SC_MODE_LANE ? P.parr[pi].idx16 : P.teams[P.team_of[sp.post_pop]].idx16;
        index_bytes += (int64_t)src.col.size() * (is16 ? 2 : 4);
        if (is16) P.syn16 += (int64_t)src.col.size();
    }
    P.algo_bytes = index_bytes + weight_bytes + (int64_t)ref * 32 + ref / 8;
    P.stream_bytes = (int64_t)ref * 32 + ref / 8;
    for (int pi = 0; pi < nproj; ++pi) {
        const HostProjArrays &A = P.parr[pi];
        if (!env->projs[pi].set || P.projs[pi].mode != SC_MODE_LANE) continue;
        P.stream_bytes += (int64_t)(A.col.size() * 4 + A.col16.size() * 2 + A.hdr.size() * 4 + A.val.size());
    }
    for (const TeamPop &T : P.teams)
        P.stream_bytes += (int64_t)(T.col.size() * 4 + T.col16.size() * 2 + T.hdr.size() * 4 + T.grp.size() * 4 +
                                    T.val.size()) - T.n_gval * wbytes_of(T.wprec);
    P.stage = opt.stage_spikes;
    // the synapse streams of one step beyond 64 MB (twice the aggregate L2) are
    // loaded nontemporal, so they do not evict descriptors and spike rows
    P.nontemporal = opt.nontemporal < 0 ? P.algo_bytes > ((int64_t)64 << 20) : opt.nontemporal != 0;
    for (auto &g : P.groups) g.self_final = 1;

    // work list: team neurons (widest teams first), lane groups, stimulus groups
    P.tasks.clear();
    std::vector<int> order;
    for (size_t ti = 0; ti < P.teams.size(); ++ti) order.push_back((int)ti);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return P.teams[a].team_waves > P.teams[b].team_waves; });
    for (int ti : order) {
        const TeamPop &T = P.teams[ti];
        const ScPop &d = P.pops[T.pop];
        for (int j = 0; j < d.num; ++j) {
            ScTask t{};
            t.kind = SC_TASK_ROW;
            t.group = d.group_base + j / SC_WAVE;
            t.proj = T.pop;
            t.post = j;
            P.tasks.push_back(t);
        }
    }
    for (int gid = 0; gid < P.n_groups; ++gid) {
        const ScGroup &g = P.groups[gid];
        const ScPop &d = P.pops[g.pop];
        if (d.kind == SCAFFOLD_INPUT || team_pop[g.pop]) continue;
        ScTask t{};
        t.kind = SC_TASK_LANE;
        t.group = gid;
        P.tasks.push_back(t);
    }
    for (int gid = 0; gid < P.n_groups; ++gid) {
        const ScGroup &g = P.groups[gid];
        if (P.pops[g.pop].kind != SCAFFOLD_INPUT) continue;
        ScTask t{};
        t.kind = SC_TASK_STIM;
        t.group = gid;
        P.tasks.push_back(t);
    }

    // stimulus ownership: a later stimulus (and a later entry) overwrites an earlier
    // one for a shared id (UserInputFunctions.cu:118)
    P.stim_s.assign(dev, -1);
    P.stim_e.assign(dev, 0);
    for (size_t s = 0; s < env->stims.size(); ++s) {
        const Stimulus &st = env->stims[s];
        const ScPop &d = P.pops[st.pop];
        const int ne = st.n_entries();
        for (int i = 0; i < ne; ++i) {
            int64_t b = st.fanout.empty() ? i : st.fanout[i];
            int64_t e = st.fanout.empty() ? i + 1 : st.fanout[i + 1];
            for (int64_t t = b; t < e; ++t) {
                int di = d.dev_base + (int)st.targets[t];
                P.stim_s[di] = (int32_t)s;
                P.stim_e[di] = i;
            }
        }
    }
    P.stims = dev_stims(env);
    return SCAFFOLD_OK;
}

}  // namespace sc
