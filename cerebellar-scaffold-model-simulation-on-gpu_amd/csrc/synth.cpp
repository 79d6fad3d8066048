// synth.cpp -- network descriptions: the reference's hard-coded scaffold tables and
// a seeded, rule-based synthetic scaffold of any area.
#include <algorithm>
#include <thread>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>

#include "model.hpp"

namespace sc {

// The reference's population tables, in registration order (init_params.cu:377-528).
// Columns: name, enum type, num, Cm, El, dt_ref, Ie, Vr, Vth, gL, device type,
// c_num, {w, E, g_bar, tau} x 2.
struct RefPop {
    const char *name;
    int type, num;
    float Cm, El, dt_ref, Ie, Vr, Vth, gL;
    int dev, c_num;
    double ch[8];
};
static const RefPop kRefPops[7] = {
    {"granule_cell", 0, 88158, 3.0f, -74.0f, 1.5f, 0.0f, -84.0f, -42.0f, 1.5f, SCAFFOLD_NORMAL, 2,
     {1., 0., 1.0, 0.5, 1., -85., 1.0, 10.0}},
    {"glomerulus", 1, 7073, 0, 0, 0, 0, 0, 0, 0, SCAFFOLD_INPUT, 0, {0, 0, 0, 0, 0, 0, 0, 0}},
    {"purkinje_cell", 2, 69, 620.0f, -62.0f, 0.8f, 600.0f, -72.0f, -47.0f, 7.0f, SCAFFOLD_NORMAL, 2,
     {1., 0., 1.0, 0.5, 1., -85., 1.0, 1.6}},
    {"golgi_cell", 3, 219, 76.0f, -65.0f, 2.0f, 36.75f, -75.0f, -55.0f, 3.6f, SCAFFOLD_NORMAL, 2,
     {1., 0.0, 1.0, 0.5, 1., -85.0, 1.0, 15.0}},
    {"stellate_cell", 4, 603, 14.6f, -68.0f, 1.6f, 15.6f, -78.0f, -53.0f, 1.0f, SCAFFOLD_NORMAL, 2,
     {1.0, 0., 1.0, 0.64, 1.0, -85.0, 1.0, 2.0}},
    {"basket_cell", 5, 603, 14.6f, -68.0f, 1.6f, 15.6f, -78.0f, -53.0f, 1.0f, SCAFFOLD_NORMAL, 2,
     {1.0, 0., 1.0, 0.64, 1.0, -85.0, 1.0, 2.0}},
    {"dcn_cell", 6, 12, 89.0f, -59.0f, 3.7f, 55.8f, -69.0f, -48.0f, 1.56f, SCAFFOLD_OUTPUT, 2,
     {1.0, 0.0, 1.0, 7.1, 4.0, -85.0, 1.0, 13.6}},
};
enum { GRC = 0, GLOM, PC, GOC, SC_, BC, DCN, MF };

// The reference's projection table in registration order (init_params.cu:541-764):
// enum type, file, pre, post, initial_weight, delay (ms).
struct RefProj {
    int type;
    const char *file;
    int pre, post;
    float w, delay;
};
static const RefProj kRefProjs[16] = {
    {0, "parallel_fiber_to_purkinje.dat", GRC, PC, 0.02f, 5.0f},
    {1, "parallel_fiber_to_golgi.dat", GRC, GOC, 0.4f, 5.0f},
    {2, "parallel_fiber_to_stellate.dat", GRC, SC_, 0.2f, 5.0f},
    {3, "parallel_fiber_to_basket.dat", GRC, BC, 0.2f, 5.0f},
    {4, "ascending_axon_to_golgi.dat", GRC, GOC, 20.0f, 2.0f},
    {5, "ascending_axon_to_purkinje.dat", GRC, PC, 75.0f, 2.0f},
    {6, "golgi_to_golgi.dat", GOC, GOC, -8.0f, 1.0f},
    {7, "stellate_to_stellate.dat", SC_, SC_, -2.0f, 1.0f},
    {8, "basket_to_basket.dat", BC, BC, -2.5f, 1.0f},
    {9, "glomerulus_to_dcn.dat", GLOM, DCN, 0.006f, 4.0f},
    {10, "purkinje_to_dcn.dat", PC, DCN, -0.0075f, 4.0f},
    {11, "basket_to_purkinje.dat", BC, PC, -9.0f, 4.0f},
    {12, "stellate_to_purkinje.dat", SC_, PC, -8.5f, 5.0f},
    {13, "glomerulus_to_granule.dat", GLOM, GRC, 9.0f, 4.0f},
    {14, "glomerulus_to_golgi.dat", GLOM, GOC, 2.0f, 4.0f},
    {15, "golgi_to_granule.dat", GOC, GRC, -5.0f, 2.0f},
};

static int register_pops(scaffold_env *env, const int *counts, int *idx)
{
    for (int i = 0; i < 7; ++i) {
        const RefPop &r = kRefPops[i];
        int rc = scaffold_set_neuron_params(env, r.type, r.name, 0, counts[i], r.Cm, r.El, r.dt_ref, r.Ie, r.Vr,
                                            r.Vth, r.gL, r.dev, r.c_num, r.ch, &idx[i]);
        if (rc) return rc;
    }
    return SCAFFOLD_OK;
}

int describe_reference(scaffold_env *env, const char *dir, uint64_t stim_seed)
{
    if (env->n_types < 7 || env->n_conn < 16 || env->n_inputs < 2)
        return fail(SCAFFOLD_ERR_CAPACITY, "the reference scaffold needs 7 populations, 16 projections, 2 inputs");
    int counts[7];
    for (int i = 0; i < 7; ++i) counts[i] = kRefPops[i].num;
    int idx[7];
    int rc = register_pops(env, counts, idx);
    if (rc) return rc;
    // init_connectivity_params passes explicit preNum / postNum; glomerulus_to_golgi
    // passes the granule count as postNum (init_params.cu:743).
    for (int i = 0; i < 16; ++i) {
        const RefProj &r = kRefProjs[i];
        std::string f = std::string(dir) + "/" + r.file;
        int post_num = (r.type == 14) ? counts[GRC] : counts[r.post];
        rc = scaffold_set_connectivity_params(env, r.type, f.c_str(), counts[r.pre], post_num, idx[r.pre],
                                              idx[r.post], r.w, r.delay, r.type == 0 ? 1 : 0, nullptr);
        if (rc) return rc;
    }
    // InitInputStimulation (UserInputFunctions.cu:72-82)
    const char *files[2] = {"background_noise.dat", "tone_stim.dat"};
    for (int s = 0; s < 2; ++s) {
        scaffold_stim_func f;
        scaffold_stim_func_reference(s, stim_seed + (uint64_t)s, &f);
        std::string path = std::string(dir) + "/" + files[s];
        rc = scaffold_set_input_stim(env, idx[GLOM], path.c_str(), &f, nullptr);
        if (rc) return rc;
    }
    return SCAFFOLD_OK;
}

// ------------------------------------------------------------ synthetic scaffold
namespace {

struct Rng {  // splitmix64-seeded xoshiro256**
    uint64_t s[4];
    static uint64_t mix(uint64_t &x)
    {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed) { for (auto &v : s) v = mix(seed); }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uni() { return (double)(next() >> 11) * 0x1p-53; }        // [0,1)
    uint64_t below(uint64_t n) { return (uint64_t)(uni() * (double)n); }
};

uint64_t hash3(uint64_t a, uint64_t b, uint64_t c)
{
    uint64_t x = a ^ (b * 0x9E3779B97F4A7C15ull) ^ (c * 0xC2B2AE3D27D4EB4Full);
    return Rng::mix(x);
}

struct Cells {
    std::vector<double> x, z;   // positions (um), sorted by z (ids are z ranks)
};

// Granule cells are numbered strip by strip: strips kStrip um thin in z, one
// after another along z, each ordered by x (the parallel-fibre axis).  A
// parallel-fibre band (|dz| < h, |dx| < 200 um) then covers one run of
// consecutive granule ids in each strip it crosses, so a Purkinje cell's fibres
// fill its presynaptic spike words densely at every sheet size (in plain z order
// the band's cells are interleaved with the rest of the slab: 0.8 x 400 / side
// of a word at most), while the id range of a band stays the z-slab's, as in z
// order.  The cells and their connections are drawn in z order; the result maps a
// z rank to the granule id.
constexpr double kStrip = 4.0;
std::vector<int32_t> strip_order(const Cells &c)
{
    const int n = (int)c.x.size();
    std::vector<int32_t> byz(n), newid(n);
    for (int i = 0; i < n; ++i) byz[i] = i;
    std::sort(byz.begin(), byz.end(), [&](int32_t a, int32_t b) {
        const int sa = (int)(c.z[a] / kStrip), sb = (int)(c.z[b] / kStrip);
        return sa != sb ? sa < sb : c.x[a] != c.x[b] ? c.x[a] < c.x[b] : a < b;
    });
    for (int k = 0; k < n; ++k) newid[byz[k]] = k;
    return newid;
}

Cells place(int n, double X, double Z, uint64_t seed)
{
    Rng r(seed);
    std::vector<std::pair<double, double>> p(n);
    for (auto &q : p) { q.first = r.uni() * Z; q.second = r.uni() * X; }
    std::sort(p.begin(), p.end());
    Cells c;
    c.x.resize(n); c.z.resize(n);
    for (int i = 0; i < n; ++i) { c.z[i] = p[i].first; c.x[i] = p[i].second; }
    return c;
}

// Uniform grid over the sheet for neighbourhood queries.
struct Grid {
    double cs, X, Z;
    int nx, nz;
    std::vector<int> start, ids;
    Grid(const Cells &c, double cell, double X_, double Z_) : cs(cell), X(X_), Z(Z_)
    {
        nx = std::max(1, (int)std::ceil(X / cs));
        nz = std::max(1, (int)std::ceil(Z / cs));
        start.assign((size_t)nx * nz + 1, 0);
        const int n = (int)c.x.size();
        std::vector<int> cid(n);
        for (int i = 0; i < n; ++i) {
            int ix = std::min(nx - 1, (int)(c.x[i] / cs)), iz = std::min(nz - 1, (int)(c.z[i] / cs));
            cid[i] = iz * nx + ix;
            start[cid[i] + 1]++;
        }
        for (size_t k = 0; k + 1 < start.size(); ++k) start[k + 1] += start[k];
        ids.resize(n);
        std::vector<int> fill(start.begin(), start.end() - 1);
        for (int i = 0; i < n; ++i) ids[fill[cid[i]]++] = i;  // ascending ids per cell
    }
};

double wrap_dx(double dx, double X)  // periodic distance along the parallel-fibre axis
{
    dx = std::fabs(dx);
    return std::min(dx, X - dx);
}

// k distinct cells drawn from the m nearest presynaptic cells of (x, z).
void nearest_pick(const Cells &pre, const Grid &g, double x, double z, int k, int m, Rng &r,
                  std::vector<int32_t> &out)
{
    const int n = (int)pre.x.size();
    m = std::min(m, n);
    k = std::min(k, m);
    std::vector<std::pair<double, int>> cand;
    int ring = 1;
    for (;;) {
        cand.clear();
        int cx = std::min(g.nx - 1, (int)(x / g.cs)), cz = std::min(g.nz - 1, (int)(z / g.cs));
        for (int iz = std::max(0, cz - ring); iz <= std::min(g.nz - 1, cz + ring); ++iz)
            for (int ix = std::max(0, cx - ring); ix <= std::min(g.nx - 1, cx + ring); ++ix) {
                int c = iz * g.nx + ix;
                for (int t = g.start[c]; t < g.start[c + 1]; ++t) {
                    int id = g.ids[t];
                    double dx = pre.x[id] - x, dz = pre.z[id] - z;
                    cand.push_back({dx * dx + dz * dz, id});
                }
            }
        const bool all = (cx - ring <= 0 && cz - ring <= 0 && cx + ring >= g.nx - 1 && cz + ring >= g.nz - 1);
        // the m nearest are certain once the m-th lies within the searched ring
        if ((int)cand.size() >= m) {
            std::nth_element(cand.begin(), cand.begin() + (m - 1), cand.end());
            double rad = std::sqrt(cand[m - 1].first);
            if (rad <= ring * g.cs || all) break;
        }
        if (all) break;
        ++ring;
    }
    std::sort(cand.begin(), cand.end());
    if ((int)cand.size() > m) cand.resize(m);
    // partial Fisher-Yates over the m candidates
    for (int i = 0; i < k; ++i) {
        int j = i + (int)r.below(cand.size() - i);
        std::swap(cand[i], cand[j]);
        out.push_back(cand[i].second);
    }
}

// Presynaptic cells within the parallel-fibre band of (x, z): |dz| < half_band and
// periodic |dx| < half_len, each kept with probability p (geometric skipping).
void band_pick(const Cells &pre, const Grid &g, double x, double z, double half_band, double half_len, double p,
               Rng &r, std::vector<int32_t> &out)
{
    if (p <= 0) return;
    const double lq = p < 1.0 ? std::log1p(-p) : 0.0;
    auto skip = [&]() -> int64_t {
        if (p >= 1.0) return 0;
        double u = r.uni();
        return (int64_t)std::floor(std::log1p(-u) / lq);
    };
    const int iz0 = std::max(0, (int)std::floor((z - half_band) / g.cs));
    const int iz1 = std::min(g.nz - 1, (int)std::floor((z + half_band) / g.cs));
    const double L = std::min(half_len, g.X / 2);
    const int nxw = (int)std::ceil(L / g.cs) + 1;
    std::vector<int> cols;
    int cx = std::min(g.nx - 1, (int)(x / g.cs));
    for (int d = -nxw; d <= nxw; ++d) {
        int ix = ((cx + d) % g.nx + g.nx) % g.nx;
        if (std::find(cols.begin(), cols.end(), ix) == cols.end()) cols.push_back(ix);
    }
    for (int iz = iz0; iz <= iz1; ++iz)
        for (int ix : cols) {
            const int c = iz * g.nx + ix;
            for (int64_t t = g.start[c] + skip(); t < g.start[c + 1]; t += 1 + skip()) {
                const int id = g.ids[t];
                if (std::fabs(pre.z[id] - z) < half_band && wrap_dx(pre.x[id] - x, g.X) < L) out.push_back(id);
            }
        }
    std::sort(out.begin(), out.end());
}

struct Edges {
    std::vector<int64_t> row_ptr;
    std::vector<int32_t> col;
    std::vector<float> val;
};

// Rows are independent (each draws from its own stream, seeded by the row index),
// so they are built in parallel over contiguous row ranges and concatenated in
// order: the result does not depend on the thread count.
template <typename F>
Edges build_rows(int post_n, uint64_t seed, int jitter, F rule)
{
    const int nt = std::max(1, std::min<int>(16, (int)std::thread::hardware_concurrency()));
    const int chunk = (post_n + nt - 1) / std::max(1, nt);
    std::vector<Edges> part(nt);
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) {
        th.emplace_back([&, t]() {
            Edges &E = part[t];
            const int j0 = std::min(post_n, t * chunk), j1 = std::min(post_n, j0 + chunk);
            E.row_ptr.assign(j1 - j0 + 1, 0);
            std::vector<int32_t> tmp;
            for (int j = j0; j < j1; ++j) {
                Rng r(hash3(seed, 0x51, (uint64_t)j));
                tmp.clear();
                rule(j, r, tmp);
                std::sort(tmp.begin(), tmp.end());
                tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
                for (int32_t c : tmp) {
                    E.col.push_back(c);
                    if (jitter) E.val.push_back(0.8f + 0.4f * (float)r.uni());  // U(0.8, 1.2)
                }
                E.row_ptr[j - j0 + 1] = (int64_t)E.col.size();
            }
        });
    }
    for (auto &x : th) x.join();
    Edges E;
    E.row_ptr.assign(post_n + 1, 0);
    size_t nnz = 0;
    for (auto &P : part) nnz += P.col.size();
    E.col.reserve(nnz);
    if (jitter) E.val.reserve(nnz);
    int j = 0;
    for (auto &P : part) {
        const int64_t base = (int64_t)E.col.size();
        for (size_t r = 1; r < P.row_ptr.size(); ++r) E.row_ptr[++j] = base + P.row_ptr[r];
        E.col.insert(E.col.end(), P.col.begin(), P.col.end());
        E.val.insert(E.val.end(), P.val.begin(), P.val.end());
        std::vector<int32_t>().swap(P.col);
        std::vector<float>().swap(P.val);
    }
    return E;
}

}  // namespace

int synth_describe(scaffold_env *env, const scaffold_synth_config *cfg)
{
    if (env->n_types < 7 || env->n_conn < 16 || env->n_inputs < 2)
        return fail(SCAFFOLD_ERR_CAPACITY, "the scaffold needs 7 populations, 16 projections, 2 inputs");
    if (!(cfg->scale > 0)) return fail(SCAFFOLD_ERR_INVALID_ARG, "scale must be positive");
    if (cfg->connectome != 0 && cfg->connectome != 1) return fail(SCAFFOLD_ERR_INVALID_ARG, "connectome is 0 or 1");
    const double side = 400.0 * std::sqrt(cfg->scale);
    const double X = side, Z = side;
    int counts[8];
    const int ref[8] = {88158, 7073, 69, 219, 603, 603, 12, 2400};
    for (int i = 0; i < 8; ++i) {
        counts[i] = cfg->counts[i] > 0 ? cfg->counts[i] : std::max(1, (int)std::lround(ref[i] * cfg->scale));
    }
    int idx[7];
    int rc = register_pops(env, counts, idx);
    if (rc) return rc;
    const uint64_t S = cfg->seed;
    Cells cell[8];
    for (int i = 0; i < 8; ++i) cell[i] = place(counts[i], X, Z, hash3(S, 0xCE11, (uint64_t)i));
    for (int i = 0; i < 7; ++i) {
        Population &P = env->pops[idx[i]];
        P.pos_x.assign(cell[i].x.begin(), cell[i].x.end());
        P.pos_z.assign(cell[i].z.begin(), cell[i].z.end());
    }
    // granule ids by strips (positions stored under the new ids)
    const std::vector<int32_t> grc_id = strip_order(cell[GRC]);
    {
        Population &P = env->pops[idx[GRC]];
        for (int i = 0; i < counts[GRC]; ++i) {
            P.pos_x[grc_id[i]] = cell[GRC].x[i];
            P.pos_z[grc_id[i]] = cell[GRC].z[i];
        }
    }
    const double dens_grc = counts[GRC] / (X * Z);
    Grid g_grc(cell[GRC], 50.0, X, Z);
    auto grid_for = [&](int p, double per_cell) {
        double dens = counts[p] / (X * Z);
        double cs = std::max(10.0, std::sqrt(per_cell / dens));
        return Grid(cell[p], cs, X, Z);
    };
    Grid g_glom = grid_for(GLOM, 8), g_goc = grid_for(GOC, 4), g_sc = grid_for(SC_, 4), g_bc = grid_for(BC, 4);
    const double half_band = 50.0, half_len = 200.0;
    const double band_area = 2 * half_band * 2 * std::min(half_len, X / 2);
    const double band_grc = dens_grc * std::min(band_area, X * Z);
    auto p_for = [&](double target) { return std::min(1.0, target / std::max(1.0, band_grc)); };
    // connectome 1: the mean fan-ins of the reference's own connectome (the sizes
    // of its data/dat_backup/latest connection files): 28.3k parallel fibres per
    // Purkinje cell (the rule's p = 0.8 on a band widened to hold them), 1.6k per
    // Golgi, 1.03k per stellate and 994 per basket cell, 26 Purkinje cells per
    // DCN cell.  The band is cut at the sheet's z edges (it wraps along x only),
    // so a band of half-width h covers 2h - h^2/Z of its length on average over
    // the sheet; the widths and probabilities below are solved for the targets.
    // connectome 0: the rule text (p = 0.8 in a 100 um band, 400 per Golgi and
    // 1000 per stellate / basket cell, Purkinje -> DCN all to all).
    const bool ref_stats = cfg->connectome == 1;
    const double L2 = 2 * std::min(half_len, X / 2);  // band length
    double pc_half_band = half_band;
    if (ref_stats) {  // 0.8 * dens * L2 * (2h - h^2/Z) = 28,300
        const double c = 28300.0 / (0.8 * std::max(1e-12, dens_grc) * L2);
        pc_half_band = c >= Z ? Z : Z * (1.0 - std::sqrt(1.0 - c / Z));
    }
    const double edge = ref_stats ? 1.0 - std::min(half_band, Z) / (2 * Z) : 1.0;  // mean covered fraction
    auto p_band = [&](double target) { return p_for(target / edge); };
    const double goc_target = ref_stats ? 1600.0 : 400.0;
    const double sc_target = ref_stats ? 1030.0 : 1000.0, bc_target = ref_stats ? 994.0 : 1000.0;
    const int J = cfg->weight_jitter;
    const int maxd = cfg->max_delay > 0 ? cfg->max_delay : 4;
    const double aa_r = 50.0;
    auto aa_pick = [&](int post, int k) {
        return [&, post, k](int j, Rng &r, std::vector<int32_t> &out) {
            // ascending-axon contacts: k granule cells from those within aa_r
            std::vector<int32_t> cand;
            band_pick(cell[GRC], g_grc, cell[post].x[j], cell[post].z[j], aa_r, aa_r, 1.0, r, cand);
            std::vector<int32_t> keep;
            for (int32_t c : cand) {
                double dx = cell[GRC].x[c] - cell[post].x[j], dz = cell[GRC].z[c] - cell[post].z[j];
                if (dx * dx + dz * dz < aa_r * aa_r) keep.push_back(c);
            }
            int kk = std::min<int>(k, (int)keep.size());
            for (int i = 0; i < kk; ++i) {
                int t = i + (int)r.below(keep.size() - i);
                std::swap(keep[i], keep[t]);
                out.push_back(keep[i]);
            }
        };
    };
    for (int i = 0; i < 16; ++i) {
        const RefProj &rp = kRefProjs[i];
        const int post = rp.post, pre = rp.pre;
        const uint64_t ps = hash3(S, 0x9F0, (uint64_t)i);
        Edges E;
        switch (rp.type) {
        case 0: E = build_rows(counts[post], ps, J, [&](int j, Rng &r, std::vector<int32_t> &o) {
                    band_pick(cell[GRC], g_grc, cell[post].x[j], cell[post].z[j], pc_half_band, half_len, 0.8, r, o); });
            break;
        case 1: E = build_rows(counts[post], ps, J, [&](int j, Rng &r, std::vector<int32_t> &o) {
                    band_pick(cell[GRC], g_grc, cell[post].x[j], cell[post].z[j], half_band, half_len,
                              p_band(goc_target), r, o); });
            break;
        case 2:
        case 3: {
            const double pb = p_band(rp.type == 2 ? sc_target : bc_target);
            E = build_rows(counts[post], ps, J, [&](int j, Rng &r, std::vector<int32_t> &o) {
                    band_pick(cell[GRC], g_grc, cell[post].x[j], cell[post].z[j], half_band, half_len, pb, r, o); });
            break;
        }
        case 4: E = build_rows(counts[post], ps, J, aa_pick(post, 361)); break;
        case 5: E = build_rows(counts[post], ps, J, aa_pick(post, 250)); break;
        case 6: E = build_rows(counts[post], ps, J, [&](int j, Rng &r, std::vector<int32_t> &o) {
                    nearest_pick(cell[GOC], g_goc, cell[post].x[j], cell[post].z[j], 34, 68, r, o); });
            break;
        case 7: E = build_rows(counts[post], ps, J, [&](int j, Rng &r, std::vector<int32_t> &o) {
                    nearest_pick(cell[SC_], g_sc, cell[post].x[j], cell[post].z[j], 4, 8, r, o); });
            break;
        case 8: E = build_rows(counts[post], ps, J, [&](int j, Rng &r, std::vector<int32_t> &o) {
                    nearest_pick(cell[BC], g_bc, cell[post].x[j], cell[post].z[j], 4, 8, r, o); });
            break;
        case 9: E = build_rows(counts[post], ps, J, [&](int, Rng &r, std::vector<int32_t> &o) {
                    const int k = std::min(counts[GLOM], 375);
                    for (int t = 0; t < k; ++t) o.push_back((int32_t)r.below(counts[GLOM])); });
            break;
        case 10: E = build_rows(counts[post], ps, J, [&](int, Rng &r, std::vector<int32_t> &o) {
                     if (!ref_stats) {
                         for (int t = 0; t < counts[PC]; ++t) o.push_back(t);
                         return;
                     }
                     std::vector<int32_t> all(counts[PC]);  // 26 distinct Purkinje cells
                     for (int t = 0; t < counts[PC]; ++t) all[t] = t;
                     const int k = std::min(counts[PC], 26);
                     for (int t = 0; t < k; ++t) {
                         std::swap(all[t], all[t + (int)r.below(all.size() - t)]);
                         o.push_back(all[t]);
                     } });
            break;
        case 11: E = build_rows(counts[post], ps, J, [&](int j, Rng &r, std::vector<int32_t> &o) {
                     nearest_pick(cell[BC], g_bc, cell[post].x[j], cell[post].z[j], 20, 40, r, o); });
            break;
        case 12: E = build_rows(counts[post], ps, J, [&](int j, Rng &r, std::vector<int32_t> &o) {
                     nearest_pick(cell[SC_], g_sc, cell[post].x[j], cell[post].z[j], 20, 40, r, o); });
            break;
        case 13: E = build_rows(counts[post], ps, J, [&](int j, Rng &r, std::vector<int32_t> &o) {
                     nearest_pick(cell[GLOM], g_glom, cell[post].x[j], cell[post].z[j], 4, 12, r, o); });
            break;
        case 14: E = build_rows(counts[post], ps, J, [&](int j, Rng &r, std::vector<int32_t> &o) {
                     nearest_pick(cell[GLOM], g_glom, cell[post].x[j], cell[post].z[j], 66, 132, r, o); });
            break;
        case 15: E = build_rows(counts[post], ps, J, [&](int j, Rng &r, std::vector<int32_t> &o) {
                     // 2 or 3 Golgi cells per granule cell, mean 2.3 as in the reference's
                     // connectome (206k synapses onto 88k granule cells)
                     const int k = r.uni() < 0.3 ? 3 : 2;
                     nearest_pick(cell[GOC], g_goc, cell[post].x[j], cell[post].z[j], k, 8, r, o); });
            break;
        }
        if (pre == GRC) {  // granule columns to granule ids, each row ascending again
            std::vector<std::pair<int32_t, float>> row;
            for (size_t j = 0; j + 1 < E.row_ptr.size(); ++j) {
                row.clear();
                for (int64_t e = E.row_ptr[j]; e < E.row_ptr[j + 1]; ++e)
                    row.push_back({grc_id[E.col[e]], J ? E.val[e] : 0.f});
                std::sort(row.begin(), row.end());
                for (size_t t = 0; t < row.size(); ++t) {
                    E.col[E.row_ptr[j] + (int64_t)t] = row[t].first;
                    if (J) E.val[E.row_ptr[j] + (int64_t)t] = row[t].second;
                }
            }
        }
        if (post == GRC) {  // rows in granule-id order
            Edges F;
            F.row_ptr.assign(E.row_ptr.size(), 0);
            std::vector<int32_t> old_of(counts[GRC]);
            for (int i = 0; i < counts[GRC]; ++i) old_of[grc_id[i]] = i;
            for (int k = 0; k < counts[GRC]; ++k)
                F.row_ptr[k + 1] = F.row_ptr[k] + (E.row_ptr[old_of[k] + 1] - E.row_ptr[old_of[k]]);
            F.col.reserve(E.col.size());
            if (J) F.val.reserve(E.val.size());
            for (int k = 0; k < counts[GRC]; ++k)
                for (int64_t e = E.row_ptr[old_of[k]]; e < E.row_ptr[old_of[k] + 1]; ++e) {
                    F.col.push_back(E.col[e]);
                    if (J) F.val.push_back(E.val[e]);
                }
            E = std::move(F);
        }
        float delay = rp.delay;
        if (!cfg->reference_delays) delay = (float)std::min(maxd, std::max(1, (int)rp.delay));
        rc = scaffold_set_connectivity_csr(env, rp.type, counts[pre], counts[post], idx[pre], idx[post], rp.w, delay,
                                           E.row_ptr.data(), E.col.data(), J ? E.val.data() : nullptr, nullptr);
        if (rc) return rc;
    }
    // mossy fibres: each glomerulus is fed by its nearest fibre
    const int nmf = counts[MF];
    Grid g_mf = grid_for(MF, 4);
    std::vector<int> owner(counts[GLOM]);
    {
        Rng r(hash3(S, 0x3F, 0));
        std::vector<int32_t> o;
        for (int gi = 0; gi < counts[GLOM]; ++gi) {
            o.clear();
            nearest_pick(cell[MF], g_mf, cell[GLOM].x[gi], cell[GLOM].z[gi], 1, 1, r, o);
            owner[gi] = o.empty() ? 0 : o[0];
        }
    }
    std::vector<int64_t> fan(nmf + 1, 0);
    for (int gi = 0; gi < counts[GLOM]; ++gi) fan[owner[gi] + 1]++;
    for (int m = 0; m < nmf; ++m) fan[m + 1] += fan[m];
    std::vector<uint32_t> tg(counts[GLOM]);
    {
        std::vector<int64_t> f(fan.begin(), fan.end() - 1);
        for (int gi = 0; gi < counts[GLOM]; ++gi) tg[f[owner[gi]]++] = (uint32_t)gi;
    }
    const double inf = std::numeric_limits<double>::infinity();
    scaffold_stim_func bg{};
    bg.poisson_hz = cfg->mf_rate_hz;
    bg.poisson_on = 0.0;
    bg.poisson_off = inf;
    bg.poisson_hz_from_freq = 1;
    bg.seed = cfg->stim_seed;
    rc = scaffold_set_input_stim_ids(env, idx[GLOM], tg.data(), nmf, fan.data(), &bg, nullptr);
    if (rc) return rc;
    int npatch = cfg->burst_n_mf > 0 ? cfg->burst_n_mf : (int)std::lround(cfg->burst_fraction * nmf);
    npatch = std::min(npatch, nmf);
    if (npatch > 0 && cfg->burst_rate_hz > 0) {
        std::vector<std::pair<double, int>> d(nmf);
        for (int m = 0; m < nmf; ++m) {
            double dx = cell[MF].x[m] - X / 2, dz = cell[MF].z[m] - Z / 2;
            d[m] = {dx * dx + dz * dz, m};
        }
        std::sort(d.begin(), d.end());
        std::vector<int> patch;
        for (int i = 0; i < npatch; ++i) patch.push_back(d[i].second);
        std::sort(patch.begin(), patch.end());
        std::vector<int64_t> pf(1, 0);
        std::vector<uint32_t> pt;
        for (int m : patch) {
            for (int64_t t = fan[m]; t < fan[m + 1]; ++t) pt.push_back(tg[t]);
            pf.push_back((int64_t)pt.size());
        }
        scaffold_stim_func bu = bg;
        bu.burst_hz = cfg->burst_rate_hz;
        bu.burst_on = cfg->burst_on_ms;
        bu.burst_off = cfg->burst_off_ms;
        bu.seed = cfg->stim_seed + 1;
        rc = scaffold_set_input_stim_ids(env, idx[GLOM], pt.data(), npatch, pf.data(), &bu, nullptr);
        if (rc) return rc;
    }
    return SCAFFOLD_OK;
}

}  // namespace sc

extern "C" int scaffold_synth_config_preset(int which, scaffold_synth_config *cfg)
{
    if (!cfg) return sc::fail(SCAFFOLD_ERR_INVALID_ARG, "null config");
    scaffold_synth_config c{};
    c.seed = 20211;
    c.weight_jitter = 1;
    c.mf_rate_hz = 4.0f;
    c.burst_rate_hz = 100.0f;
    c.burst_on_ms = 300.0;
    c.burst_off_ms = 350.0;
    c.stim_seed = 0xC0FFEE;
    c.max_delay = 4;
    switch (which) {
    case 0: {
        c.scale = 0.1;
        const int k[8] = {8800, 700, 7, 22, 60, 60, 2, 240};
        memcpy(c.counts, k, sizeof(k));
        c.burst_n_mf = 0;
        c.burst_rate_hz = 0.f;
        break;
    }
    // the reference-size and tiled networks carry the reference connectome's
    // fan-ins (4.23M synapses at reference size, 67M at 16x, 421M at 100x)
    case 1: c.scale = 1.0; c.burst_n_mf = 50; c.connectome = 1; break;
    case 2: c.scale = 16.0; c.burst_n_mf = 800; c.connectome = 1; break;
    case 4: c.scale = 100.0; c.burst_n_mf = 0; c.burst_fraction = 0.05; c.connectome = 1; break;
    default: return sc::fail(SCAFFOLD_ERR_INVALID_ARG, "preset %d: use 0, 1, 2 or 4", which);
    }
    *cfg = c;
    return SCAFFOLD_OK;
}
