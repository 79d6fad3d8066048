/*
 * scaffold_oracle.c -- TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
 *
 * CPU restatement of the reference's per-exchange-step algorithm, used only to
 * check the engine by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline.
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, strict IEEE, no FMA).
 *
 * Pinning (checks that would catch a plausible mistake in each part):
 *   orc_philox4x32_10 / orc_philox_u32 : Random123 known-answer vectors and the
 *       stream of two independent open engines (rocRAND philox4x32_10_engine,
 *       PyTorch at::philox_engine) under the same (seed, subsequence, offset),
 *       committed as tests/golden/philox_vectors.json.
 *   orc_uniform : rocRAND's uniform_distribution over the same u32 inputs
 *       (golden file).  Its equality with cuRAND's curand_uniform mapping is
 *       PARITY UNPINNED: cuRAND's mapping is not published as source and
 *       cannot be run here.  All thresholds used are below 2^-8, where
 *       v * 2^-32 is exact, so at most one v per threshold can differ.
 *   orc_lif_step_f32 / _f64 : closed forms -- free-running inter-spike interval
 *       from the geometric Euler recursion, the fixed point V = El + Ie/gL,
 *       conductance decay (1-DT/tau)^k, refractory hold of exactly ref_step
 *       sub-steps, a hand-computed single step.
 *   orc_stim_entry_fires : exact Poisson thresholds (1e-4 per sub-step at 1 Hz),
 *       the 140 Hz periodic spike list of tone_stim computed by hand, window
 *       edges, and the u32 draw pattern (3 blocks per step, 10 words used).
 *   propagation (inside orc_step) : brute-force recount on tiny graphs, the
 *       count*w_bar identity for unit multipliers, independence of order.
 *   orc_step schedule : delay/ring indexing checked by single-spike probes
 *       arriving exactly `delay` steps later.
 */
#include "scaffold_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* option.h:4-7 */
#define ORC_DT 0.1
#define ORC_SEC 1000.
#define ORC_STEP_MAX 10

/* ------------------------------------------------------------ Philox4x32-10 */
/* Salmon, Moraes, Dror, Shaw, "Parallel Random Numbers: As Easy as 1, 2, 3",
 * SC'11: multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key increments
 * 0x9E3779B9 / 0xBB67AE85, 10 rounds.  The reference draws through cuRAND's
 * Philox4_32_10 (Simulation_header.cu:32, UserInputFunctions.cu:104). */
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void orc_philox_block(uint64_t seed, uint64_t subsequence, uint64_t block, uint32_t out[4])
{
    uint32_t ctr[4] = {(uint32_t)block, (uint32_t)(block >> 32),
                       (uint32_t)subsequence, (uint32_t)(subsequence >> 32)};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox4x32_10(ctr, key, out);
}

uint32_t orc_philox_u32(uint64_t seed, uint64_t subsequence, uint64_t offset)
{
    uint32_t o[4];
    orc_philox_block(seed, subsequence, offset >> 2, o);
    return o[offset & 3];
}

float orc_uniform(uint32_t v)
{
    const float inv = 2.3283064365386963e-10f; /* 2^-32, exact in float */
    return inv + (float)v * inv;
}

/* --------------------------------------------------------------- structure */
struct orc_sim {
    int n_pops, n_projs, n_stims;
    orc_pop  *pops;
    orc_proj *projs;
    orc_stim *stims;
    float    *rate_override;   /* per stimulus, SetInputFreq value        */
    int      *base;            /* base id per population                  */
    int      *type_of;         /* population of each global id            */
    int       total;
    int       R;
    uint64_t  neuron_seed;
    int       precision, sum_mode;
    int64_t   init_calls;      /* InitParams calls so far                 */
    int64_t   draws;           /* stimulus draws (steps) since init       */
    /* FP32 state */
    float    *u, *g, *dge, *dgi;
    /* FP64 state */
    double   *u64, *g64, *dge64, *dgi64;
    int32_t  *refr;
    uint8_t  *ring;            /* R x total bytes (GPUScaffoldSimulation.cu:361) */
    int       threads;         /* OpenMP threads of orc_step (1: serial)  */
    uint8_t  *active;          /* driven mode: 1 = simulated here; NULL = all */
    /* plasticity rules and the multipliers they own (per projection, NULL: not plastic) */
    int       n_plast;
    struct orc_plast {
        int rule, target, teacher, window, grid_exp;
        unsigned present, affine;   /* affine bit c: slot c = (float)(ac + aa * (ab - (double)w)) */
        float coef[4];
        double ac[4], aa[4], ab[4];
    } *plast;
    float   **own_val;
};

/* Threads used by orc_step for propagation and the neuron update (both are
 * independent per post neuron, so the result does not depend on the count). */
void orc_set_threads(orc_sim *s, int threads) { s->threads = threads > 0 ? threads : 1; }

int orc_ref_step(float dt_ref)
{
    return (int)((double)dt_ref / ORC_DT);
}

orc_sim *orc_create(int n_pops, const orc_pop *pops, int n_projs, const orc_proj *projs,
                    int n_stims, const orc_stim *stims, int ring_rows,
                    uint64_t neuron_seed, int precision, int sum_mode)
{
    orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
    s->n_pops = n_pops; s->n_projs = n_projs; s->n_stims = n_stims;
    s->pops = (orc_pop *)malloc(sizeof(orc_pop) * (size_t)(n_pops > 0 ? n_pops : 1));
    if (n_pops > 0) memcpy(s->pops, pops, sizeof(orc_pop) * (size_t)n_pops);
    s->projs = (orc_proj *)malloc(sizeof(orc_proj) * (size_t)(n_projs > 0 ? n_projs : 1));
    if (n_projs > 0) memcpy(s->projs, projs, sizeof(orc_proj) * (size_t)n_projs);
    s->stims = (orc_stim *)malloc(sizeof(orc_stim) * (size_t)(n_stims > 0 ? n_stims : 1));
    if (n_stims > 0) memcpy(s->stims, stims, sizeof(orc_stim) * (size_t)n_stims);
    s->rate_override = (float *)calloc((size_t)(n_stims > 0 ? n_stims : 1), sizeof(float));
    for (int i = 0; i < n_stims; ++i) s->rate_override[i] = stims[i].poisson_hz;
    /* set_base_id (init_params.cu:334-341) */
    s->base = (int *)malloc(sizeof(int) * (size_t)(n_pops + 1));
    int b = 0;
    for (int p = 0; p < n_pops; ++p) { s->base[p] = b; b += pops[p].num; }
    s->base[n_pops] = b;
    s->total = b;
    s->type_of = (int *)malloc(sizeof(int) * (size_t)(b > 0 ? b : 1));
    for (int p = 0; p < n_pops; ++p)
        for (int i = 0; i < pops[p].num; ++i) s->type_of[s->base[p] + i] = p;
    s->R = ring_rows;
    s->neuron_seed = neuron_seed;
    s->precision = precision;
    s->sum_mode = sum_mode;
    s->threads = 1;
    size_t N = (size_t)(b > 0 ? b : 1);
    s->refr = (int32_t *)calloc(N, sizeof(int32_t));
    s->ring = (uint8_t *)calloc(N * (size_t)ring_rows, 1); /* defined zero (reference: cudaMalloc) */
    if (precision == 0) {
        s->u = (float *)calloc(N, sizeof(float));
        s->g = (float *)calloc(2 * N, sizeof(float));
        s->dge = (float *)calloc(N, sizeof(float));
        s->dgi = (float *)calloc(N, sizeof(float));
    } else {
        s->u64 = (double *)calloc(N, sizeof(double));
        s->g64 = (double *)calloc(2 * N, sizeof(double));
        s->dge64 = (double *)calloc(N, sizeof(double));
        s->dgi64 = (double *)calloc(N, sizeof(double));
    }
    return s;
}

void orc_destroy(orc_sim *s)
{
    if (!s) return;
    free(s->pops); free(s->projs); free(s->stims); free(s->rate_override);
    free(s->base); free(s->type_of); free(s->refr); free(s->ring);
    free(s->u); free(s->g); free(s->dge); free(s->dgi);
    free(s->u64); free(s->g64); free(s->dge64); free(s->dgi64);
    free(s->active);
    if (s->own_val)
        for (int i = 0; i < s->n_projs; ++i) free(s->own_val[i]);
    free(s->own_val);
    free(s->plast);
    free(s);
}

int orc_total_neurons(const orc_sim *s) { return s->total; }
int orc_base_id(const orc_sim *s, int pop) { return s->base[pop]; }
int64_t orc_steps_since_init(const orc_sim *s) { return s->draws; }

static inline int orc_row(const orc_sim *s, int n)
{
    int r = n % s->R;
    return r < 0 ? r + s->R : r;
}

/* InitParams (init_params.cu:343-352):
 *   u = Vr + (Vth - Vr) * curand_uniform(state[i]);  dg = 0; refr = 0; spike[i] = 0
 * spike[i] for i < total_nn clears ring row 0 only.  Each call consumes the next
 * 32-bit output of neuron i's stream curand_init(seed, i, 0)
 * (GPUScaffoldSimulation.cu:358). */
void orc_init_params(orc_sim *s)
{
    uint64_t off = (uint64_t)s->init_calls;
    for (int i = 0; i < s->total; ++i) {
        const orc_pop *p = &s->pops[s->type_of[i]];
        float U = orc_uniform(orc_philox_u32(s->neuron_seed, (uint64_t)i, off));
        if (s->precision == 0) {
            float span = p->Vth - p->Vr;
            float prod = span * U;
            s->u[i] = p->Vr + prod;
            s->dge[i] = 0.f; s->dgi[i] = 0.f;
        } else {
            s->u64[i] = (double)p->Vr + ((double)p->Vth - (double)p->Vr) * (double)U;
            s->dge64[i] = 0.0; s->dgi64[i] = 0.0;
        }
        s->refr[i] = 0;
        s->ring[i] = 0;
    }
    s->init_calls++;
}

void orc_reset(orc_sim *s) { orc_init_params(s); }

void orc_set_freq(orc_sim *s, const float *freq_list)
{
    for (int i = 0; i < s->n_stims; ++i)
        if (s->stims[i].poisson_hz_from_freq) s->rate_override[i] = freq_list[i];
}

/* ---------------------------------------------------------------- stimulus */
/* PoissonProcess (UserInputFunctions.cu:4-6): r < freq/sec*DT, with freq a
 * float promoted to double. */
static inline int orc_poisson(float r, float freq)
{
    return (double)r < (double)freq / ORC_SEC * ORC_DT;
}

/* PeriodicFiring (UserInputFunctions.cu:8-10):
 *   (int)((time - start - DT)*freq/sec) != (int)((time - start)*freq/sec)
 * time, start, freq are float: (time-start) is a float difference; the first
 * product is in double (DT promotes), the second in float. */
static inline int orc_periodic(float t, float freq, float start)
{
    float d = t - start;
    int a = (int)(((double)d - ORC_DT) * (double)freq / ORC_SEC);
    float fp = d * freq;
    int b = (int)((double)fp / ORC_SEC);
    return a != b;
}

/* The stimulus cosine: Cody-Waite reduction by pi/2 = HI + LO (HI with 33
 * significant bits, so k*HI and x - k*HI are exact for |k| < 2^20; pi/2 - HI - LO
 * < 4e-27), Taylor series of cos and sin on |r| <= pi/4 (truncation < 1e-17), in
 * double with basic operations in a fixed order, rounded once to float.  The
 * reference calls CUDA's cosf, whose bits are not published: parity with it is
 * unpinned; this function is pinned to the float rounding of numpy's double cos. */
float orc_cosf(float x)
{
    const double HI = 0x1.921fb544p+0, LO = 0x1.0b4611a626331p-34, TWO_OVER_PI = 0x1.45f306dc9c883p-1;
    double xd = (double)x;
    double k = rint(xd * TWO_OVER_PI);
    double r = (xd - k * HI) - k * LO;
    double z = r * r;
    double c = 1.0 + z * (-1.0 / 2 + z * (1.0 / 24 + z * (-1.0 / 720 + z * (1.0 / 40320 +
               z * (-1.0 / 3628800 + z * (1.0 / 479001600 + z * (-1.0 / 87178291200.0 +
               z * (1.0 / 20922789888000.0))))))));
    double s = r + r * z * (-1.0 / 6 + z * (1.0 / 120 + z * (-1.0 / 5040 + z * (1.0 / 362880 +
               z * (-1.0 / 39916800 + z * (1.0 / 6227020800.0 + z * (-1.0 / 1307674368000.0 +
               z * (1.0 / 355687428096000.0))))))));
    int q = (int)((long long)k & 3);
    double v = q == 0 ? c : q == 1 ? -s : q == 2 ? -c : s;
    return (float)v;
}

/* SinusoidalOscillation (UserInputFunctions.cu:12-14):
 *   max/2 * (1 - cosf(2*M_PI*(time/sec)/osci + shift))
 * the phase in double (M_PI and sec promote), converted to float for cosf; the
 * rest in float.  puff_stim (:33-39) passes it to PoissonProcess as freq. */
static inline float orc_sine_rate(float t, float max, float cycle, float shift)
{
    double phase = 0x1.921fb54442d18p+2 * ((double)t / ORC_SEC) / (double)cycle + (double)shift;
    return max / 2.f * (1.f - orc_cosf((float)phase));
}

static inline int orc_stim_fn(const orc_stim *st, float rate, float r, float t)
{
    double td = (double)t;
    int f = 0;
    if (rate != 0.f && st->poisson_on <= td && td < st->poisson_off && orc_poisson(r, rate)) f = 1;
    if (st->burst_hz != 0.f && st->burst_on <= td && td < st->burst_off && orc_poisson(r, st->burst_hz)) f = 1;
    if (st->periodic_hz != 0.f && st->periodic_on <= td &&
        (st->periodic_off_inclusive ? td <= st->periodic_off : td < st->periodic_off) &&
        orc_periodic(t, st->periodic_hz, st->periodic_start)) f = 1;
    if (st->sine_max_hz != 0.f && st->sine_on <= td && td < st->sine_off &&
        orc_poisson(r, orc_sine_rate(t, st->sine_max_hz, st->sine_cycle_s, st->sine_shift))) f = 1;
    return f;
}

/* InputStimulation (UserInputFunctions.cu:94-118): three curand_uniform4 blocks
 * per entry and step -> r[0..9] = (x,y,z,w)(x,y,z,w)(x,y); sub-step time
 * t_k = (CTYPE)n + ((CTYPE)k)*DT; spike = OR over the 10 sub-steps. */
int orc_stim_entry_fires(const orc_stim *st, float rate, int entry, int n, int64_t m)
{
    float r[ORC_STEP_MAX];
    for (int b = 0; b < 3; ++b) {
        uint32_t o[4];
        orc_philox_block(st->seed, (uint64_t)entry, (uint64_t)(3 * m + b), o);
        for (int w = 0; w < 4 && 4 * b + w < ORC_STEP_MAX; ++w) r[4 * b + w] = orc_uniform(o[w]);
    }
    int spk = 0;
    for (int k = 0; k < ORC_STEP_MAX; ++k) {
        float t = (float)((double)(float)n + (double)(float)k * ORC_DT);
        spk |= orc_stim_fn(st, rate, r[k], t);
    }
    return spk;
}

/* --------------------------------------------------------------------- LIF */
/* update_lif_ch (Simulation_header.cu:305-381), host twin host_update_lif_ch
 * (:385-465), with each expression promoted exactly as written:
 *   Isyn = sum_c g_c*(u - E_c)                                   (float)
 *   du   = (float)((DT/Cm) * (double)((Ie - gL*(u - El)) - Isyn))
 *   k==0: g_c = (float)((double)g_c + (-(DT/tau_c)*(double)g_c + (double)((g_bar_c*w_c)*dg)))
 *   k>0 : g_c = (float)((double)g_c + (-(DT/tau_c)*(double)g_c))
 *   fired = u > Vth; refr = fired ? ref_step : max(refr-1, 0); u = refr ? Vr : u + du
 * dg is dg_exc when E_c > El, else dg_inh (:347). */
int orc_lif_step_f32(const orc_pop *p, float *u, float *g, int32_t *refr,
                     float dg_exc, float dg_inh)
{
    const double dtc = ORC_DT / (double)p->Cm;
    const int ref_step = orc_ref_step(p->dt_ref);
    float u_ = *u;
    int32_t r_ = *refr;
    int spk = 0;
    for (int k = 0; k < ORC_STEP_MAX; ++k) {
        float Isyn = 0.f;
        for (int c = 0; c < p->c_num; ++c) {
            float diff = u_ - p->E[c];
            float prod = g[c] * diff;
            Isyn = Isyn + prod;
        }
        float leak = p->gL * (u_ - p->El);
        float inner = (p->Ie - leak) - Isyn;
        float du = (float)(dtc * (double)inner);
        for (int c = 0; c < p->c_num; ++c) {
            double dtt = ORC_DT / (double)p->tau[c];
            double decay = (-dtt) * (double)g[c];
            if (k == 0) {
                float gw = p->g_bar[c] * p->w[c];
                float inj = gw * ((p->E[c] > p->El) ? dg_exc : dg_inh);
                g[c] = (float)((double)g[c] + (decay + (double)inj));
            } else {
                g[c] = (float)((double)g[c] + decay);
            }
        }
        int fired = u_ > p->Vth;
        spk |= fired;
        r_ = fired ? ref_step : (r_ > 0 ? r_ - 1 : 0);
        u_ = r_ ? p->Vr : u_ + du;
    }
    *u = u_;
    *refr = r_;
    return spk;
}

/* The same rule with every quantity in double (the FP64 oracle). */
int orc_lif_step_f64(const orc_pop *p, double *u, double *g, int32_t *refr,
                     double dg_exc, double dg_inh)
{
    const double dtc = ORC_DT / (double)p->Cm;
    const int ref_step = orc_ref_step(p->dt_ref);
    double u_ = *u;
    int32_t r_ = *refr;
    int spk = 0;
    for (int k = 0; k < ORC_STEP_MAX; ++k) {
        double Isyn = 0.0;
        for (int c = 0; c < p->c_num; ++c) Isyn = Isyn + g[c] * (u_ - (double)p->E[c]);
        double du = dtc * (((double)p->Ie - (double)p->gL * (u_ - (double)p->El)) - Isyn);
        for (int c = 0; c < p->c_num; ++c) {
            double dtt = ORC_DT / (double)p->tau[c];
            if (k == 0) {
                double gw = (double)p->g_bar[c] * (double)p->w[c];
                g[c] = g[c] + ((-dtt) * g[c] + gw * (((double)p->E[c] > (double)p->El) ? dg_exc : dg_inh));
            } else {
                g[c] = g[c] + (-dtt) * g[c];
            }
        }
        int fired = u_ > (double)p->Vth;
        spk |= fired;
        r_ = fired ? ref_step : (r_ > 0 ? r_ - 1 : 0);
        u_ = r_ ? (double)p->Vr : u_ + du;
    }
    *u = u_;
    *refr = r_;
    return spk;
}

/* --------------------------------------------------------------------- step */
/* One exchange step (loop_n_steps body, GPUScaffoldSimulation.cu:717-765):
 *   1. stimuli in order, entries in order, overwrite (:720-725, UserInputFunctions.cu:118)
 *   2. projections in index order into dg_exc / dg_inh (:726-744), reading ring
 *      row (n - delay) mod R (:735); S_j = sum over the row of val*[pre fired]
 *      (Simulation_header.cu:91-97), dg[j] += S_j * |w_bar| (:98)
 *   3. LIF of every non-INPUT population, writing ring row n mod R and zeroing
 *      dg (:746-751; OUTPUT populations run the same rule on the host,
 *      GPUScaffoldSimulation.cu:41-46, Simulation_header.cu:385-465). */
static void step_impl(orc_sim *s, int n, const uint8_t *ext);
static void plasticity_step(orc_sim *s, int n);

void orc_step(orc_sim *s, int n) { step_impl(s, n, NULL); }

/* Driven mode (a test device, not a reference function): only the neurons with
 * active[id] = 1 are simulated; the spikes of every other neuron at step n are
 * taken from ext[id] (orc_step_driven).  Propagation and the neuron update skip
 * the driven neurons, whose state stays as it is. */
void orc_set_active(orc_sim *s, const uint8_t *active)
{
    free(s->active);
    s->active = NULL;
    if (!active) return;
    s->active = (uint8_t *)malloc((size_t)(s->total > 0 ? s->total : 1));
    memcpy(s->active, active, (size_t)s->total);
}

void orc_step_driven(orc_sim *s, int n, const uint8_t *ext) { step_impl(s, n, ext); }

static void step_impl(orc_sim *s, int n, const uint8_t *ext)
{
    const uint8_t *act = ext ? s->active : NULL;
    const int row = orc_row(s, n);
    uint8_t *out = s->ring + (size_t)row * (size_t)s->total;
    /* 1. stimulus */
    for (int si = 0; si < s->n_stims; ++si) {
        const orc_stim *st = &s->stims[si];
        const int pb = s->base[st->pop];
        for (int i = 0; i < st->n_entries; ++i) {
            int spk = orc_stim_entry_fires(st, s->rate_override[si], i, n, s->draws);
            if (st->fanout_ptr) {
                for (int64_t t = st->fanout_ptr[i]; t < st->fanout_ptr[i + 1]; ++t)
                    out[pb + (int)st->targets[t]] = (uint8_t)spk;
            } else {
                out[pb + (int)st->targets[i]] = (uint8_t)spk;
            }
        }
    }
    s->draws++;
    /* 2. propagation */
    for (int pi = 0; pi < s->n_projs; ++pi) {
        const orc_proj *pr = &s->projs[pi];
        const uint8_t *in = s->ring + (size_t)orc_row(s, n - pr->delay) * (size_t)s->total
                            + (size_t)s->base[pr->pre];
        const int post_base = s->base[pr->post];
        const int post_num = s->pops[pr->post].num;
        const int exc = pr->w_bar > 0;
        const float W = exc ? pr->w_bar : pr->w_bar * (-1.f);
        const double ss = pr->sum_scale == 0.f ? 1.0 : (double)pr->sum_scale;
#pragma omp parallel for schedule(static) num_threads(s->threads) if (s->threads > 1)
        for (int j = 0; j < post_num; ++j) {
            if (act && !act[post_base + j]) continue;
            if (s->precision == 0) {
                float Sf;
                if (s->sum_mode == 0) {
                    double S = 0.0;
                    for (int64_t e = pr->row_ptr[j]; e < pr->row_ptr[j + 1]; ++e)
                        if (in[pr->col[e]]) S += pr->val ? (double)pr->val[e] : 1.0;
                    Sf = (float)(S * ss);
                } else {
                    Sf = 0.f;
                    for (int64_t e = pr->row_ptr[j]; e < pr->row_ptr[j + 1]; ++e)
                        Sf = Sf + ((in[pr->col[e]]) ? (pr->val ? pr->val[e] : 1.f) : 0.f);
                    if (ss != 1.0) Sf = (float)((double)Sf * ss);
                }
                float add = Sf * W;
                float *buf = exc ? s->dge : s->dgi;
                buf[post_base + j] = buf[post_base + j] + add;
            } else {
                double S = 0.0;
                for (int64_t e = pr->row_ptr[j]; e < pr->row_ptr[j + 1]; ++e)
                    if (in[pr->col[e]]) S += pr->val ? (double)pr->val[e] : 1.0;
                double *buf = exc ? s->dge64 : s->dgi64;
                buf[post_base + j] = buf[post_base + j] + S * ss * (double)W;
            }
        }
    }
    /* 3. LIF */
    for (int p = 0; p < s->n_pops; ++p) {
        const orc_pop *pp = &s->pops[p];
        if (pp->kind == ORC_INPUT) continue;
#pragma omp parallel for schedule(static) num_threads(s->threads) if (s->threads > 1)
        for (int i = s->base[p]; i < s->base[p] + pp->num; ++i) {
            int spk;
            if (act && !act[i]) continue;
            if (s->precision == 0) {
                spk = orc_lif_step_f32(pp, &s->u[i], &s->g[2 * (size_t)i], &s->refr[i],
                                       s->dge[i], s->dgi[i]);
                s->dge[i] = 0.f; s->dgi[i] = 0.f;
            } else {
                spk = orc_lif_step_f64(pp, &s->u64[i], &s->g64[2 * (size_t)i], &s->refr[i],
                                       s->dge64[i], s->dgi64[i]);
                s->dge64[i] = 0.0; s->dgi64[i] = 0.0;
            }
            out[i] = (uint8_t)(spk ? 1 : 0);
        }
    }
    /* driven neurons (input populations included) take the supplied spikes */
    if (ext) {
#pragma omp parallel for schedule(static) num_threads(s->threads) if (s->threads > 1)
        for (int i = 0; i < s->total; ++i)
            if (!act || !act[i]) out[i] = ext[i];
    }
    /* 4. plasticity (invoke_stdp_plasticity, GPUScaffoldSimulation.cu:754) */
    if (s->n_plast > 0) plasticity_step(s, n);
}

/* Plasticity.  Pinning: tests/test_oracle_plasticity.py drives spike trains
 * (driven mode) and checks the multipliers against dw counted by hand from the
 * rule text (A0, A1_J, A1_I, the A2_JI window count), both clamps, the window
 * edges, the ring wrap, the teacher's delay and an empty teacher row, and the
 * grid rounding. */
int orc_add_plasticity(orc_sim *s, int rule, int target, int teacher, const float coef[4],
                       unsigned present, int window, int grid_exp)
{
    if (target < 0 || target >= s->n_projs || (rule == 1 && (teacher < 0 || teacher >= s->n_projs))) return 1;
    struct orc_plast *p = (struct orc_plast *)realloc(s->plast, sizeof(*p) * (size_t)(s->n_plast + 1));
    if (!p) return 1;
    s->plast = p;
    struct orc_plast *r = &s->plast[s->n_plast++];
    r->rule = rule; r->target = target; r->teacher = teacher; r->window = window;
    r->grid_exp = grid_exp; r->present = present; r->affine = 0u;
    for (int c = 0; c < 4; ++c) { r->coef[c] = coef[c]; r->ac[c] = r->aa[c] = r->ab[c] = 0.0; }
    if (!s->own_val) s->own_val = (float **)calloc((size_t)s->n_projs, sizeof(float *));
    orc_proj *pr = &s->projs[target];
    if (!s->own_val[target]) {
        const int64_t nnz = pr->row_ptr[s->pops[pr->post].num];
        float *v = (float *)malloc(sizeof(float) * (size_t)(nnz > 0 ? nnz : 1));
        for (int64_t e = 0; e < nnz; ++e) v[e] = pr->val ? pr->val[e] : 1.f;
        s->own_val[target] = v;
        pr->val = v;
    }
    return 0;
}

/* Weight-dependent coefficients of rule `index` (the template's coefficient
 * functions, convertHDF52dat/base_file/SynapticPlasticity.cu:14-19): bit c of
 * `affine` makes slot c the float value of the double expression
 * ac[c] + aa[c] * (ab[c] - (double)w) for the synapse's multiplier w before the
 * update.  Pinning: tests/test_oracle_plasticity.py (the fixed point the rule
 * converges to, and the template's functions evaluated independently). */
int orc_set_plasticity_affine(orc_sim *s, int index, unsigned affine, const double *ac, const double *aa,
                              const double *ab)
{
    if (index < 0 || index >= s->n_plast) return 1;
    struct orc_plast *r = &s->plast[index];
    r->affine = affine & 15u;
    for (int c = 0; c < 4; ++c) { r->ac[c] = ac[c]; r->aa[c] = aa[c]; r->ab[c] = ab[c]; }
    return 0;
}

static float plast_coef(const struct orc_plast *P, int c, float w)
{
    return ((P->affine >> c) & 1u) ? (float)(P->ac[c] + P->aa[c] * (P->ab[c] - (double)w)) : P->coef[c];
}

void orc_get_multipliers(const orc_sim *s, int proj, float *out)
{
    const orc_proj *pr = &s->projs[proj];
    const int64_t nnz = pr->row_ptr[s->pops[pr->post].num];
    for (int64_t e = 0; e < nnz; ++e) out[e] = pr->val ? pr->val[e] : 1.f;
}

/* Hebb_plasticity / Teacher_plasticity (SynapticPlasticity.cu:38-121) after step n:
 * dw summed in float in the kernels' order, then the clamp (Hebb [0, 1], :73;
 * Teacher >= 0, :117) and the grid rounding. */
static void plasticity_step(orc_sim *s, int n)
{
    const int row = orc_row(s, n);
    const uint8_t *cur = s->ring + (size_t)row * (size_t)s->total;
    for (int r = 0; r < s->n_plast; ++r) {
        const struct orc_plast *P = &s->plast[r];
        const orc_proj *pr = &s->projs[P->target];
        float *val = s->own_val[P->target];
        const int pre_base = s->base[pr->pre], post_base = s->base[pr->post];
        const int post_num = s->pops[pr->post].num;
        const uint8_t *pre_row = s->ring + (size_t)orc_row(s, n - pr->delay) * (size_t)s->total;
        const orc_proj *tp = P->rule == 1 ? &s->projs[P->teacher] : NULL;
        const uint8_t *t_row = tp ? s->ring + (size_t)orc_row(s, n - tp->delay) * (size_t)s->total : NULL;
        for (int j = 0; j < post_num; ++j) {
            int s_post;
            if (!tp) {
                s_post = cur[post_base + j];
            } else {
                const int64_t t0 = tp->row_ptr[j], t1 = tp->row_ptr[j + 1];
                s_post = t0 < t1 ? t_row[s->base[tp->pre] + tp->col[t0]] : 0;
            }
            for (int64_t e = pr->row_ptr[j]; e < pr->row_ptr[j + 1]; ++e) {
                const int pre = pre_base + pr->col[e];
                const int s_pre = pre_row[pre];
                const float w0 = val[e];
                float dw = 0.f;
                if (P->present & 1u) dw += plast_coef(P, 0, w0);
                if ((P->present & 4u) && s_pre) dw += plast_coef(P, 2, w0);
                if ((P->present & 2u) && s_post) dw += plast_coef(P, 1, w0);
                if ((P->present & 8u) && s_post)
                    for (int k = 0; k < P->window; ++k)
                        if (s->ring[(size_t)orc_row(s, n - pr->delay - k) * (size_t)s->total + (size_t)pre])
                            dw += plast_coef(P, 3, w0);
                float w = w0 + dw;
                if (P->rule == 0) w = w < 0.f ? 0.f : (w > 1.f ? 1.f : w);
                else w = w < 0.f ? 0.f : w;
                val[e] = ldexpf(rintf(ldexpf(w, -P->grid_exp)), P->grid_exp);
            }
        }
    }
}

void orc_run(orc_sim *s, int n0, int k)
{
    for (int i = 0; i < k; ++i) orc_step(s, n0 + i);
}

void orc_get_state(const orc_sim *s, float *V, float *g, int32_t *refr)
{
    size_t N = (size_t)s->total;
    if (s->precision == 0) {
        if (V) memcpy(V, s->u, N * sizeof(float));
        if (g) memcpy(g, s->g, 2 * N * sizeof(float));
    } else {
        for (size_t i = 0; i < N; ++i) {
            if (V) V[i] = (float)s->u64[i];
            if (g) { g[2 * i] = (float)s->g64[2 * i]; g[2 * i + 1] = (float)s->g64[2 * i + 1]; }
        }
    }
    if (refr) memcpy(refr, s->refr, N * sizeof(int32_t));
}

void orc_set_state(orc_sim *s, const float *V, const float *g, const int32_t *refr)
{
    size_t N = (size_t)s->total;
    if (s->precision == 0) {
        if (V) memcpy(s->u, V, N * sizeof(float));
        if (g) memcpy(s->g, g, 2 * N * sizeof(float));
    } else {
        for (size_t i = 0; i < N; ++i) {
            if (V) s->u64[i] = V[i];
            if (g) { s->g64[2 * i] = g[2 * i]; s->g64[2 * i + 1] = g[2 * i + 1]; }
        }
    }
    if (refr) memcpy(s->refr, refr, N * sizeof(int32_t));
}

void orc_get_state_f64(const orc_sim *s, double *V, double *g, int32_t *refr)
{
    size_t N = (size_t)s->total;
    for (size_t i = 0; i < N; ++i) {
        if (s->precision == 0) {
            if (V) V[i] = s->u[i];
            if (g) { g[2 * i] = s->g[2 * i]; g[2 * i + 1] = s->g[2 * i + 1]; }
        } else {
            if (V) V[i] = s->u64[i];
            if (g) { g[2 * i] = s->g64[2 * i]; g[2 * i + 1] = s->g64[2 * i + 1]; }
        }
    }
    if (refr) memcpy(refr, s->refr, N * sizeof(int32_t));
}

const uint8_t *orc_spike_row(const orc_sim *s, int n)
{
    return s->ring + (size_t)orc_row(s, n) * (size_t)s->total;
}
