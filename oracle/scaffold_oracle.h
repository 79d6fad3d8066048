/*
 * scaffold_oracle.h -- TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
 *
 * A plain-C, single-threaded CPU restatement of the reference simulator's
 * per-exchange-step algorithm.  Only the tests, the smoke check and the
 * benchmark's CPU-baseline leg load it, and only to check results.  The product
 * (the HIP engine) never links or calls it.
 *
 * Every function cites the reference lines it restates (paths relative to the
 * reference repository, SimulationModule/Scaffold/ unless stated otherwise).
 *
 * Pinning status (which independent checks pin each part) is listed in
 * scaffold_oracle.c next to each function.
 */
#ifndef SCAFFOLD_ORACLE_H
#define SCAFFOLD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ Philox4x32-10 */
/* One Philox4x32-10 block (Salmon et al., SC'11).  ctr[4], key[2] -> out[4]. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* The u32 number `offset` of the stream (seed, subsequence), with the layout of
 * curand_init(seed, subsequence, offset): key = seed, counter = (block, subsequence)
 * where block = offset/4 and the word is offset%4. */
uint32_t orc_philox_u32(uint64_t seed, uint64_t subsequence, uint64_t offset);
/* u32 -> float uniform in (0,1]: 2^-32 + v*2^-32 in float (the rocRAND mapping,
 * stated in place of cuRAND's, which is not restated here). */
float orc_uniform(uint32_t v);
/* cos of a float argument rounded to float; the stimulus's own cosine in place
 * of cosf in SinusoidalOscillation (UserInputFunctions.cu:12-14) */
float orc_cosf(float x);

/* ----------------------------------------------------------- description */
enum { ORC_NORMAL = 0, ORC_INPUT = 1, ORC_OUTPUT = 2 };

typedef struct {
    int   num, kind;
    float Cm, El, dt_ref, Ie, Vr, Vth, gL;
    int   c_num;
    float w[2], E[2], g_bar[2], tau[2];
} orc_pop;

typedef struct {
    int   pre, post;          /* population indices                            */
    float w_bar;              /* initial_weight: sign selects exc/inh buffer   */
    int   delay;              /* exchange steps                                */
    const int64_t *row_ptr;   /* post population size + 1                      */
    const int32_t *col;       /* population-local pre ids                      */
    const float   *val;       /* per-synapse multiplier, NULL = 1.0            */
    float sum_scale;          /* factor on each row sum (0 is read as 1); the
                                 engine's correction for quantized multipliers,
                                 1 in the reference                            */
} orc_proj;

typedef struct {
    int    pop;
    float  poisson_hz;  double poisson_on, poisson_off;
    float  burst_hz;    double burst_on, burst_off;
    float  periodic_hz; float periodic_start; double periodic_on, periodic_off;
    int    periodic_off_inclusive;
    int    poisson_hz_from_freq;
    float  sine_max_hz; float sine_cycle_s; float sine_shift; double sine_on, sine_off;
    uint64_t seed;
    int    n_entries;
    const int64_t  *fanout_ptr;  /* NULL: entry i drives targets[i]             */
    const uint32_t *targets;     /* population-local ids                        */
} orc_stim;

/* precision: 0 = the reference's FP32 semantics (mixed f32/f64 exactly as the
 * reference's expressions promote), 1 = everything in FP64.
 * sum_mode (FP32 only): 0 = row sum taken exactly (in f64) then rounded once to
 * f32; 1 = f32 accumulation in ascending synapse order (the reference's literal
 * loop, Simulation_header.cu:91-97).  With unit multipliers both are identical. */
typedef struct orc_sim orc_sim;
orc_sim *orc_create(int n_pops, const orc_pop *pops, int n_projs, const orc_proj *projs,
                    int n_stims, const orc_stim *stims, int ring_rows,
                    uint64_t neuron_seed, int precision, int sum_mode);
void orc_destroy(orc_sim *s);

int  orc_total_neurons(const orc_sim *s);
int  orc_base_id(const orc_sim *s, int pop);
/* InitParams (init_params.cu:343-352): next u32 of each neuron's stream. */
void orc_init_params(orc_sim *s);
/* ResetNetworkEnv (GPUScaffoldSimulation.cu:802-854): InitParams only. */
void orc_reset(orc_sim *s);
/* One exchange step at trial-relative step n (GPUScaffoldSimulation.cu:717-765). */
void orc_step(orc_sim *s, int n);
/* Steps n0 .. n0+k-1. */
void orc_run(orc_sim *s, int n0, int k);
void orc_set_threads(orc_sim *s, int threads);
/* Driven mode (a test device with no reference counterpart): only neurons with
 * active[id] = 1 (total_neurons bytes, copied; NULL: all) are simulated;
 * orc_step_driven takes the spikes of step n of every other neuron from ext
 * (total_neurons bytes).  Used to run a sub-network in FP64 inside a larger
 * network simulated elsewhere. */
void orc_set_active(orc_sim *s, const uint8_t *active);
void orc_step_driven(orc_sim *s, int n, const uint8_t *ext);
void orc_set_freq(orc_sim *s, const float *freq_list);
/* Plasticity (Hebb_plasticity / Teacher_plasticity, SynapticPlasticity.cu:38-121;
 * rule 0 Hebb, 1 Teacher), applied after every step to the multipliers of
 * projection `target`, rules in the order added.  coef[c] (A0, A1_I, A1_J, A2_JI)
 * is used when bit c of `present` is set.  The teacher of post neuron j is the
 * first entry of row j of projection `teacher`.  Updated multipliers are rounded
 * to multiples of 2^grid_exp (the engine's rule that keeps row sums exact; the
 * reference stores w + dw unrounded).  The oracle keeps its own copy of the
 * target's multipliers (1.0 where val is NULL). */
int  orc_add_plasticity(orc_sim *s, int rule, int target, int teacher, const float coef[4],
                        unsigned present, int window, int grid_exp);
/* Weight-dependent coefficients of rule `index`: bit c of `affine` makes slot c
 * (float)(ac[c] + aa[c] * (ab[c] - (double)w)). */
int orc_set_plasticity_affine(orc_sim *s, int index, unsigned affine, const double *ac, const double *aa,
                              const double *ab);
/* Current multipliers of projection `proj` (nnz floats). */
void orc_get_multipliers(const orc_sim *s, int proj, float *out);

/* State access (FP32 oracle: float arrays; FP64 oracle: the *_f64 calls). */
void orc_get_state(const orc_sim *s, float *V, float *g, int32_t *refr);
void orc_set_state(orc_sim *s, const float *V, const float *g, const int32_t *refr);
void orc_get_state_f64(const orc_sim *s, double *V, double *g, int32_t *refr);
/* Spike bytes of ring row for step n (total_neurons bytes). */
const uint8_t *orc_spike_row(const orc_sim *s, int n);
int64_t orc_steps_since_init(const orc_sim *s);

/* Single-neuron exchange step in FP32 reference semantics (update_lif_ch body,
 * Simulation_header.cu:315-377), exposed for closed-form tests.  Returns the
 * spike flag; updates *u, g[0..c_num), *refr. */
int orc_lif_step_f32(const orc_pop *p, float *u, float *g, int32_t *refr,
                     float dg_exc, float dg_inh);
int orc_lif_step_f64(const orc_pop *p, double *u, double *g, int32_t *refr,
                     double dg_exc, double dg_inh);
/* ref_step = (int)(dt_ref / DT) with DT the double 0.1 (Simulation_header.cu:322). */
int orc_ref_step(float dt_ref);
/* Spike flag of one stimulus entry for trial step n and draw index m
 * (InputStimulation, UserInputFunctions.cu:94-118). */
int orc_stim_entry_fires(const orc_stim *st, float rate_override, int entry, int n, int64_t m);

#ifdef __cplusplus
}
#endif
#endif
