/*
 * scaffold.h -- C ABI of the MI355X (gfx950) cerebellar-scaffold exchange-step engine.
 *
 * The entry points mirror the reference simulator's lifecycle API and its
 * network-description plug-ins so that a caller of the reference can switch
 * over call by call.  Each declaration names the reference interface it
 * replaces (paths relative to the reference repository,
 * SimulationModule/Scaffold/ unless stated otherwise).
 *
 * Conventions
 *   - Every function returns SCAFFOLD_OK (0) or a positive SCAFFOLD_ERR_* code;
 *     scaffold_last_error() returns a thread-local message for the last failure.
 *     (The reference returns nothing and calls exit() on errors:
 *     Simulation_header.h:11-18, init_params.cu:33-36.)
 *   - Plain pointers and sizes only.  All device memory is owned by the library.
 *   - Host pointers returned through out-parameters stay valid until the next
 *     call on the same environment.
 *   - An environment is not re-entrant: one caller thread at a time
 *     (as the reference's static state, GPUScaffoldSimulation.cu:697).
 *   - Neuron indices in queries are the reference's global ids: populations in
 *     internal index order, each occupying [base_id, base_id + num)
 *     (init_params.cu:334-341).
 */
#ifndef SCAFFOLD_H
#define SCAFFOLD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define SCAFFOLD_OK                0
#define SCAFFOLD_ERR_INVALID_ARG   1  /* bad argument value or null pointer           */
#define SCAFFOLD_ERR_STATE         2  /* call out of lifecycle order                   */
#define SCAFFOLD_ERR_IO            3  /* file cannot be opened / read                  */
#define SCAFFOLD_ERR_FORMAT        4  /* .dat malformed, unsorted or ids out of range  */
#define SCAFFOLD_ERR_UNSUPPORTED   5  /* feature outside this engine (e.g. >1 device)  */
#define SCAFFOLD_ERR_HIP           6  /* HIP runtime error                             */
#define SCAFFOLD_ERR_NO_DEVICE     7  /* no gfx950 device visible                      */
#define SCAFFOLD_ERR_WEIGHT_RANGE  8  /* weights whose row sums cannot be exact in f64 */
#define SCAFFOLD_ERR_CAPACITY      9  /* a size limit of the engine is exceeded        */

const char *scaffold_last_error(void);
/* Library version string, e.g. "scaffold-mi355x 0.1 gfx950". */
const char *scaffold_version(void);

/* ------------------------------------------------------------ enums/types */
/* DeviceType (struct_enum_def.h:54-58). */
#define SCAFFOLD_NORMAL 0
#define SCAFFOLD_INPUT  1
#define SCAFFOLD_OUTPUT 2

/* Synaptic weight storage (the per-synapse multiplier of w_bar). */
#define SCAFFOLD_WEIGHTS_F32  0
#define SCAFFOLD_WEIGHTS_BF16 1
#define SCAFFOLD_WEIGHTS_FP8  2   /* OCP e4m3 with a per-projection power-of-two scale */

/* Stimulus function expressed as data.  It replaces the device function-pointer
 * table d_pInputFunctions[] (UserInputFunctions.cu:83) and its entries
 * background_noise / tone_stim / puff_stim (UserInputFunctions.cu:19-39), built
 * from PoissonProcess / PeriodicFiring / SinusoidalOscillation
 * (UserInputFunctions.cu:4-14).
 *
 * For a list entry and sub-step time t (ms, float), with uniform r:
 *   fire = (poisson_on  <= t < poisson_off  && r < poisson_hz/1000*0.1)
 *       || (burst_on    <= t < burst_off    && r < burst_hz/1000*0.1)
 *       || (periodic_on <= t < periodic_off(+inclusive) && PeriodicFiring(t, periodic_hz, periodic_start))
 *       || (sine_on     <= t < sine_off     && r < sine(t)/1000*0.1)
 * with the rate of SinusoidalOscillation (UserInputFunctions.cu:12-14), in float:
 *   sine(t) = sine_max_hz/2 * (1 - cos((float)(2*pi*(t/1000)/sine_cycle_s + sine_shift)))
 * (the reference's `mean` argument does not enter its formula).  cos is this
 * library's own float cosine (Cody-Waite reduction by pi/2 and Taylor series in
 * double, rounded once to float), not CUDA's cosf.
 * A rate of 0 disables a component.  Windows are compared in double precision. */
typedef struct {
    float  poisson_hz;
    double poisson_on, poisson_off;
    float  burst_hz;
    double burst_on, burst_off;
    float  periodic_hz;
    float  periodic_start;
    double periodic_on, periodic_off;
    int    periodic_off_inclusive;
    /* 1: scaffold_set_input_freq() overrides poisson_hz for this stimulus.
     * 0: the rate is fixed, as in the reference kernel, which ignores the
     *    SetInputFreq value (UserInputFunctions.cu:89,116). */
    int    poisson_hz_from_freq;
    /* Philox4x32-10 key of this stimulus: curand_init(seed, subsequence=i, 0)
     * per list entry i (GPUScaffoldSimulation.cu:329, Simulation_header.cu:29-34). */
    uint64_t seed;
    /* Sinusoidally modulated Poisson component (0 = off); sine_cycle_s is the
     * period in seconds (finite, non-zero when the component is on), sine_shift
     * the phase in radians. */
    float  sine_max_hz;
    float  sine_cycle_s;
    float  sine_shift;
    double sine_on, sine_off;
} scaffold_stim_func;

/* Fill *out with the reference's stimulus function func_id:
 * 0 = background_noise (UserInputFunctions.cu:19-23), 1 = tone_stim (:25-29),
 * 2 = puff_stim (:33-39; in the reference's alternative table, :84). */
int scaffold_stim_func_reference(int func_id, uint64_t seed, scaffold_stim_func *out);

typedef struct scaffold_env   scaffold_env;
typedef struct scaffold_batch scaffold_batch;

/* ------------------------------------------------------------- lifecycle */
/* CreateNetworkEnv (GPUScaffoldSimulation.h:41, GPUScaffoldSimulation.cu:634-657).
 * num_of_dev must be 1: this engine drives one GPU. */
int scaffold_create_network_env(scaffold_env **env, int num_of_dev,
                                int num_of_neuron_types, int num_of_connectivity_types,
                                int num_of_inputs, int num_of_plasticities);

/* set_neuron_params (init_params.cu:255-303).  NORMAL populations receive the
 * internal indices 0,1,2,... in call order; INPUT/OUTPUT populations receive
 * num_of_neuron_types-1, -2, ... (init_params.cu:259-265).  `channels` holds
 * c_num x {w, E, g_bar, tau} (the variadic doubles of the reference); c_num <= 2.
 * *index_out receives the internal index (the reference's NeuronTypeID[type]). */
int scaffold_set_neuron_params(scaffold_env *env, int type, const char *name,
                               int duplicate, int num,
                               float Cm, float El, float dt_ref, float Ie,
                               float Vr, float Vth, float gL,
                               int device_type, int c_num, const double *channels,
                               int *index_out);

/* set_connectivity_params (init_params.cu:304-332) with the .dat loader
 * GetMaxConv + LoadConnectivityFile_ELL (init_params.cu:13-67,96-128).
 * File format: one "pre<TAB>post" pair per line, population-local ids, sorted by
 * (post, pre) (convertHDF52dat/convert.py:79-81).  preType/postType are internal
 * population indices.  initial_weight > 0 feeds the excitatory buffer, < 0 the
 * inhibitory one with |w| (GPUScaffoldSimulation.cu:732-733).  delay is in ms
 * (whole exchange steps, truncated as (int)delay, :734).  Every stored synapse
 * has weight multiplier 1.0 as in the reference (init_params.cu:46).
 * postNum may exceed the post population size: pairs whose post id lies in
 * [post population size, postNum) are dropped, as the reference's kernels never
 * read those ELL rows (init_params.cu:743 passes granule_cell_num for the Golgi
 * cells); pairs with pre >= preNum or post >= postNum are a format error here
 * (the reference writes out of bounds). */
int scaffold_set_connectivity_params(scaffold_env *env, int type, const char *filename,
                                     int preNum, int postNum, int preType, int postType,
                                     float initial_weight, float delay,
                                     int use_parallel_reduction, int *index_out);

/* Synaptic plasticity: one row of the rule table that Init_Plasticity fills
 * (SynapticPlasticity.cu:20-33; the reference ships the table empty).  Each
 * exchange step, after the neuron update, every synapse of projection `target`
 * is updated as Hebb_plasticity / Teacher_plasticity do
 * (SynapticPlasticity.cu:38-121, invoked at GPUScaffoldSimulation.cu:754):
 *   dw = A0 + A1_J [pre spiked at n - delay]
 *           + A1_I [Hebb: post spiked at n | Teacher: teacher spiked at n - t_delay]
 *           + that spike x sum over s < time_window of A2_JI [pre spiked at n - delay - s]
 * summed in float in that order; Hebb: w = clamp(w + dw, 0, 1); Teacher:
 * w = max(w + dw, 0).  The teacher of a post neuron is the first entry of its
 * row in projection `teacher` (the climbing fibre, SynapticPlasticity.cu:93),
 * which must share the target's post population.  Coefficients are constants
 * (the reference's coefficient functions return constants); slot c of
 * `coefficients` is used when bit c of `present` is set (the reference's NULL
 * entries are clear bits; A2_IJ is unused by both rules).  Rules apply in table
 * order.  A plastic projection keeps f32 multipliers (1.0 where none were
 * given), as do the other weighted projections onto its post population, and
 * updated multipliers are rounded to multiples of 2^grid_exp (see
 * scaffold_projection_info), which keeps every row sum exact; a Teacher rule
 * must not potentiate (its multipliers would have no bound).  The spike ring
 * is grown to delay + time_window + 2 rows where 2 x T_print is shorter.
 * time_window is the number of steps the A2_JI sum covers, as the reference's
 * table field intends ("should be time window"); its kernels use a fixed
 * s_upper_bound of 50 steps instead (SynapticPlasticity.cu:45,86), which
 * time_window = 0 selects (the reference's own table sets 50 as well).
 * Environments with plasticity cannot join a batch.  At most num_of_plasticities
 * rules. */
#define SCAFFOLD_STDP_HEBB    0  /* STDP_RULE (struct_enum_def.h:121) */
#define SCAFFOLD_STDP_TEACHER 1
#define SCAFFOLD_STDP_A0    0    /* STDP_COEFFICIENTS (struct_enum_def.h:120) */
#define SCAFFOLD_STDP_A1_I  1
#define SCAFFOLD_STDP_A1_J  2
#define SCAFFOLD_STDP_A2_JI 3
#define SCAFFOLD_STDP_A2_IJ 4
int scaffold_set_plasticity(scaffold_env *env, int rule, int target, int teacher,
                            const float coefficients[5], uint32_t present, int time_window,
                            int *index_out);

/* Weight-dependent coefficients: the reference's coefficient functions
 * CTYPE f(CTYPE w, CTYPE s) (stdp_Coefficient_t, struct_enum_def.h:119) of the
 * code-generation template, e.g. PF_PC_A1_J = 0.001*(1.0 - w) and
 * PF_PC_A2_TI = -0.01*w (convertHDF52dat/base_file/SynapticPlasticity.cu:14-19).
 * Slot c is the float value of the double expression
 *   c + a * (b - (double)w)
 * evaluated for each synapse with its multiplier w before the update (the
 * template's functions are {0, 0.001, 1.0} and {0, 0.01, 0.0}: -0.01*w and
 * 0.01*(0.0 - w) are the same double).  a = 0 gives the constant c.  Otherwise
 * as scaffold_set_plasticity.  A Teacher rule must stay bounded: every term with
 * a != 0 needs a > 0 (a term that decreases as w grows), a constant term must
 * not be positive, and the potentiating terms' a, each times the number of times
 * it can apply in one step (1, the window for A2_JI), may sum to at most 1; the
 * multipliers then stay below the largest of the initial ones and of the points
 * (c + a*b)/a where the potentiating terms vanish. */
typedef struct {
    double c, a, b;
} scaffold_stdp_affine;
int scaffold_set_plasticity_affine(scaffold_env *env, int rule, int target, int teacher,
                                   const scaffold_stdp_affine coefficients[5], uint32_t present,
                                   int time_window, int *index_out);

/* In-memory variant: CSR by post row.  row_ptr has postNum+1 entries (int64),
 * col holds population-local pre ids, val the per-synapse multipliers of
 * |initial_weight| (NULL = all 1.0, the reference loader's value). */
int scaffold_set_connectivity_csr(scaffold_env *env, int type,
                                  int preNum, int postNum, int preType, int postType,
                                  float initial_weight, float delay,
                                  const int64_t *row_ptr, const int32_t *col, const float *val,
                                  int *index_out);

/* setInputStim (UserInputFunctions.cu:41-70): read one unsigned id per line from
 * id_file; entry i drives neuron id_i of population neuron_index.  Stimuli are
 * applied in registration order; a later stimulus overwrites an earlier one for
 * a shared id (UserInputFunctions.cu:118). */
int scaffold_set_input_stim(scaffold_env *env, int neuron_index, const char *id_file,
                            const scaffold_stim_func *func, int *index_out);

/* In-memory variant.  With fanout_ptr == NULL entry i drives ids[i] (n entries).
 * Otherwise entry i drives ids[fanout_ptr[i] .. fanout_ptr[i+1]) (a mossy fibre
 * feeding several glomeruli); fanout_ptr has n+1 entries. */
int scaffold_set_input_stim_ids(scaffold_env *env, int neuron_index,
                                const uint32_t *ids, int n, const int64_t *fanout_ptr,
                                const scaffold_stim_func *func, int *index_out);

/* Engine options; set before scaffold_initialize_network_env.  Keys:
 *   "neuron_seed"       Philox key of the initial-potential draw (InitParams,
 *                       init_params.cu:343-352); default 0x5eed
 *   "weight_precision"  SCAFFOLD_WEIGHTS_*; default F32
 *   "index_bits"        32 (default) or 16: column indices stored as 16-bit
 *                       offsets from a base device index held once per 64- or
 *                       128-entry group; a population whose groups span 65,535
 *                       or more presynaptic neurons keeps 32-bit indices
 *                       (scaffold_info.synapses_16bit counts the converted ones)
 *   "lane_max_fanin"    populations whose mean total fan-in is <= this are
 *                       gathered one lane per post neuron (sliced ELL); others by a
 *                       team of wavefronts per post neuron (merged CSR row)
 *   "lane_synapses"     team mode: target synapses per lane (sets the team width);
 *                       0 (default): 8 when one step's synapse and state streams
 *                       fit in 64 MB (the step is latency-bound: wide teams),
 *                       else 64 (throughput-bound: more synapses per block)
 *   "graph"             0 (default): launch every step kernel directly; 1: replay
 *                       captured hipGraphs of whole runs (on ROCm 7 the direct
 *                       launches measured as fast or faster at every size: the
 *                       host keeps ahead of 10 us step kernels either way, and
 *                       graph-launched step kernels ran ~2 % longer)
 *   "graph_steps"       exchange steps per run: a run's statistics and counter
 *                       advance are one launch at its end (the length of a
 *                       captured graph with "graph" 1; a batch's runs are always
 *                       replayed as graphs).  The output train of a call is
 *                       gathered by that launch in a graph, else by a launch of
 *                       its own after the call's last step, and the call returns
 *                       once the train is written.
 *   "reset_g"           1: also zero conductances in scaffold_reset_network_env
 *                       (the reference keeps them, GPUScaffoldSimulation.cu:802-854)
 *   "stage_spikes"      1 (default): blocks whose gathers scatter over a bounded
 *                       range of presynaptic neurons first copy those spike words
 *                       into LDS; 0: every gather reads the spike ring in memory
 *   "stage_ratio"       a block is staged when its windows fit in LDS (64 KiB) and
 *                       hold at most this many 32-bit words per 64-byte segment
 *                       its gathers would touch; 0 (default): 8 up to 64 MB of
 *                       streams per step, 32 above
 *   "nontemporal"       -1 (default): the synapse streams (columns, multipliers)
 *                       are loaded nontemporal when one step's exceed 64 MB, so
 *                       they do not evict descriptors and spike rows from L2;
 *                       0: never; 1: always.  Results are identical either way
 *   "word_groups"       n >= 1: a team-mode projection with at least n synapses
 *                       per 64-bit presynaptic spike word, over all its rows
 *                       (the parallel fibres onto Purkinje cells of the
 *                       reference-size scaffold hold 51), is stored as one
 *                       16-byte record per (post neuron, presynaptic word) --
 *                       word index, connection mask, first multiplier -- and its
 *                       multipliers are read only for connected neurons that
 *                       spiked; 0: every synapse in the merged rows; -1
 *                       (default): 2 when one step's streams fit in 64 MB, else
 *                       8 (at most 64).  Results are identical */
int scaffold_set_option(scaffold_env *env, const char *key, int64_t value);
int scaffold_get_option(scaffold_env *env, const char *key, int64_t *value);

/* InitializeNetworkEnv (GPUScaffoldSimulation.cu:659-691): builds the device
 * layout, uploads it, draws the initial state twice (Initialization :381 and the
 * ResetNetworkEnv(env,0) at :689), so trial 0 uses each neuron's second Philox
 * output.  T_print is raised to the maximum delay; T_print must be a multiple of
 * T_cpu_sim (:674-684).  The spike ring has 2*T_print rows (:361), or more
 * where a step reads further back: at least max_delay + 2 rows, and
 * delay + time_window + 2 for a plasticity rule (each step clears the row of the
 * step after it); scaffold_info.ring_rows gives the count. */
int scaffold_initialize_network_env(scaffold_env *env, int T_print, int T_cpu_sim);

/* loop_n_steps (GPUScaffoldSimulation.cu:693-791).  Advances n_steps >= 0
 * exchange steps (1 ms = 10 Euler sub-steps of 0.1 ms each) starting at
 * trial-relative step start_step, and returns when they are done.  As in the
 * reference (:768-783), the output spike train is handed over after every step
 * that completes a window of T_cpu_sim steps (step count n with
 * n % T_cpu_sim == 0): if output_spike_train is not NULL and this call completes
 * at least one window, *output_spike_train receives a library-owned host buffer
 * (pinned) of T_cpu_sim x (number of OUTPUT neurons) bytes holding the
 * OUTPUT-population spikes of the last window completed (row-major by step),
 * gathered on the GPU and copied with one transfer, valid until the next call; a
 * call that completes no window leaves it unchanged.  (The reference hands over
 * the buffer of its host thread that simulates the OUTPUT cells, :777; its rows
 * are T_cpu_sim x all neurons wide.) */
int scaffold_loop_n_steps(scaffold_env *env, int start_step, int n_steps, int trial,
                          const char **output_spike_train);

/* SetInputFreq (GPUScaffoldSimulation.cu:883-890): freq_list[i] is the Poisson
 * rate (Hz) of stimulus i; it takes effect for stimuli registered with
 * poisson_hz_from_freq = 1. */
int scaffold_set_input_freq(scaffold_env *env, const float *freq_list);

/* ResetNetworkEnv (GPUScaffoldSimulation.cu:802-854): redraws the membrane
 * potentials (next Philox output), zeroes dg and refractory counters and ring
 * row 0.  Conductances and stimulus streams carry over, as in the reference. */
int scaffold_reset_network_env(scaffold_env *env, int trial);

/* FinalizeNetworkEnv (GPUScaffoldSimulation.cu:856-880); also frees *env. */
int scaffold_finalize_network_env(scaffold_env *env);

/* ---------------------------------------------------- description helpers */
/* The reference's hard-coded description plug-ins init_neurons_params +
 * init_connectivity_params (init_params.cu:367-766) and InitInputStimulation
 * (UserInputFunctions.cu:72-82), reading the .dat files from data_dir.
 * seeds: neuron_seed applies to InitParams, stim_seed to stimulus i as
 * stim_seed + i. */
int scaffold_describe_reference(scaffold_env *env, const char *data_dir,
                                uint64_t stim_seed);

/* Seeded synthetic scaffold (the network the benchmark configurations name):
 * populations placed on an x-z sheet of side 400*sqrt(scale) um, rule-based
 * connectivity, mossy fibres each feeding a set of glomeruli.  Cells are
 * numbered by z; granule cells by 4 um strips of z, each ordered by x (the
 * parallel-fibre axis), so a parallel-fibre band holds runs of consecutive
 * granule ids. */
typedef struct {
    double   scale;              /* sheet area / (400 um x 400 um)                       */
    int32_t  counts[8];          /* 0 = derive from scale.  Order: granule, glomerulus,
                                    purkinje, golgi, stellate, basket, dcn, mossy fibre */
    uint64_t seed;               /* connectome seed                                      */
    int32_t  weight_jitter;      /* 1: multipliers U(0.8,1.2); 0: all 1.0                */
    float    mf_rate_hz;         /* background rate of every mossy fibre                 */
    float    burst_rate_hz;      /* burst rate of the patch                              */
    double   burst_on_ms, burst_off_ms;
    int32_t  burst_n_mf;         /* patch size (mossy fibres nearest the sheet centre)   */
    double   burst_fraction;     /* used when burst_n_mf == 0                            */
    uint64_t stim_seed;
    int32_t  max_delay;          /* delays are clipped to [1, max_delay] steps           */
    int32_t  reference_delays;   /* 1: use the reference's delays (1..5 ms)              */
    int32_t  connectome;         /* 0: the benchmark rules as written (parallel fibres at
                                    p = 0.8 in a 100 um band, 400 per Golgi cell; 3.03M
                                    synapses at reference size); 1: the mean fan-ins of
                                    the reference's connectome files (28.3k parallel fibres
                                    per Purkinje cell, 1.6k per Golgi cell; 4.07M)        */
} scaffold_synth_config;

/* Presets: which = 0 (small, ~9.9k cells, connectome 0), 1 (reference size,
 * ~96.7k cells), 2 (16x area), 4 (100x area) -- 1, 2 and 4 with connectome 1.
 * Other fields get the benchmark defaults. */
int scaffold_synth_config_preset(int which, scaffold_synth_config *cfg);
int scaffold_describe_synthetic(scaffold_env *env, const scaffold_synth_config *cfg);

/* Write projection `index` in the reference .dat format (pre<TAB>post, sorted). */
int scaffold_write_dat(scaffold_env *env, int index, const char *filename);

/* ----------------------------------------------------------------- queries */
typedef struct {
    int32_t  n_populations, n_projections, n_stimuli;
    int32_t  total_neurons;        /* reference global id space                */
    int32_t  device_neurons;       /* padded to 64 per population on the GPU   */
    int32_t  T_print, T_cpu_sim, ring_rows, max_delay;
    int64_t  total_synapses;
    int64_t  steps_since_init;     /* stimulus Philox draws so far             */
    int32_t  current_step;         /* trial-relative step of the next call     */
    int32_t  n_output_neurons;
    int64_t  device_bytes;         /* device memory held by this environment   */
    int64_t  algorithmic_bytes_per_step; /* nnz*(idx+w bytes) + N*32 + N/8     */
    int32_t  n_tasks, n_groups;
    int64_t  synapses_16bit;       /* synapses stored as 16-bit column offsets */
    int32_t  n_blocks;             /* blocks of the step kernel                */
    int32_t  n_staged_blocks;      /* of which stage spike windows in LDS      */
    int32_t  lane_synapses;        /* team width target in effect (option or by size) */
    int32_t  stage_ratio;          /* staging threshold in effect (option or by size) */
    int64_t  word_groups;          /* word-group records of the team rows      */
    int64_t  word_group_synapses;  /* synapses they hold                        */
    int64_t  stream_bytes_per_step; /* bytes the stored layout streams per step:
                                      every column, header and multiplier of the
                                      gather layouts (padding included), 16 per
                                      word group (not the multipliers read for
                                      spiking connected neurons), N*32 + N/8 */
} scaffold_info;
int scaffold_get_info(scaffold_env *env, scaffold_info *out);

typedef struct {
    char     name[32];
    int32_t  type, device_type, num, base_id, c_num;
    float    Cm, El, dt_ref, Ie, Vr, Vth, gL;
    float    w[2], E[2], g_bar[2], tau[2];
    int32_t  ref_step;
} scaffold_population_info;
int scaffold_get_population(scaffold_env *env, int index, scaffold_population_info *out);
/* Cell positions (um) of population `index` of a synthetic network (x, z: num
 * floats each; either may be NULL).  SCAFFOLD_ERR_STATE for a population without
 * positions (networks described from .dat files).  No counterpart in the
 * reference's simulator: its cell positions stay in the scaffold's HDF5 file. */
int scaffold_get_positions(scaffold_env *env, int index, float *x, float *z);

typedef struct {
    int32_t type, pre, post, pre_num, post_num, delay;
    float   initial_weight, w_bar;
    int64_t nnz;
    int32_t max_conv;          /* largest row (the reference's ELL width)      */
    int32_t lane_mode;         /* 1: gathered one lane per post neuron         */
    float   fp8_scale;         /* power-of-two scale of FP8 multipliers        */
    float   sum_scale;         /* BF16 / FP8: each row sum of the stored
                                  multipliers is multiplied by this (the
                                  generated over the stored multiplier sum of
                                  the projection), else 1                      */
    int32_t plastic;           /* 1: a plasticity rule updates the multipliers */
    int32_t grid_exp;          /* plastic: updated multipliers are rounded to
                                  multiples of 2^grid_exp (keeps row sums exact
                                  in f64); set at initialization               */
} scaffold_projection_info;
int scaffold_get_projection(scaffold_env *env, int index, scaffold_projection_info *out);
/* Copy the CSR of a projection (row_ptr: post_num+1, col, val: nnz).  val is the
 * multiplier as the engine applies it (after BF16/FP8 rounding when selected;
 * for a plastic projection after initialization, its current value on the
 * device). */
int scaffold_get_projection_csr(scaffold_env *env, int index,
                                int64_t *row_ptr, int32_t *col, float *val);

typedef struct {
    int32_t  pop, n_entries, n_targets;
    scaffold_stim_func func;
} scaffold_stim_info;
int scaffold_get_stimulus(scaffold_env *env, int index, scaffold_stim_info *out,
                          int64_t *fanout_ptr /* n_entries+1 or NULL */,
                          uint32_t *targets /* n_targets or NULL */);

/* Neuron state in global-id order (V: total_neurons, g: total_neurons x 2,
 * refr: total_neurons).  Any pointer may be NULL. */
int scaffold_get_state(scaffold_env *env, float *V, float *g, int32_t *refr);
int scaffold_set_state(scaffold_env *env, const float *V, const float *g, const int32_t *refr);

/* Spike row of trial-relative step `step` (one byte per neuron, global ids);
 * the step must be among the last ring_rows - 1 steps (each step clears the ring
 * row of the step after it, which holds the step ring_rows earlier).  The same
 * window applies to scaffold_get_counts, scaffold_get_fired and the text
 * writers. */
int scaffold_get_spike_row(scaffold_env *env, int step, uint8_t *spikes);
/* Per-population spike counts (n_populations) and synaptic events
 * (sum over fired neurons of their out-degree) of step `step`. */
int scaffold_get_counts(scaffold_env *env, int step, int32_t *counts, uint64_t *events);
/* Spike totals per population (n_populations) and the synaptic-event total,
 * accumulated on the GPU since initialization. */
int scaffold_get_totals(scaffold_env *env, uint64_t *spikes, uint64_t *events);
/* Compacted fired-neuron list of step `step` (global ids, ascending), built on
 * the GPU with wavefront ballot/popcount and a prefix sum. */
int scaffold_get_fired(scaffold_env *env, int step, int32_t *ids, int32_t cap, int32_t *n);
/* scaffold_get_fired without allocating device memory (no counterpart in the
 * reference): for callers that cannot allocate -- code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc.  The compaction's count and
 * ids go through a static scratch array of each GPU (16 MiB, held until the
 * process ends), so cap may be at most 4,194,304 ids (SCAFFOLD_ERR_CAPACITY
 * otherwise).  Every call on one GPU shares that array: calls that overlap on
 * one GPU (from several threads or environments at once) race. */
int scaffold_get_fired_noalloc(scaffold_env *env, int step, int32_t *ids, int32_t cap, int32_t *n);
/* Append the spikes of trial-relative steps [step0, step0+n) to filename as the
 * reference's printer thread print_spike_train writes them
 * (GPUScaffoldSimulation.cu:76-84): "%f\t%d\t%d\n" = (step as ms,
 * population-local id, population type), step by step, neurons in global-id
 * order, for every population except the OUTPUT ones (which the reference
 * simulates on the host and prints there).  The last column is the NeuronType
 * the population was registered with: the printer's table is
 * ReverseNeuronTypeID (:179-183, :429). */
int scaffold_write_spike_text(scaffold_env *env, int step0, int n, const char *filename);
/* The OUTPUT-population lines of the same steps as host_update_lif_ch writes
 * them (Simulation_header.cu:456): (step as ms, global id, population type).
 * The reference keeps them in a file of their own (spike_host*.dat) and appends
 * it to the device file at the end of a trial (GPUScaffoldSimulation.cu:826). */
int scaffold_write_output_spike_text(scaffold_env *env, int step0, int n, const char *filename);
/* The reference's PRINT mode (printer threads print_spike_train,
 * GPUScaffoldSimulation.cu:56-89, started at :404-432; host thread
 * SimulationOnCPU :7-53; merge in ResetNetworkEnv :809-846): from this call on,
 * scaffold_loop_n_steps writes the spike files into directory `dir` (the
 * reference's "spike_output") while the simulation runs.  After each step that
 * completes a window of T_cpu_sim steps the window's rows are copied to pinned
 * host memory on a stream of their own and a writer thread formats them, so the
 * steps go on meanwhile.  Window h (steps [h*T_cpu_sim, (h+1)*T_cpu_sim) of the
 * trial) belongs to thread k = (h / (T_print/T_cpu_sim)) % 2:
 *   spike_dev0_<k>.dat  lines of scaffold_write_spike_text, written per whole
 *                       T_print window once it completes;
 *   spike_host<k>.dat   lines of scaffold_write_output_spike_text, per window.
 * scaffold_reset_network_env(env, trial) then writes spike_dev0.dat (the two
 * device files) and spike_result_trial<trial>.dat (spike_dev0.dat, spike_host0.dat,
 * spike_host1.dat), as the reference's `cat` calls do, and starts the four files
 * empty for the next trial.  Lines of a T_print window left incomplete at the
 * reset are dropped.  Steps advanced by scaffold_time_steps are not recorded.
 * dir == NULL stops recording (pending windows are written first).  Rejected for
 * an environment in a batch. */
int scaffold_record_spikes(scaffold_env *env, const char *dir);

/* Timing helper for benchmarks: advance n_steps (continuing the trial) and
 * return the device time measured with HIP events on the engine stream, and the
 * mean duration of one step kernel measured around individual launches. */
int scaffold_time_steps(scaffold_env *env, int n_steps, double *ms_total,
                        double *ms_step_kernel_mean);

/* Step-kernel timer (no counterpart in the reference): on = n >= 1 (at most
 * 1000) clears and starts it, 0 stops it, -1 only reads it.  While it runs, one
 * run of steps in n (scaffold_loop_n_steps, scaffold_time_steps; the first one
 * included) records HIP events on the engine's stream around its step kernels
 * (launched directly, or with "graph" 1 as a replay of a graph that holds the
 * run's step kernels alone; the run's end follows as a direct launch).  *ms
 * receives the measured time, *launches the step kernels it covers (either may
 * be NULL). */
int scaffold_step_timer(scaffold_env *env, int on, double *ms, int64_t *launches);

/* Autotuner (no counterpart in the reference, whose layout is fixed at compile
 * time).  Before scaffold_initialize_network_env: builds private trial engines of
 * this network under candidate layouts -- lane_max_fanin (which populations use
 * one lane per neuron over sliced-ELL slices, i.e. the ELL width) in {16, 48, 128}
 * and lane_synapses (synapses per lane, i.e. wavefronts per team-mode neuron) in
 * {8, 16, 32, 64}, besides the environment's own options -- times n_steps step
 * kernels of each after 50 warm-up steps, then tries, one at a time on the
 * fastest layout, the other column index width (index_bits 32 / 16), the other
 * stage_spikes setting, graph_steps 50 and word groups off (or on at their
 * default threshold when the environment turned them off), and sets the winning
 * options on env (17 trials; ms_default is the time with the environment's own
 * options, and a lane_synapses or word_groups of -1 / 0 picking by size is
 * reported as the value it resolves to for this network, so the result can be
 * applied to other networks as is).  The network's own state, seeds and
 * counters are untouched. */
typedef struct {
    int32_t lane_max_fanin, lane_synapses, graph_steps;
    int32_t n_tried;
    double ms_per_step;     /* of the winner, run ends included */
    double ms_default;      /* of the default options, for comparison */
    int32_t index_bits, stage_spikes;
    int32_t word_groups, reserved;
} scaffold_tune_result;
int scaffold_autotune(scaffold_env *env, int n_steps, scaffold_tune_result *out);
/* One trial of the autotuner under the environment's current options (no
 * counterpart in the reference): before scaffold_initialize_network_env, builds a
 * private engine of the described network, initializes it, runs 50 warm-up steps
 * and times n_steps steps (launched as the "graph" option says); *ms_per_step is
 * their device time per step, *kernel_ms (may be NULL) the mean step-kernel
 * duration.  The environment
 * is left as it was. */
int scaffold_time_layout(scaffold_env *env, int n_steps, double *ms_per_step, double *kernel_ms);

/* --------------------------------------------------------------- batching */
/* Many initialized environments of uneven size advanced together: every step
 * kernel covers all instances, and runs of steps are replayed as hipGraphs. */
int scaffold_batch_create(scaffold_batch **batch);
int scaffold_batch_add(scaffold_batch *batch, scaffold_env *env);
int scaffold_batch_build(scaffold_batch *batch);
int scaffold_batch_run(scaffold_batch *batch, int n_steps);
int scaffold_batch_time(scaffold_batch *batch, int n_steps, double *ms_total);
int scaffold_batch_destroy(scaffold_batch *batch);

/* ------------------------------------------------------------ diagnostics */
/* The multipliers of one projection as the engine stores them with `precision`
 * (SCAFFOLD_WEIGHTS_*), dequantized into out (which may alias in): BF16 rounded to
 * nearest even; FP8-e4m3 (OCP) as in / scale rounded to nearest even, saturating
 * at 448, with the projection's power-of-two scale 2^ceil(log2(max|in| / 448))
 * (*scale).  *sum_scale: sum(in) / sum(out) for BF16 and FP8 (the factor each
 * row sum is multiplied by), else 1.  The planner runs exactly this step on every
 * projection; no counterpart in the reference, whose ELL values are all 1.0. */
int scaffold_encode_multipliers(int precision, const float *in, int64_t n, float *out, float *scale,
                                float *sum_scale);

/* Device copy of `bytes` bytes between two library-allocated scratch buffers;
 * used to calibrate HBM counters against a transfer of known size. */
int scaffold_copy_probe(int64_t bytes, int iters, double *ms);

#ifdef __cplusplus
}
#endif
#endif /* SCAFFOLD_H */
