// device_types.hpp -- data layout shared by the host planner and the gfx950 kernels.
//
// Layout in HBM (one network instance):
//   * Neurons live in "device order": populations in internal index order, each
//     padded to a multiple of 64, so one wavefront (64 lanes) covers one
//     64-neuron group of a single population.
//   * Neuron state is SoA: V[Ndev] f32, g[2*Ndev] f32 (exc, inh interleaved),
//     refr[Ndev] i32.
//   * Spike ring: R rows x Wr 64-bit words.  Word g of row r holds the spikes of
//     group g at the step s with s mod R == r.  The last word of every row stays
//     zero; padding entries of sliced-ELL slices point into it.
//   * Synapses are gathered in one of two layouts, chosen per post population:
//       lane mode: sliced ELL per projection, one slice per 64 post neurons,
//                  column-major inside the slice (lane l reads entry k of its row
//                  at slice + 64*k + l), so every wavefront load is one contiguous
//                  256-byte run; one lane per post neuron;
//       team mode: one merged CSR row per post neuron with the synapses of all its
//                  incoming projections (projection after projection, in index
//                  order), and the word groups of its dense projections (below);
//                  a team of 1..16 wavefronts strides over both.
//     Column entries are device bit indices of the presynaptic neuron.
//   * Row sums are accumulated in f64 (team rows: LDS atomics on f64 per
//     projection).  With multipliers that pass the exactness check every sum is
//     exact, so the order in which lanes and wavefronts combine them cannot change
//     the result; no floating-point atomics touch global memory.
#pragma once
#include <stdint.h>

#define SC_WAVE     64
#define SC_MAX_IN   8      // incoming projections per population
#define SC_MAX_POPS 32

// weight storage
#define SC_W_F32  0
#define SC_W_BF16 1
#define SC_W_FP8  2
#define SC_W_UNIT 3

// gather layout of a projection (that of its post population)
#define SC_MODE_LANE 0
#define SC_MODE_ROW  1   // team mode

// task kinds
#define SC_TASK_STIM 0
#define SC_TASK_LANE 1
#define SC_TASK_ROW  2   // one team-mode neuron

struct ScPop {
    double  dtc;           // DT / Cm
    double  dtt[2];        // DT / tau_c
    float   El, Vr, Vth, Ie, gL;
    float   E[2];
    float   gw[2];         // g_bar_c * w_c (float product, as the reference writes it)
    float   Vr_span;       // Vth - Vr (float)
    int32_t exc_ch[2];     // 1: channel c reads dg_exc (E_c > El)
    int32_t c_num;
    int32_t ref_step;
    int32_t kind;          // 0 NORMAL, 1 INPUT, 2 OUTPUT
    int32_t dev_base;      // first device index (multiple of 64)
    int32_t group_base;    // dev_base / 64
    int32_t ref_base;      // base id in the reference's global id space
    int32_t num;
    int32_t n_in;
    int32_t in_proj[SC_MAX_IN];
};

struct ScProj {            // host-side summary of one projection
    float   W;             // |w_bar|
    float   scale;         // FP8 multiplier scale (a power of two), else 1
    float   sum_scale;     // BF16 / FP8: generated over stored multiplier sum, else 1
    int32_t delay;
    int32_t exc;
    int32_t mode;          // SC_MODE_LANE / SC_MODE_ROW, -1: not registered
    int32_t wprec;
    int32_t pre_pop, post_pop;
};

struct ScGroup {
    int32_t pop;
    int32_t first_local;   // population-local id of lane 0
    int32_t nvalid;        // valid lanes
    int32_t self_final;
};

struct ScStim {
    uint32_t key0, key1;
    double  p_thr;               // Poisson threshold rate/1000*0.1 (0 = off)
    double  poisson_on, poisson_off;
    double  b_thr;               // burst threshold (0 = off)
    double  burst_on, burst_off;
    double  periodic_on, periodic_off;
    float   periodic_hz, periodic_start;
    int32_t periodic_incl;
    float   sine_max;            // sinusoidal Poisson peak rate, Hz (0 = off)
    double  sine_on, sine_off;
    float   sine_cycle, sine_shift;   // period (s), phase (rad)
};

struct ScCtl {
    int64_t n;      // trial-relative step of launch offset k = 0
    int64_t m;      // stimulus draws since init (3 Philox blocks per draw)
    int32_t row;    // ring row of step n (n mod R)
    int32_t recount;  // 1: the ring changed outside the step sequence (initialization,
                      // reset, a jump of the step count); the plasticity window counts
                      // of the next run are recomputed from the ring, not updated
};

// 64-neuron groups per wavefront of the plasticity window-count class
#define SC_HIST_GROUPS 8

// Team-mode column words: projection index (segment of the merged row) in the top
// bits, presynaptic device bit index below.
#define SC_Q_SHIFT  29
#define SC_BIT_MASK 0x1FFFFFFFu

// Word groups: a team-mode projection with enough synapses per 64-bit spike word
// of its presynaptic neurons (over all its rows) is held as one
// 16-byte record per (post neuron, presynaptic word): x = slot << SC_Q_SHIFT |
// word index, y = index of the first multiplier, z/w = mask of the connected
// neurons of the word (bits 0-31 / 32-63), their multipliers in ascending bit
// order.  A gather reads the record and the spike word, and multipliers only
// for the set bits of (spikes & mask).

// Planner form of one unit of work.
struct ScTask {
    int32_t kind;
    int32_t inst;
    int32_t group;      // instance-wide group index
    int32_t proj;       // team mode: population
    int32_t post;       // team mode: population-local neuron id
    int32_t pad_[3];
};

// ---- lowered work descriptors: each holds absolute device pointers, so a
// wavefront reaches its data after one descriptor load instead of a chain of
// table lookups (every dependent global load costs on the order of a microsecond).
#define SC_MAX_SLICES 3    // projections of a lane-mode population
#define SC_MAX_TEAM   16   // wavefronts per team-mode neuron (one 1024-thread block)
#define SC_BLOCK      1024 // threads per block of the step kernel
#define SC_LANE_PER_BLOCK 16  // lane-mode groups per block (the other wavefronts idle)
#define SC_STIM_PER_BLOCK 16  // stimulus groups per block

// 16-bit column form: columns are stored as 16-bit offsets from base device bits
// held in headers (lane slices: one per column pair; team rows: one per 8-entry
// chunk); 0xFFFF marks padding.
#define SC_PAD16 0xFFFFu

struct ScSlice {           // one lane-mode projection of one 64-neuron group
    const void *col;       // slice columns, 64 x width, column-major (32-bit bits or 16-bit offsets)
    const uint32_t *hdr;   // 16-bit form: base device bit of each column k (nullptr: 32-bit)
    const void *val;       // slice multipliers (nullptr: unit)
    int32_t width, delay, wprec, exc;
    float W, scale;
    int32_t lds;           // staged blocks: LDS word that stands for word 0 of the ring row
    int32_t pad_;
};

struct ScLaneTask {        // 64 neurons of a lane-mode population
    int32_t inst, group, pop, n_sl;
    int32_t nvalid;
    int32_t wp;            // storage of the slices that have multipliers (SC_W_UNIT: none)
    int32_t idx16;         // slices in the 16-bit column form
    int32_t pad2;
    ScSlice sl[SC_MAX_SLICES];
};

struct ScStimTask {        // 64 neurons of an INPUT population
    int32_t inst, group, pop, nvalid;
};

// ScTeamIn.lds of a projection that a staged block reads from the ring in
// memory (a block of 16-bit team rows whose windows did not all fit in LDS).
#define SC_LDS_GLOBAL ((int32_t)0x80000000)

struct ScTeamIn {          // one incoming projection of a team-mode neuron
    int32_t end;           // end of its segment in the merged row (32-bit form; informational)
    int32_t delay, exc;
    int32_t lds;           // staged blocks: LDS word that stands for word 0 of the ring row,
                           // or SC_LDS_GLOBAL
    float W, scale;
};

// Spike-bitmap windows staged in LDS by one block before its gathers: words
// [w0, w0 + nw) of ring row (n - delay) of instance inst go to LDS words
// [off, off + nw); w0, nw and off are multiples of 4 (16-byte copies).  LDS words
// 0..3 are kept zero (padding entries read word 0).
#define SC_MAX_WIN   8
#define SC_WIN_WORDS 16384   // 64 KB of LDS per block
struct ScWin {
    int32_t inst, delay, w0, nw, off, pad_;
};
struct ScStage {
    int32_t n, pad_;
    ScWin w[SC_MAX_WIN];
};
#define SC_STAGE_DWORDS ((int)(sizeof(ScStage) / 4))  // at most 64: one dword per lane
static_assert(sizeof(ScStage) / 4 <= 64, "a block's window table is loaded with one dword per lane");

struct ScTeamTask {        // one neuron of a team-mode population
    const void *col;       // merged row columns (32-bit words, or 16-bit offsets)
    const uint32_t *hdr;   // 16-bit form: group headers of the row (nullptr: 32-bit)
    const void *val;       // merged row multipliers (nullptr: unit); the word groups' follow
    const uint32_t *grp;   // the neuron's word groups (4 words each; nullptr: none)
    int32_t inst, group, bit, pop;
    int32_t n_in, wprec, len, n_grp;
    ScTeamIn in[SC_MAX_IN];
};

// Block layout of the step kernel (a kernel argument).  Blocks are grouped in
// classes: team neurons of width 16, 8, 4, 2, 1 wavefronts (classes 0..4, with
// 16/width neurons per block), lane groups (class 5), stimulus groups (class 6)
// and the window counts of the plasticity rules (class 7: one 64-neuron group of
// a rule's presynaptic population per wavefront), 16 wavefronts per block.
#define SC_CLASSES    8
#define SC_CLASS_LANE 5
#define SC_CLASS_STIM 6
#define SC_CLASS_HIST 7
struct ScLayout {
    int32_t cls_end[SC_CLASSES];    // cumulative block count at the end of each class
    int32_t cls_ntask[SC_CLASSES];  // tasks of each class
    int32_t cls_task0[SC_CLASSES];  // first descriptor index of each class
};

// State of one device neuron: one 16-byte record, read and written by one load
// and one store per exchange step.
struct alignas(16) ScState {
    float   V, g_exc, g_inh;
    int32_t refr;
};

// One plasticity rule over the synapses of its target projection
// (Hebb_plasticity / Teacher_plasticity, SynapticPlasticity.cu:38-121), applied
// with every exchange step.  Blocks of the step kernel itself (class
// SC_CLASS_HIST) write, per neuron of the presynaptic population, one byte
// `hist`: bits 0..6 count its spikes at n - delay - s, s < window (window <=
// 127), bit 7 is its spike at n - delay -- rows that are complete before step n
// starts, as every delay is at least one step.  k_plastic then updates the
// synapses (source order, grouped by post row) in
// chunks of at most SC_PLASTIC_CHUNK entries of one row, one wavefront each; a
// chunk carries its row's trigger: the postsynaptic device bit (Hebb) or the
// teacher's device bit (Teacher; -1: none).  Per synapse: the presynaptic
// population-local id and the index of its multiplier in `val` (the f32
// multiplier array of the layout that holds the projection).  skip_quiet: a
// synapse whose trigger did not spike keeps its multiplier (no A0 term, a zero
// or absent A1_J, every multiplier inside the clamp range), so quiet chunks stop
// after reading the trigger.
#define SC_STDP_HEBB     0
#define SC_STDP_TEACHER  1
#define SC_PLASTIC_CHUNK 128
#define SC_PLASTIC_WINDOW_MAX 127
struct ScPlastic {
    const uint32_t *pre;       // population-local presynaptic id
    const uint32_t *vidx;
    const uint32_t *ch_begin;  // n_chunks + 1 synapse offsets
    const int32_t  *ch_trig;   // per chunk
    float *val;
    uint8_t *hist;             // per presynaptic neuron (groups x 64 bytes)
    int32_t n_chunks;
    int32_t pre_group0, pre_groups;  // the presynaptic population's 64-neuron groups
    int32_t rule, delay, t_delay, window;
    int32_t present;           // bit c: coefficient c is set (A0, A1_I, A1_J, A2_JI)
    int32_t grid_exp;          // updated multipliers are rounded to multiples of 2^grid_exp
    int32_t skip_quiet;
    float coef[4];             // A0, A1_I, A1_J, A2_JI
    int32_t affine;            // bit c: coefficient c is (float)(ac[c] + aa[c] * (ab[c] - (double)w))
    double ac[4], aa[4], ab[4];
};

// Fired-list queries without allocation: ids at most, and the word offset of the
// ids in the static scratch (word 0 holds the count; 64 words keep the ids
// 256-byte aligned).
#define SC_FIRED_NOALLOC_MAX (1 << 22)
#define SC_SCRATCH_IDS       64

struct ScInstance {
    ScState  *st;              // per device index
    uint64_t *ring;
    int32_t  *counts;          // R x n_pops
    unsigned long long *events; // R
    unsigned long long *totals; // n_pops spike totals, then the event total (since initialization)
    const ScGroup *groups;
    const ScPop   *pops;
    const ScStim  *stims;
    const int32_t *stim_s;     // per device index of INPUT groups: stimulus (-1: none)
    const int32_t *stim_e;     // entry index
    const uint32_t *outdeg;    // per device index
    ScCtl   *ctl;
    int32_t R, Wr, n_groups, n_pops, n_stims, Ndev, pad0, pad1;
};
