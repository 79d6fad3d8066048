// device_types.hpp -- data layout shared by the host planner and the gfx950 kernels.
//
// Layout in HBM (one network instance):
//   * Neurons live in "device order": populations in internal index order, each
//     padded to a multiple of 64, so one wavefront (64 lanes) covers one
//     64-neuron group of a single population.
//   * Neuron state is SoA: V[Ndev] f32, g[2*Ndev] f32 (exc, inh interleaved),
//     refr[Ndev] i32.
//   * Spike ring: R rows x Wr 64-bit words.  Word g of row r is the wavefront
//     ballot of group g at the step s with s mod R == r.  The last word of every
//     row stays zero; padding entries of sliced-ELL slices point into it.
//   * Projections are gathered in one of two modes:
//       lane mode: sliced ELL, one slice per 64 post neurons, column-major inside
//                  the slice (lane l reads entry k of its row at slice + 64*k + l),
//                  so every load of a wavefront is one contiguous 256-byte run;
//       row mode:  CSR rows cut into chunks of `chunk` synapses, one wavefront per
//                  chunk, lanes striding over the chunk (coalesced).
//     Column entries are device bit indices of the presynaptic neuron (32-bit), or
//     16-bit offsets from a per-slice / per-chunk base.
//   * Row-mode chunks (and lane-mode projections of groups that also receive
//     row-mode input) leave f64 partial sums in slots.  With multipliers that pass
//     the exactness check every partial is the exact sum, so the order in which
//     slots are combined cannot change the result and no floating-point atomics
//     are used anywhere.
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

// projection gather modes
#define SC_MODE_LANE 0
#define SC_MODE_ROW  1

// task kinds
#define SC_TASK_STIM 0
#define SC_TASK_LANE 1
#define SC_TASK_ROW  2

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

struct ScProj {
    const uint32_t *col;       // 32-bit device bit indices
    const uint16_t *col16;     // 16-bit offsets from a per-slice / per-chunk base
    const void     *val;       // f32 / bf16 / fp8 multipliers; nullptr when unit
    const int64_t  *row_ptr;   // row mode: CSR row pointers by post local id
    const int64_t  *sell_off;  // lane mode: slice offset per post group
    const int32_t  *sell_w;    // lane mode: slice width per post group
    const int32_t  *sell_base; // lane mode, 16-bit columns: base bit index per group
    int64_t slot;              // first partial slot of this projection (-1: none)
    float   W;                 // |w_bar|
    float   scale;             // FP8 multiplier scale (a power of two), else 1
    int32_t delay;
    int32_t exc;
    int32_t mode;
    int32_t wprec;
    int32_t idx16;
    int32_t pre_pop, post_pop;
    int32_t chunk;
    int32_t maxch;
    int32_t pad_;
};

struct ScGroup {
    int32_t pop;
    int32_t first_local;   // population-local id of lane 0
    int32_t nvalid;        // valid lanes
    int32_t self_final;    // 1: the lane task finalizes the group directly
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
    int32_t pad_;
};

struct ScCtl {
    int64_t n;      // trial-relative step of launch offset k = 0
    int64_t m;      // stimulus draws since init (3 Philox blocks per draw)
};

// One task is processed by one wavefront.
struct ScTask {
    int32_t kind;
    int32_t inst;
    int32_t group;      // instance-wide group index
    int32_t proj;       // row mode: projection
    int32_t post;       // row mode: post population-local id
    int32_t chunk_idx;  // row mode: chunk index within the row
    int32_t len;        // row mode: synapses in the chunk
    int32_t col_base;   // row mode, 16-bit columns: base bit index
    int64_t e0;         // row mode: first synapse of the chunk
    int64_t pad_;
};

struct ScInstance {
    float    *V;
    float    *g;
    int32_t  *refr;
    uint64_t *ring;
    double   *partial;
    int32_t  *counts;          // R x n_pops
    unsigned long long *events; // R
    const ScGroup *groups;
    const ScPop   *pops;
    const ScProj  *projs;
    const ScStim  *stims;
    const int32_t *stim_s;     // per device index of INPUT groups: stimulus (-1: none)
    const int32_t *stim_e;     // entry index
    const uint32_t *outdeg;    // per device index
    ScCtl   *ctl;
    int32_t R, Wr, n_groups, n_pops, n_projs, n_stims, Ndev, pad_;
};
