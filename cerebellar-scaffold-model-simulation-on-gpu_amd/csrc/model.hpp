// model.hpp -- host-side network description, planner output and engine interface.
#pragma once
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/scaffold.h"
#include "device_types.hpp"

namespace sc {

// Reference time constants (option.h:4-7).
constexpr double kDT = 0.1;
constexpr double kSec = 1000.;
constexpr int kStepMax = 10;

// Thread-local error reporting behind scaffold_last_error().
int fail(int code, const char *fmt, ...);
void clear_error();

struct Population {
    bool set = false;
    std::string name;
    int type = -1, duplicate = 0, num = 0, kind = SCAFFOLD_NORMAL, c_num = 0;
    float Cm = 0, El = 0, dt_ref = 0, Ie = 0, Vr = 0, Vth = 0, gL = 0;
    float w[2] = {0, 0}, E[2] = {0, 0}, g_bar[2] = {0, 0}, tau[2] = {1, 1};
    std::vector<float> pos_x, pos_z;  // synthetic networks: cell positions (um)
};

struct Projection {
    bool set = false;
    int type = -1;
    std::string file;
    int preNum = 0, postNum = 0, pre = -1, post = -1;
    float weight = 0.f;   // initial_weight (sign selects the buffer)
    float delay_ms = 0.f;
    int use_pr = 0;
    // CSR by post local id over the post population (rows beyond it dropped)
    std::vector<int64_t> row_ptr;
    std::vector<int32_t> col;   // pre local ids
    std::vector<float> val;     // empty: unit multipliers
    int max_conv = 0;
    int delay() const { return (int)delay_ms; }
};

// One plasticity rule (a row of the table Init_Plasticity fills,
// SynapticPlasticity.cu:20-33): rule, target and teacher projections (indices),
// constant coefficients (slot c used when bit c of `present` is set) and the
// window of the A2_JI term in exchange steps.
struct Plasticity {
    int rule = 0, target = -1, teacher = -1, window = 50;
    uint32_t present = 0;
    float coef[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    // bit c of `affine`: slot c is (float)(ac[c] + aa[c] * (ab[c] - (double)w)) instead of coef[c]
    uint32_t affine = 0;
    double ac[5] = {0, 0, 0, 0, 0}, aa[5] = {0, 0, 0, 0, 0}, ab[5] = {0, 0, 0, 0, 0};
};

struct Stimulus {
    int pop = -1;
    std::vector<uint32_t> targets;
    std::vector<int64_t> fanout;   // empty: entry i drives targets[i]
    scaffold_stim_func f{};
    float rate = 0.f;              // current Poisson rate (SetInputFreq may change it)
    int n_entries() const { return fanout.empty() ? (int)targets.size() : (int)fanout.size() - 1; }
};

struct Options {
    uint64_t neuron_seed = 0x5eed;
    int weight_precision = SCAFFOLD_WEIGHTS_F32;
    int index_bits = 32;
    int lane_max_fanin = 48;
    int lane_synapses = 0;    // team mode: target synapses per lane of a neuron's team (0: by network size)
    int graph = 0;            // 1: replay captured hipGraphs of whole runs (0: direct launches)
    int word_groups = -1;     // team rows hold projections with at least this many synapses per spike word
                              // as word groups (0: none; -1: by network size)
    int graph_steps = 25;
    int reset_g = 0;
    int stage_spikes = 1;
    int stage_ratio = 0;      // stage a block when window words <= stage_ratio x its gathers' 64-byte segments (0: by size)
    int nontemporal = -1;     // synapse streams as nontemporal loads: -1 when they exceed the L2, 0 never, 1 always
};

// Entry (column k, lane l) of a lane slice that starts at entry `off`: the
// columns of a lane come in groups of 4 adjacent entries (one 16-byte load of
// 32-bit columns; multipliers the same way), groups of the 64 lanes one after
// another; slice widths are padded to a multiple of 4.
inline int64_t sell_pos(int64_t off, int k, int l) { return off + (int64_t)(k >> 2) * (4 * SC_WAVE) + l * 4 + (k & 3); }

// Everything the engine uploads, built once by plan_build().
struct HostProjArrays {
    std::vector<uint32_t> col;
    // 16-bit form of lane slices: the offsets of columns 2p and 2p+1 of a lane
    // share one 32-bit word (low half: column 2p), both relative to the pair
    // header hdr[p] (base device bit); the words of pairs 2g and 2g+1 of a lane
    // are adjacent, at word (g*64 + lane)*2 of the slice; 0xFFFF marks padding.
    // sell_off16: start of each slice in 32-bit words.
    std::vector<uint16_t> col16;
    std::vector<uint32_t> hdr;
    std::vector<int64_t> sell_off16;
    bool idx16 = false;
    std::vector<uint8_t> val;       // raw bytes of f32 / bf16 / fp8 multipliers
    std::vector<int64_t> row_ptr;
    std::vector<int64_t> sell_off;
    std::vector<int32_t> sell_w;
    // per slice (64-neuron group): lowest and highest device bit it reads (lo > hi:
    // none) and the 64-byte segments its gathers touch, estimated per column
    std::vector<uint32_t> sl_lo, sl_hi;
    std::vector<int64_t> sl_seg;
};

// Team-mode population: one merged CSR row per post neuron holding its synapses of
// every incoming projection, projection after projection in index order.
struct TeamPop {
    int pop = -1;
    int team_waves = 1;
    int wprec = SC_W_UNIT;
    std::vector<int64_t> row_ptr;   // num + 1
    std::vector<int32_t> seg_end;   // num x n_in: end of each projection's segment, relative to the row
    std::vector<uint32_t> col;      // device bit indices
    std::vector<uint8_t> val;       // encoded multipliers
    // 16-bit form: rows of 8-entry chunks of one projection each; one header word
    // per chunk (projection slot in the top 3 bits, base device bit); an entry is
    // its 16-bit offset from the base (0xFFFF: padding); row_ptr counts chunks
    bool idx16 = false;
    std::vector<uint16_t> col16;
    std::vector<uint32_t> hdr;
    // per neuron and projection (num x n_in): lowest and highest device bit read
    // (lo > hi: none); per neuron: estimated 64-byte segments its gathers touch
    std::vector<uint32_t> q_lo, q_hi;
    std::vector<int64_t> row_seg;
    std::vector<int64_t> q_seg;     // num x n_in: the same per incoming projection
    // word groups of the dense incoming projections: 4 words per group (projection
    // slot << 29 | presynaptic 64-bit spike word, index of its first multiplier in
    // val, mask bits 0-31, mask bits 32-63); grp_ptr: num + 1 group offsets
    std::vector<uint32_t> grp;
    std::vector<int64_t> grp_ptr;
    int64_t n_grp = 0;
    int64_t n_gval = 0;  // synapses in word groups (their multipliers end val)
};

// A plasticity rule lowered to the device layout: per synapse of the target
// projection (source order) its presynaptic device bit and the index of its
// multiplier in the f32 array that holds it (a lane-mode projection's slices or
// a team-mode population's merged rows); per chunk of one row's synapses the
// row's trigger bit.
struct PlasticPlan {
    Plasticity rule;
    int delay = 0, t_delay = 0, grid_exp = 0;
    bool skip_quiet = false;  // synapses whose trigger did not spike keep their multiplier
    bool team = false;
    int owner = -1;           // team index (team) or projection index (lane)
    int64_t n_val = 0;        // f32 elements of the owner's multiplier array
    std::vector<uint32_t> pre, vidx;
    std::vector<uint32_t> ch_begin;  // chunks of <= SC_PLASTIC_CHUNK synapses of one row
    std::vector<int32_t> ch_trig;    // per chunk: post bit (Hebb) / teacher bit (Teacher), -1: none
};

struct Plan {
    int Ndev = 0, n_groups = 0, Wr = 0, R = 0;
    int total_ref = 0;
    std::vector<ScPop> pops;
    std::vector<ScProj> projs;        // pointers filled by the engine
    std::vector<HostProjArrays> parr;
    std::vector<ScGroup> groups;
    std::vector<ScStim> stims;
    std::vector<int32_t> stim_s, stim_e;   // per device index
    std::vector<uint32_t> outdeg;         // per device index
    std::vector<int32_t> dev_to_ref;      // -1 for padding
    std::vector<int32_t> ref_to_dev;
    std::vector<TeamPop> teams;
    std::vector<int> team_of;             // per population: index into teams or -1
    std::vector<ScTask> tasks;            // step kernel work list (planner form)
    int64_t nnz = 0;
    int64_t algo_bytes = 0;
    int stage = 1;  // stage spike windows in LDS where the plan finds it pays
    int stage_ratio = 8;
    int lane_synapses = 8;  // team width target in effect
    int group_min = 0;      // word-group threshold in effect (0: none)
    bool nontemporal = false;  // load the synapse streams nontemporal
    int64_t syn16 = 0;  // synapses stored as 16-bit column offsets
    int64_t n_grp = 0, grp_syn = 0;  // word groups of the team rows and the synapses they hold
    // bytes the stored layout streams per step: every column / header / multiplier
    // of the lane slices and merged rows (padding included), 16 per word-group
    // record (not the multipliers of spiking connected neurons, which depend on
    // activity), and the neuron state and spike bits as in algo_bytes
    int64_t stream_bytes = 0;
    int max_delay = 0;
    std::vector<int> out_pops;            // OUTPUT populations
    int n_output = 0;
    std::vector<int32_t> out_bits;        // device bit of each OUTPUT neuron, in output-train order
    int T_out = 0;                        // steps per output-train hand-off (T_cpu_sim)
    std::vector<PlasticPlan> plastic;     // plasticity rules in table order
};

struct Engine;  // device side (engine.hip)
struct Recorder;  // spike files (recorder.cpp)

// Engine entry points used by the C ABI layer.
int engine_create(Engine **e, Plan &plan);
void engine_destroy(Engine *e);
int engine_init_params(Engine *e, uint64_t seed, int64_t call_index);
int engine_zero_g(Engine *e);
int engine_set_ctl(Engine *e, int64_t n, int64_t m);
int engine_upload_stims(Engine *e, const std::vector<ScStim> &stims);
int engine_run(Engine *e, int n_steps, int use_graph, int graph_steps);
// engine_run, then (train != nullptr) the output spike train of the T_out steps
// from first_out_step on, gathered on the device and copied with one D2H copy into
// the engine's pinned buffer (*train), valid until the next call.
// With a recorder, the steps run in pieces that end at multiples of the
// recorder's window (step numbers counted from step0 = the first step of the
// call), each completed window handed to the recorder.
int engine_run_output(Engine *e, int n_steps, int use_graph, int graph_steps, int64_t first_out_step,
                      const uint8_t **train, Recorder *rec = nullptr, int rec_T = 0, int64_t step0 = 0);
// Orders the engine's stream after every copy handed to the recorder.
int engine_record_sync(Engine *e, Recorder *rec);
// Step-kernel timer: on 1 starts (clears) it, 0 stops it, -1 leaves it; returns
// the measured milliseconds and the step launches they cover.
int engine_timer(Engine *e, int on, double *ms, int64_t *launches);
int engine_sync(Engine *e);
int engine_get_state(Engine *e, float *V, float *g, int32_t *refr);
int engine_set_state(Engine *e, const float *V, const float *g, const int32_t *refr);
int engine_get_ring_row(Engine *e, int row, uint64_t *words);
int engine_get_counts(Engine *e, int row, int32_t *counts, unsigned long long *events);
int engine_get_fired(Engine *e, int row, int32_t *ids, int32_t cap, int32_t *n);
// engine_get_fired through the static scratch of the engine's device (no allocation).
int engine_get_fired_noalloc(Engine *e, int row, int32_t *ids, int32_t cap, int32_t *n);
int engine_get_totals(Engine *e, unsigned long long *totals);
int engine_time(Engine *e, int n_steps, int use_graph, int graph_steps,
                double *ms_total, double *ms_kernel_mean);
int64_t engine_device_bytes(Engine *e);
void engine_blocks(Engine *e, int *n_blocks, int *n_staged);
int engine_copy_probe(int64_t bytes, int iters, double *ms);
// Current multipliers of a plastic projection in source order; returns -1 if no
// plasticity rule targets the projection.
int engine_read_plastic(Engine *e, int proj, float *out);

// Batching (engine.hip).
struct Batch;
int batch_build(Batch **b, std::vector<Engine *> &engines, const std::vector<const Plan *> &plans);
int batch_run(Batch *b, int n_steps, int graph_steps, double *ms);
void batch_destroy(Batch *b);

}  // namespace sc

// The opaque environment behind the C ABI.
struct scaffold_env {
    int n_types = 0, n_conn = 0, n_inputs = 0, n_plast = 0;
    std::vector<sc::Population> pops;
    int normal_next = 0, other_next = 0;
    std::vector<sc::Projection> projs;
    int gpu_next = 0, cpu_next = 0;
    std::vector<sc::Stimulus> stims;
    std::vector<sc::Plasticity> plast;
    sc::Options opt;
    bool initialized = false;
    bool finalized = false;
    int T_print = 0, T_cpu_sim = 0;
    int64_t init_calls = 0;
    int64_t draws = 0;      // stimulus draws since init (mirror of the device counter)
    int64_t cur_step = 0;   // trial-relative step of the next launch
    sc::Plan plan;
    sc::Engine *eng = nullptr;
    std::vector<uint64_t> tmp_words;
    int in_batches = 0;           // batches referencing this environment's device state
    sc::Recorder *rec = nullptr;  // spike files written as the run goes (scaffold_record_spikes)
};

namespace sc {
// describe.cpp
int plan_build(scaffold_env *env, Plan &plan);
int load_dat(const char *file, int preNum, int postNum, int postPopNum, Projection &p);
int synth_describe(scaffold_env *env, const scaffold_synth_config *cfg);
int describe_reference(scaffold_env *env, const char *dir, uint64_t stim_seed);
// capi.cpp / planner.cpp
std::vector<ScStim> dev_stims(scaffold_env *env);
float plan_multiplier(const scaffold_env *env, int proj, int64_t e);
void encode_values(int wprec, float *v, int64_t n, float *scale, float *sum_scale);
}  // namespace sc
