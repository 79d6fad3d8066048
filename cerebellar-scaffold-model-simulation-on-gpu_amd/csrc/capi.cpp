// capi.cpp -- the extern "C" boundary (include/scaffold.h): lifecycle, description
// registration, options and queries.  Device work is delegated to engine.hip.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <new>

#include "model.hpp"
#include "recorder.hpp"

namespace sc {

static thread_local char g_err[1024] = "";

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

void clear_error() { g_err[0] = 0; }

static int check_env(scaffold_env *env)
{
    if (!env) return fail(SCAFFOLD_ERR_INVALID_ARG, "null environment");
    if (env->finalized) return fail(SCAFFOLD_ERR_STATE, "environment already finalized");
    return SCAFFOLD_OK;
}

static int check_described(scaffold_env *env)
{
    int rc = check_env(env);
    if (rc) return rc;
    if (env->initialized)
        return fail(SCAFFOLD_ERR_STATE, "network description is frozen after initialization");
    return SCAFFOLD_OK;
}

static int check_ready(scaffold_env *env)
{
    int rc = check_env(env);
    if (rc) return rc;
    if (!env->initialized) return fail(SCAFFOLD_ERR_STATE, "environment not initialized");
    return SCAFFOLD_OK;
}

// Stimulus thresholds as the reference evaluates them: PoissonProcess compares
// r < freq/sec*DT with freq a float promoted to double (UserInputFunctions.cu:4-6).
static double poisson_thr(float hz) { return (double)hz / kSec * kDT; }

static ScStim make_dev_stim(const Stimulus &s)
{
    ScStim d{};
    d.key0 = (uint32_t)s.f.seed;
    d.key1 = (uint32_t)(s.f.seed >> 32);
    d.p_thr = poisson_thr(s.rate);
    d.poisson_on = s.f.poisson_on;
    d.poisson_off = s.f.poisson_off;
    d.b_thr = poisson_thr(s.f.burst_hz);
    d.burst_on = s.f.burst_on;
    d.burst_off = s.f.burst_off;
    d.periodic_on = s.f.periodic_on;
    d.periodic_off = s.f.periodic_off;
    d.periodic_hz = s.f.periodic_hz;
    d.periodic_start = s.f.periodic_start;
    d.periodic_incl = s.f.periodic_off_inclusive;
    d.sine_max = s.f.sine_max_hz;
    d.sine_on = s.f.sine_on;
    d.sine_off = s.f.sine_off;
    d.sine_cycle = s.f.sine_cycle_s;
    d.sine_shift = s.f.sine_shift;
    return d;
}

std::vector<ScStim> dev_stims(scaffold_env *env)
{
    std::vector<ScStim> v;
    for (auto &s : env->stims) v.push_back(make_dev_stim(s));
    return v;
}

}  // namespace sc

using namespace sc;

extern "C" {

const char *scaffold_last_error(void) { return g_err; }

const char *scaffold_version(void) { return "scaffold-mi355x 0.1 gfx950"; }

int scaffold_stim_func_reference(int func_id, uint64_t seed, scaffold_stim_func *out)
{
    if (!out) return fail(SCAFFOLD_ERR_INVALID_ARG, "null output");
    const double inf = std::numeric_limits<double>::infinity();
    scaffold_stim_func f{};
    f.seed = seed;
    f.poisson_off = inf;
    f.burst_on = 0;
    f.burst_off = 0;
    f.periodic_on = 0;
    f.periodic_off = 0;
    switch (func_id) {
    case 0:  // background_noise: (0.0 <= time) && PoissonProcess(r, time, 1.0, 0.0)
        f.poisson_hz = 1.0f;
        f.poisson_on = 0.0;
        break;
    case 1:  // tone_stim: ((300 <= t) && (t < 351) && PeriodicFiring(r,t,140,292.8)) || Poisson(1 Hz)
        f.poisson_hz = 1.0f;
        f.poisson_on = -inf;
        f.periodic_hz = 140.0f;
        f.periodic_start = 292.8f;
        f.periodic_on = 300.0;
        f.periodic_off = 351.0;
        break;
    case 2:  // puff_stim: (0.0 <= time) && PoissonProcess(r, t, SinusoidalOscillation(3, 1.5, 2, 0, t), 0)
        f.poisson_on = 0.0;
        f.poisson_off = 0.0;
        f.sine_max_hz = 3.0f;
        f.sine_cycle_s = 2.0f;
        f.sine_shift = 0.0f;
        f.sine_on = 0.0;
        f.sine_off = inf;
        break;
    default:
        return fail(SCAFFOLD_ERR_INVALID_ARG, "reference stimulus function %d does not exist (0, 1, 2)", func_id);
    }
    *out = f;
    return SCAFFOLD_OK;
}

int scaffold_create_network_env(scaffold_env **env, int num_of_dev, int num_of_neuron_types,
                                int num_of_connectivity_types, int num_of_inputs,
                                int num_of_plasticities)
{
    clear_error();
    if (!env) return fail(SCAFFOLD_ERR_INVALID_ARG, "null env pointer");
    *env = nullptr;
    if (num_of_dev != 1)
        return fail(SCAFFOLD_ERR_UNSUPPORTED, "num_of_dev=%d: this engine drives exactly one GPU", num_of_dev);
    if (num_of_neuron_types <= 0 || num_of_neuron_types > SC_MAX_POPS)
        return fail(SCAFFOLD_ERR_INVALID_ARG, "num_of_neuron_types must be in [1, %d]", SC_MAX_POPS);
    if (num_of_connectivity_types < 0 || num_of_inputs < 0 || num_of_plasticities < 0)
        return fail(SCAFFOLD_ERR_INVALID_ARG, "negative table size");
    scaffold_env *e = new (std::nothrow) scaffold_env();
    if (!e) return fail(SCAFFOLD_ERR_CAPACITY, "out of host memory");
    e->n_types = num_of_neuron_types;
    e->n_conn = num_of_connectivity_types;
    e->n_inputs = num_of_inputs;
    e->n_plast = num_of_plasticities;
    e->pops.resize(num_of_neuron_types);
    e->projs.resize(num_of_connectivity_types);
    *env = e;
    return SCAFFOLD_OK;
}

int scaffold_set_neuron_params(scaffold_env *env, int type, const char *name, int duplicate, int num,
                               float Cm, float El, float dt_ref, float Ie, float Vr, float Vth,
                               float gL, int device_type, int c_num, const double *channels,
                               int *index_out)
{
    clear_error();
    int rc = check_described(env);
    if (rc) return rc;
    if (num < 0) return fail(SCAFFOLD_ERR_INVALID_ARG, "negative population size");
    if (device_type < SCAFFOLD_NORMAL || device_type > SCAFFOLD_OUTPUT)
        return fail(SCAFFOLD_ERR_INVALID_ARG, "unknown device type %d", device_type);
    if (c_num < 0 || c_num > 2) return fail(SCAFFOLD_ERR_UNSUPPORTED, "c_num=%d (at most 2 channels)", c_num);
    if (c_num > 0 && !channels) return fail(SCAFFOLD_ERR_INVALID_ARG, "null channel table");
    if (device_type != SCAFFOLD_INPUT) {
        if (!(Cm > 0.f)) return fail(SCAFFOLD_ERR_INVALID_ARG, "Cm must be positive");
        for (int c = 0; c < c_num; ++c)
            if (!(channels[4 * c + 3] > 0.0)) return fail(SCAFFOLD_ERR_INVALID_ARG, "tau must be positive");
    }
    if (env->normal_next + env->other_next >= env->n_types)
        return fail(SCAFFOLD_ERR_CAPACITY, "more populations than num_of_neuron_types=%d", env->n_types);
    // init_params.cu:259-265: NORMAL from the front, INPUT/OUTPUT from the back.
    int target = (device_type == SCAFFOLD_NORMAL) ? env->normal_next++ : env->n_types - 1 - env->other_next++;
    Population &p = env->pops[target];
    p = Population();
    p.set = true;
    p.name = name ? name : "";
    p.type = type;
    p.duplicate = duplicate;
    p.num = num;
    p.kind = device_type;
    p.c_num = c_num;
    p.Cm = Cm; p.El = El; p.dt_ref = dt_ref; p.Ie = Ie; p.Vr = Vr; p.Vth = Vth; p.gL = gL;
    for (int c = 0; c < c_num; ++c) {  // (CTYPE) va_arg(args, double), init_params.cu:293-296
        p.w[c] = (float)channels[4 * c + 0];
        p.E[c] = (float)channels[4 * c + 1];
        p.g_bar[c] = (float)channels[4 * c + 2];
        p.tau[c] = (float)channels[4 * c + 3];
    }
    if (index_out) *index_out = target;
    return SCAFFOLD_OK;
}

static int assign_projection(scaffold_env *env, int postType, int *target)
{
    if (env->gpu_next + env->cpu_next >= env->n_conn)
        return fail(SCAFFOLD_ERR_CAPACITY, "more projections than num_of_connectivity_types=%d", env->n_conn);
    // init_params.cu:309-315: projections onto OUTPUT populations from the back.
    if (env->pops[postType].kind == SCAFFOLD_OUTPUT) *target = env->n_conn - 1 - env->cpu_next++;
    else *target = env->gpu_next++;
    return SCAFFOLD_OK;
}

static int check_projection_args(scaffold_env *env, int preNum, int postNum, int preType, int postType,
                                 float delay)
{
    if (preType < 0 || preType >= env->n_types || !env->pops[preType].set)
        return fail(SCAFFOLD_ERR_INVALID_ARG, "preType %d is not a registered population", preType);
    if (postType < 0 || postType >= env->n_types || !env->pops[postType].set)
        return fail(SCAFFOLD_ERR_INVALID_ARG, "postType %d is not a registered population", postType);
    if (env->pops[postType].kind == SCAFFOLD_INPUT)
        return fail(SCAFFOLD_ERR_INVALID_ARG, "INPUT population %d cannot receive projections", postType);
    if (preNum < 0 || postNum < 0) return fail(SCAFFOLD_ERR_INVALID_ARG, "negative preNum/postNum");
    if ((int)delay < 1)
        return fail(SCAFFOLD_ERR_INVALID_ARG, "delay %.3f ms: at least one exchange step is required", delay);
    return SCAFFOLD_OK;
}

int scaffold_set_connectivity_params(scaffold_env *env, int type, const char *filename, int preNum,
                                     int postNum, int preType, int postType, float initial_weight,
                                     float delay, int use_parallel_reduction, int *index_out)
{
    clear_error();
    int rc = check_described(env);
    if (rc) return rc;
    if (!filename) return fail(SCAFFOLD_ERR_INVALID_ARG, "null file name");
    rc = check_projection_args(env, preNum, postNum, preType, postType, delay);
    if (rc) return rc;
    Projection p;
    p.type = type;
    p.file = filename;
    p.preNum = preNum; p.postNum = postNum; p.pre = preType; p.post = postType;
    p.weight = initial_weight; p.delay_ms = delay; p.use_pr = use_parallel_reduction;
    if (preNum > env->pops[preType].num)
        return fail(SCAFFOLD_ERR_INVALID_ARG, "preNum %d exceeds population %d size %d", preNum, preType,
                    env->pops[preType].num);
    rc = load_dat(filename, preNum, postNum, env->pops[postType].num, p);
    if (rc) return rc;
    int target = -1;
    rc = assign_projection(env, postType, &target);
    if (rc) return rc;
    p.set = true;
    env->projs[target] = std::move(p);
    if (index_out) *index_out = target;
    return SCAFFOLD_OK;
}

int scaffold_set_connectivity_csr(scaffold_env *env, int type, int preNum, int postNum, int preType,
                                  int postType, float initial_weight, float delay,
                                  const int64_t *row_ptr, const int32_t *col, const float *val,
                                  int *index_out)
{
    clear_error();
    int rc = check_described(env);
    if (rc) return rc;
    rc = check_projection_args(env, preNum, postNum, preType, postType, delay);
    if (rc) return rc;
    if (!row_ptr) return fail(SCAFFOLD_ERR_INVALID_ARG, "null row_ptr");
    const int pre_n = env->pops[preType].num, post_n = env->pops[postType].num;
    if (preNum > pre_n) return fail(SCAFFOLD_ERR_INVALID_ARG, "preNum exceeds the population size");
    if (row_ptr[0] != 0) return fail(SCAFFOLD_ERR_FORMAT, "row_ptr[0] must be 0");
    for (int j = 0; j < postNum; ++j)
        if (row_ptr[j + 1] < row_ptr[j]) return fail(SCAFFOLD_ERR_FORMAT, "row_ptr decreases at row %d", j);
    const int64_t nnz = row_ptr[postNum];
    if (nnz > 0 && !col) return fail(SCAFFOLD_ERR_INVALID_ARG, "null col");
    Projection p;
    p.type = type;
    p.preNum = preNum; p.postNum = postNum; p.pre = preType; p.post = postType;
    p.weight = initial_weight; p.delay_ms = delay;
    // rows beyond the post population are dropped (the reference allocates them
    // and never reads them, init_params.cu:743)
    const int rows = postNum < post_n ? postNum : post_n;
    p.row_ptr.assign(post_n + 1, 0);
    for (int j = 0; j <= post_n; ++j) p.row_ptr[j] = row_ptr[j <= rows ? j : rows];
    const int64_t kept = p.row_ptr[post_n];
    p.col.assign(col, col + kept);
    for (int64_t e = 0; e < kept; ++e)
        if (p.col[e] < 0 || p.col[e] >= preNum)
            return fail(SCAFFOLD_ERR_FORMAT, "pre id %d out of range [0,%d)", p.col[e], preNum);
    if (val) p.val.assign(val, val + kept);
    for (int j = 0; j < post_n; ++j) {
        int len = (int)(p.row_ptr[j + 1] - p.row_ptr[j]);
        if (len > p.max_conv) p.max_conv = len;
    }
    int target = -1;
    rc = assign_projection(env, postType, &target);
    if (rc) return rc;
    p.set = true;
    env->projs[target] = std::move(p);
    if (index_out) *index_out = target;
    return SCAFFOLD_OK;
}

// The checks and the table row shared by both registration calls; the
// coefficients are filled in by the caller.
static int plasticity_row(scaffold_env *env, int rule, int target, int teacher, uint32_t present, int time_window,
                          Plasticity &r)
{
    int rc = check_described(env);
    if (rc) return rc;
    if ((int)env->plast.size() >= env->n_plast)
        return fail(SCAFFOLD_ERR_CAPACITY, "more plasticity rules than num_of_plasticities=%d", env->n_plast);
    if (rule != SCAFFOLD_STDP_HEBB && rule != SCAFFOLD_STDP_TEACHER)
        return fail(SCAFFOLD_ERR_INVALID_ARG, "rule %d: Hebb (0) and Teacher (1) are the reference's rules", rule);
    if (target < 0 || target >= env->n_conn || !env->projs[target].set)
        return fail(SCAFFOLD_ERR_INVALID_ARG, "target projection %d not registered", target);
    if (rule == SCAFFOLD_STDP_TEACHER) {
        if (teacher < 0 || teacher >= env->n_conn || !env->projs[teacher].set)
            return fail(SCAFFOLD_ERR_INVALID_ARG, "teacher projection %d not registered", teacher);
        if (env->projs[teacher].post != env->projs[target].post)
            return fail(SCAFFOLD_ERR_INVALID_ARG, "the teacher projection must share the target's post population");
    }
    if (present >> 5) return fail(SCAFFOLD_ERR_INVALID_ARG, "present: bits 0-4 only");
    if (time_window < 0) return fail(SCAFFOLD_ERR_INVALID_ARG, "time_window must be >= 0");
    r = Plasticity();
    r.rule = rule;
    r.target = target;
    r.teacher = rule == SCAFFOLD_STDP_TEACHER ? teacher : -1;
    // 0: the window the reference's kernels use whatever the table says,
    // s_upper_bound = 50 (SynapticPlasticity.cu:45,86)
    r.window = time_window == 0 ? 50 : time_window;
    r.present = present;
    return SCAFFOLD_OK;
}

int scaffold_set_plasticity(scaffold_env *env, int rule, int target, int teacher, const float coefficients[5],
                            uint32_t present, int time_window, int *index_out)
{
    clear_error();
    Plasticity r;
    int rc = plasticity_row(env, rule, target, teacher, present, time_window, r);
    if (rc) return rc;
    if (present && !coefficients) return fail(SCAFFOLD_ERR_INVALID_ARG, "null coefficients");
    for (int c = 0; c < 5; ++c) {
        r.coef[c] = (present >> c) & 1 ? coefficients[c] : 0.f;
        if (!std::isfinite(r.coef[c])) return fail(SCAFFOLD_ERR_INVALID_ARG, "coefficient %d is not finite", c);
    }
    env->plast.push_back(r);
    if (index_out) *index_out = (int)env->plast.size() - 1;
    return SCAFFOLD_OK;
}

int scaffold_set_plasticity_affine(scaffold_env *env, int rule, int target, int teacher,
                                   const scaffold_stdp_affine coefficients[5], uint32_t present, int time_window,
                                   int *index_out)
{
    clear_error();
    Plasticity r;
    int rc = plasticity_row(env, rule, target, teacher, present, time_window, r);
    if (rc) return rc;
    if (present && !coefficients) return fail(SCAFFOLD_ERR_INVALID_ARG, "null coefficients");
    for (int c = 0; c < 5; ++c) {
        if (!((present >> c) & 1)) continue;
        const scaffold_stdp_affine &f = coefficients[c];
        if (!std::isfinite(f.c) || !std::isfinite(f.a) || !std::isfinite(f.b))
            return fail(SCAFFOLD_ERR_INVALID_ARG, "coefficient %d is not finite", c);
        if (f.a == 0.0) {  // a constant: (float)(c + 0 * (b - w)) is (float)c
            r.coef[c] = (float)f.c;
            if (!std::isfinite(r.coef[c])) return fail(SCAFFOLD_ERR_INVALID_ARG, "coefficient %d overflows float", c);
            continue;
        }
        r.affine |= 1u << c;
        r.ac[c] = f.c;
        r.aa[c] = f.a;
        r.ab[c] = f.b;
    }
    env->plast.push_back(r);
    if (index_out) *index_out = (int)env->plast.size() - 1;
    return SCAFFOLD_OK;
}

static int add_stim(scaffold_env *env, int neuron_index, std::vector<uint32_t> &&ids,
                    std::vector<int64_t> &&fan, const scaffold_stim_func *func, int *index_out)
{
    if (neuron_index < 0 || neuron_index >= env->n_types || !env->pops[neuron_index].set)
        return fail(SCAFFOLD_ERR_INVALID_ARG, "stimulus population %d is not registered", neuron_index);
    if (env->pops[neuron_index].kind != SCAFFOLD_INPUT)
        return fail(SCAFFOLD_ERR_INVALID_ARG, "stimuli drive INPUT populations only");
    if ((int)env->stims.size() >= env->n_inputs)
        return fail(SCAFFOLD_ERR_CAPACITY, "more stimuli than num_of_inputs=%d", env->n_inputs);
    if (func->sine_max_hz != 0.f &&
        !(std::isfinite(func->sine_max_hz) && std::isfinite(func->sine_cycle_s) && func->sine_cycle_s != 0.f &&
          std::isfinite(func->sine_shift)))
        return fail(SCAFFOLD_ERR_INVALID_ARG, "sinusoidal stimulus needs a finite rate, phase and non-zero period");
    const int num = env->pops[neuron_index].num;
    for (uint32_t id : ids)
        if ((int64_t)id >= num) return fail(SCAFFOLD_ERR_FORMAT, "stimulus id %u out of range [0,%d)", id, num);
    Stimulus s;
    s.pop = neuron_index;
    s.targets = std::move(ids);
    s.fanout = std::move(fan);
    s.f = *func;
    s.rate = func->poisson_hz;
    env->stims.push_back(std::move(s));
    if (index_out) *index_out = (int)env->stims.size() - 1;
    return SCAFFOLD_OK;
}

int scaffold_set_input_stim(scaffold_env *env, int neuron_index, const char *id_file,
                            const scaffold_stim_func *func, int *index_out)
{
    clear_error();
    int rc = check_described(env);
    if (rc) return rc;
    if (!id_file || !func) return fail(SCAFFOLD_ERR_INVALID_ARG, "null argument");
    FILE *fp = fopen(id_file, "r");
    if (!fp) return fail(SCAFFOLD_ERR_IO, "cannot open file: %s", id_file);
    std::vector<uint32_t> ids;
    char line[256];
    int lineno = 0;
    while (fgets(line, sizeof(line), fp)) {  // UserInputFunctions.cu:53-57
        ++lineno;
        char *p = line;
        while (*p == ' ' || *p == '\t') ++p;
        if (*p == '\n' || *p == '\r' || *p == 0) continue;
        unsigned long v;
        char *end;
        v = strtoul(p, &end, 10);
        if (end == p || v > 0xFFFFFFFFul) {
            fclose(fp);
            return fail(SCAFFOLD_ERR_FORMAT, "%s:%d: not an unsigned id", id_file, lineno);
        }
        ids.push_back((uint32_t)v);
    }
    fclose(fp);
    return add_stim(env, neuron_index, std::move(ids), {}, func, index_out);
}

int scaffold_set_input_stim_ids(scaffold_env *env, int neuron_index, const uint32_t *ids, int n,
                                const int64_t *fanout_ptr, const scaffold_stim_func *func, int *index_out)
{
    clear_error();
    int rc = check_described(env);
    if (rc) return rc;
    if (!func || n < 0 || (n > 0 && !ids && !fanout_ptr)) return fail(SCAFFOLD_ERR_INVALID_ARG, "bad arguments");
    std::vector<int64_t> fan;
    std::vector<uint32_t> v;
    if (fanout_ptr) {
        if (fanout_ptr[0] != 0) return fail(SCAFFOLD_ERR_FORMAT, "fanout_ptr[0] must be 0");
        for (int i = 0; i < n; ++i)
            if (fanout_ptr[i + 1] < fanout_ptr[i]) return fail(SCAFFOLD_ERR_FORMAT, "fanout_ptr decreases");
        fan.assign(fanout_ptr, fanout_ptr + n + 1);
        if (fan[n] > 0 && !ids) return fail(SCAFFOLD_ERR_INVALID_ARG, "null ids");
        v.assign(ids, ids + fan[n]);
    } else {
        v.assign(ids, ids + n);
    }
    return add_stim(env, neuron_index, std::move(v), std::move(fan), func, index_out);
}

int scaffold_set_option(scaffold_env *env, const char *key, int64_t value)
{
    clear_error();
    int rc = check_env(env);
    if (rc) return rc;
    if (!key) return fail(SCAFFOLD_ERR_INVALID_ARG, "null key");
    Options &o = env->opt;
    const bool frozen = env->initialized;
    auto frozen_err = [&]() { return fail(SCAFFOLD_ERR_STATE, "option %s must be set before initialization", key); };
    if (!strcmp(key, "neuron_seed")) { if (frozen) return frozen_err(); o.neuron_seed = (uint64_t)value; }
    else if (!strcmp(key, "weight_precision")) {
        if (frozen) return frozen_err();
        if (value < 0 || value > 2) return fail(SCAFFOLD_ERR_INVALID_ARG, "weight_precision in {0,1,2}");
        o.weight_precision = (int)value;
    } else if (!strcmp(key, "index_bits")) {
        if (frozen) return frozen_err();
        if (value != 16 && value != 32) return fail(SCAFFOLD_ERR_INVALID_ARG, "index_bits is 16 or 32");
        o.index_bits = (int)value;
    } else if (!strcmp(key, "lane_max_fanin")) {
        if (frozen) return frozen_err();
        if (value < 0 || value > 4096) return fail(SCAFFOLD_ERR_INVALID_ARG, "lane_max_fanin in [0,4096]");
        o.lane_max_fanin = (int)value;
    } else if (!strcmp(key, "lane_synapses")) {
        if (frozen) return frozen_err();
        if (value < 0 || value > 4096) return fail(SCAFFOLD_ERR_INVALID_ARG, "lane_synapses in [0, 4096]");
        o.lane_synapses = (int)value;
    } else if (!strcmp(key, "graph")) o.graph = value ? 1 : 0;
    else if (!strcmp(key, "graph_steps")) {
        if (value < 1 || value > 100000) return fail(SCAFFOLD_ERR_INVALID_ARG, "graph_steps in [1,100000]");
        o.graph_steps = (int)value;
    } else if (!strcmp(key, "reset_g")) o.reset_g = value ? 1 : 0;
    else if (!strcmp(key, "word_groups")) {
        if (frozen) return frozen_err();
        if (value < -1 || value > 64) return fail(SCAFFOLD_ERR_INVALID_ARG, "word_groups in [-1, 64]");
        o.word_groups = (int)value;
    } else if (!strcmp(key, "stage_spikes")) {
        if (frozen) return frozen_err();
        o.stage_spikes = value ? 1 : 0;
    } else if (!strcmp(key, "stage_ratio")) {
        if (frozen) return frozen_err();
        if (value < 0 || value > 1000000) return fail(SCAFFOLD_ERR_INVALID_ARG, "stage_ratio in [0, 1000000]");
        o.stage_ratio = (int)value;
    } else if (!strcmp(key, "nontemporal")) {
        if (frozen) return frozen_err();
        if (value < -1 || value > 1) return fail(SCAFFOLD_ERR_INVALID_ARG, "nontemporal is -1, 0 or 1");
        o.nontemporal = (int)value;
    } else return fail(SCAFFOLD_ERR_INVALID_ARG, "unknown option %s", key);
    return SCAFFOLD_OK;
}

int scaffold_get_option(scaffold_env *env, const char *key, int64_t *value)
{
    clear_error();
    int rc = check_env(env);
    if (rc) return rc;
    if (!key || !value) return fail(SCAFFOLD_ERR_INVALID_ARG, "null argument");
    const Options &o = env->opt;
    if (!strcmp(key, "neuron_seed")) *value = (int64_t)o.neuron_seed;
    else if (!strcmp(key, "weight_precision")) *value = o.weight_precision;
    else if (!strcmp(key, "index_bits")) *value = o.index_bits;
    else if (!strcmp(key, "lane_max_fanin")) *value = o.lane_max_fanin;
    else if (!strcmp(key, "lane_synapses")) *value = o.lane_synapses;
    else if (!strcmp(key, "graph")) *value = o.graph;
    else if (!strcmp(key, "graph_steps")) *value = o.graph_steps;
    else if (!strcmp(key, "reset_g")) *value = o.reset_g;
    else if (!strcmp(key, "stage_spikes")) *value = o.stage_spikes;
    else if (!strcmp(key, "word_groups")) *value = o.word_groups;
    else if (!strcmp(key, "stage_ratio")) *value = o.stage_ratio;
    else if (!strcmp(key, "nontemporal")) *value = o.nontemporal;
    else return fail(SCAFFOLD_ERR_INVALID_ARG, "unknown option %s", key);
    return SCAFFOLD_OK;
}

int scaffold_initialize_network_env(scaffold_env *env, int T_print, int T_cpu_sim)
{
    clear_error();
    int rc = check_described(env);
    if (rc) return rc;
    if (T_cpu_sim <= 0) return fail(SCAFFOLD_ERR_INVALID_ARG, "T_cpu_sim must be positive");
    // GPUScaffoldSimulation.cu:674-684
    float max_delay = 0.f;
    for (auto &p : env->projs)
        if (p.set && p.delay_ms > max_delay) max_delay = p.delay_ms;
    env->T_print = ((float)T_print > max_delay) ? T_print : (int)max_delay;
    env->T_cpu_sim = T_cpu_sim;
    if (env->T_print < env->T_cpu_sim || (env->T_print % env->T_cpu_sim) != 0)
        return fail(SCAFFOLD_ERR_INVALID_ARG,
                    "T_print (%d) must be a multiple of T_cpu_sim (%d)", env->T_print, env->T_cpu_sim);
    env->plan = Plan();
    env->plan.R = 2 * env->T_print;  // N_PRINT_THREDS * T_print rows (GPUScaffoldSimulation.cu:361)
    rc = plan_build(env, env->plan);
    if (rc) return rc;
    rc = engine_create(&env->eng, env->plan);
    if (rc) return rc;
    // Initialization draws once (:381); InitializeNetworkEnv then resets trial 0 (:689 -> :851).
    rc = engine_init_params(env->eng, env->opt.neuron_seed, 0);
    if (!rc) rc = engine_init_params(env->eng, env->opt.neuron_seed, 1);
    if (!rc) rc = engine_set_ctl(env->eng, 0, 0);
    if (rc) {
        engine_destroy(env->eng);
        env->eng = nullptr;
        return rc;
    }
    env->init_calls = 2;
    env->draws = 0;
    env->cur_step = 0;
    env->initialized = true;
    return SCAFFOLD_OK;
}

int scaffold_loop_n_steps(scaffold_env *env, int start_step, int n_steps, int trial,
                          const char **output_spike_train)
{
    (void)trial;
    clear_error();
    int rc = check_ready(env);
    if (rc) return rc;
    if (start_step < 0 || n_steps < 0) return fail(SCAFFOLD_ERR_INVALID_ARG, "negative step");
    if (start_step != env->cur_step) {
        if (env->rec && ((rc = engine_record_sync(env->eng, env->rec)) || (rc = recorder_restart(env->rec))))
            return rc;
        rc = engine_set_ctl(env->eng, start_step, env->draws);
        if (rc) return rc;
        env->cur_step = start_step;
    }
    // Hand-off (GPUScaffoldSimulation.cu:768-783): after every step that completes a
    // window of T_cpu_sim steps (step count n with n % T_cpu_sim == 0; the ring
    // holds a multiple of T_cpu_sim rows there, so this is its target_row test).
    // The train of the last window completed in this call is gathered on the
    // device and copied once; a call that completes no window leaves
    // *output_spike_train as it was, as the reference does.
    const int T = env->T_cpu_sim;
    const int64_t end = env->cur_step + n_steps;
    const int64_t last = end - end % T;  // last window boundary at or before end
    const bool handoff = output_spike_train && n_steps > 0 && last > env->cur_step;
    const uint8_t *train = nullptr;
    rc = engine_run_output(env->eng, n_steps, env->opt.graph, env->opt.graph_steps, last - T,
                           handoff ? &train : nullptr, env->rec, T, env->cur_step);
    if (rc) return rc;
    env->cur_step = end;
    env->draws += n_steps;
    if (handoff) *output_spike_train = reinterpret_cast<const char *>(train);
    return SCAFFOLD_OK;
}

int scaffold_set_input_freq(scaffold_env *env, const float *freq_list)
{
    clear_error();
    int rc = check_env(env);
    if (rc) return rc;
    if (!freq_list) return fail(SCAFFOLD_ERR_INVALID_ARG, "null freq_list");
    bool changed = false;
    for (size_t i = 0; i < env->stims.size(); ++i) {
        Stimulus &s = env->stims[i];
        if (s.f.poisson_hz_from_freq && s.rate != freq_list[i]) {
            s.rate = freq_list[i];
            changed = true;
        }
    }
    if (changed && env->initialized) {
        std::vector<ScStim> v = dev_stims(env);
        env->plan.stims = v;
        return engine_upload_stims(env->eng, v);
    }
    return SCAFFOLD_OK;
}

int scaffold_reset_network_env(scaffold_env *env, int trial)
{
    clear_error();
    int rc = check_ready(env);
    if (rc) return rc;
    // PRINT mode (GPUScaffoldSimulation.cu:809-846): the trial's spike files are
    // merged before the state is re-initialized.
    if (env->rec && ((rc = engine_record_sync(env->eng, env->rec)) || (rc = recorder_merge(env->rec, trial))))
        return rc;
    rc = engine_init_params(env->eng, env->opt.neuron_seed, env->init_calls);
    if (rc) return rc;
    env->init_calls++;
    if (env->opt.reset_g) {
        rc = engine_zero_g(env->eng);
        if (rc) return rc;
    }
    return SCAFFOLD_OK;
}

int scaffold_finalize_network_env(scaffold_env *env)
{
    clear_error();
    if (!env) return fail(SCAFFOLD_ERR_INVALID_ARG, "null environment");
    if (env->in_batches > 0)
        return fail(SCAFFOLD_ERR_STATE, "environment is part of a batch; destroy the batch first");
    if (env->rec) {  // the last windows are written; the files are left as they are
        if (env->eng) {
            (void)engine_record_sync(env->eng, env->rec);
            (void)engine_sync(env->eng);
        }
        (void)recorder_finish(env->rec);
        recorder_destroy(env->rec);
        env->rec = nullptr;
    }
    if (env->eng) engine_destroy(env->eng);
    env->eng = nullptr;
    delete env;
    return SCAFFOLD_OK;
}

int scaffold_describe_reference(scaffold_env *env, const char *data_dir, uint64_t stim_seed)
{
    clear_error();
    int rc = check_described(env);
    if (rc) return rc;
    return describe_reference(env, data_dir ? data_dir : ".", stim_seed);
}

int scaffold_describe_synthetic(scaffold_env *env, const scaffold_synth_config *cfg)
{
    clear_error();
    int rc = check_described(env);
    if (rc) return rc;
    if (!cfg) return fail(SCAFFOLD_ERR_INVALID_ARG, "null config");
    return synth_describe(env, cfg);
}

int scaffold_write_dat(scaffold_env *env, int index, const char *filename)
{
    clear_error();
    int rc = check_env(env);
    if (rc) return rc;
    if (index < 0 || index >= env->n_conn || !env->projs[index].set)
        return fail(SCAFFOLD_ERR_INVALID_ARG, "projection %d not registered", index);
    FILE *fp = fopen(filename, "w");
    if (!fp) return fail(SCAFFOLD_ERR_IO, "cannot open %s for writing", filename);
    const Projection &p = env->projs[index];
    const int post_n = (int)p.row_ptr.size() - 1;
    // pre<TAB>post sorted by (post, pre), convert.py:79-81
    std::vector<int32_t> row;
    for (int j = 0; j < post_n; ++j) {
        row.assign(p.col.begin() + p.row_ptr[j], p.col.begin() + p.row_ptr[j + 1]);
        std::sort(row.begin(), row.end());
        for (int32_t c : row) fprintf(fp, "%d\t%d\n", c, j);
    }
    fclose(fp);
    return SCAFFOLD_OK;
}

int scaffold_get_info(scaffold_env *env, scaffold_info *out)
{
    clear_error();
    int rc = check_env(env);
    if (rc) return rc;
    if (!out) return fail(SCAFFOLD_ERR_INVALID_ARG, "null output");
    scaffold_info i{};
    i.n_populations = env->n_types;
    i.n_projections = env->n_conn;
    i.n_stimuli = (int)env->stims.size();
    int total = 0;
    for (auto &p : env->pops) total += p.set ? p.num : 0;
    i.total_neurons = total;
    int64_t nnz = 0;
    int maxd = 0;
    for (auto &p : env->projs)
        if (p.set) {
            nnz += (int64_t)p.col.size();
            if (p.delay() > maxd) maxd = p.delay();
        }
    i.total_synapses = nnz;
    i.max_delay = maxd;
    if (env->initialized) {
        i.device_neurons = env->plan.Ndev;
        i.T_print = env->T_print;
        i.T_cpu_sim = env->T_cpu_sim;
        i.ring_rows = env->plan.R;
        i.steps_since_init = env->draws;
        i.current_step = (int32_t)env->cur_step;
        i.n_output_neurons = env->plan.n_output;
        i.device_bytes = engine_device_bytes(env->eng);
        i.algorithmic_bytes_per_step = env->plan.algo_bytes;
        i.n_tasks = (int)env->plan.tasks.size();
        i.n_groups = env->plan.n_groups;
        i.synapses_16bit = env->plan.syn16;
        engine_blocks(env->eng, &i.n_blocks, &i.n_staged_blocks);
        i.lane_synapses = env->plan.lane_synapses;
        i.stage_ratio = env->plan.stage_ratio;
        i.word_groups = env->plan.n_grp;
        i.word_group_synapses = env->plan.grp_syn;
        i.stream_bytes_per_step = env->plan.stream_bytes;
    }
    *out = i;
    return SCAFFOLD_OK;
}

int scaffold_get_population(scaffold_env *env, int index, scaffold_population_info *out)
{
    clear_error();
    int rc = check_env(env);
    if (rc) return rc;
    if (!out || index < 0 || index >= env->n_types) return fail(SCAFFOLD_ERR_INVALID_ARG, "bad index");
    const Population &p = env->pops[index];
    scaffold_population_info i{};
    snprintf(i.name, sizeof(i.name), "%s", p.name.c_str());
    i.type = p.type;
    i.device_type = p.kind;
    i.num = p.set ? p.num : 0;
    int base = 0;
    for (int k = 0; k < index; ++k) base += env->pops[k].set ? env->pops[k].num : 0;
    i.base_id = base;
    i.c_num = p.c_num;
    i.Cm = p.Cm; i.El = p.El; i.dt_ref = p.dt_ref; i.Ie = p.Ie; i.Vr = p.Vr; i.Vth = p.Vth; i.gL = p.gL;
    for (int c = 0; c < 2; ++c) { i.w[c] = p.w[c]; i.E[c] = p.E[c]; i.g_bar[c] = p.g_bar[c]; i.tau[c] = p.tau[c]; }
    i.ref_step = (int)((double)p.dt_ref / kDT);
    *out = i;
    return SCAFFOLD_OK;
}

int scaffold_get_positions(scaffold_env *env, int index, float *x, float *z)
{
    clear_error();
    int rc = check_env(env);
    if (rc) return rc;
    if (index < 0 || index >= env->n_types) return fail(SCAFFOLD_ERR_INVALID_ARG, "bad index");
    const Population &p = env->pops[index];
    if (!p.set || (int)p.pos_x.size() != p.num) return fail(SCAFFOLD_ERR_STATE, "population %d has no positions", index);
    if (x) memcpy(x, p.pos_x.data(), sizeof(float) * p.pos_x.size());
    if (z) memcpy(z, p.pos_z.data(), sizeof(float) * p.pos_z.size());
    return SCAFFOLD_OK;
}

int scaffold_get_projection(scaffold_env *env, int index, scaffold_projection_info *out)
{
    clear_error();
    int rc = check_env(env);
    if (rc) return rc;
    if (!out || index < 0 || index >= env->n_conn) return fail(SCAFFOLD_ERR_INVALID_ARG, "bad index");
    const Projection &p = env->projs[index];
    scaffold_projection_info i{};
    i.type = p.type;
    i.pre = p.set ? p.pre : -1;
    i.post = p.set ? p.post : -1;
    i.pre_num = p.preNum;
    i.post_num = p.set ? (int)p.row_ptr.size() - 1 : 0;
    i.delay = p.delay();
    i.initial_weight = p.weight;
    i.w_bar = p.weight > 0 ? p.weight : p.weight * (-1.f);
    i.nnz = (int64_t)p.col.size();
    i.max_conv = p.max_conv;
    i.fp8_scale = 1.f;
    i.sum_scale = 1.f;
    for (const Plasticity &r : env->plast) i.plastic |= r.target == index ? 1 : 0;
    if (env->initialized && p.set) {
        const ScProj &sp = env->plan.projs[index];
        i.lane_mode = sp.mode == SC_MODE_LANE;
        i.fp8_scale = sp.scale;
        i.sum_scale = sp.sum_scale;
        for (const PlasticPlan &x : env->plan.plastic)
            if (x.rule.target == index) i.grid_exp = x.grid_exp;
    }
    *out = i;
    return SCAFFOLD_OK;
}

int scaffold_get_projection_csr(scaffold_env *env, int index, int64_t *row_ptr, int32_t *col, float *val)
{
    clear_error();
    int rc = check_env(env);
    if (rc) return rc;
    if (index < 0 || index >= env->n_conn || !env->projs[index].set)
        return fail(SCAFFOLD_ERR_INVALID_ARG, "projection %d not registered", index);
    const Projection &p = env->projs[index];
    if (row_ptr) std::copy(p.row_ptr.begin(), p.row_ptr.end(), row_ptr);
    if (col) std::copy(p.col.begin(), p.col.end(), col);
    if (val) {
        const int64_t nnz = (int64_t)p.col.size();
        if (env->initialized) {
            rc = env->eng ? engine_read_plastic(env->eng, index, val) : -1;
            if (rc > 0) return rc;
            if (rc < 0)  // not plastic: the multipliers as planned
                for (int64_t e = 0; e < nnz; ++e) val[e] = plan_multiplier(env, index, e);
        } else {
            for (int64_t e = 0; e < nnz; ++e) val[e] = p.val.empty() ? 1.f : p.val[e];
        }
    }
    return SCAFFOLD_OK;
}

int scaffold_get_stimulus(scaffold_env *env, int index, scaffold_stim_info *out, int64_t *fanout_ptr,
                          uint32_t *targets)
{
    clear_error();
    int rc = check_env(env);
    if (rc) return rc;
    if (!out || index < 0 || index >= (int)env->stims.size()) return fail(SCAFFOLD_ERR_INVALID_ARG, "bad index");
    const Stimulus &s = env->stims[index];
    out->pop = s.pop;
    out->n_entries = s.n_entries();
    out->n_targets = (int)s.targets.size();
    out->func = s.f;
    out->func.poisson_hz = s.rate;
    if (fanout_ptr) {
        if (s.fanout.empty())
            for (int i = 0; i <= out->n_entries; ++i) fanout_ptr[i] = i;
        else
            std::copy(s.fanout.begin(), s.fanout.end(), fanout_ptr);
    }
    if (targets) std::copy(s.targets.begin(), s.targets.end(), targets);
    return SCAFFOLD_OK;
}

int scaffold_get_state(scaffold_env *env, float *V, float *g, int32_t *refr)
{
    clear_error();
    int rc = check_ready(env);
    if (rc) return rc;
    return engine_get_state(env->eng, V, g, refr);
}

int scaffold_set_state(scaffold_env *env, const float *V, const float *g, const int32_t *refr)
{
    clear_error();
    int rc = check_ready(env);
    if (rc) return rc;
    return engine_set_state(env->eng, V, g, refr);
}

// Ring row of a queryable step: one of the last R - 1 steps.  (Step n clears the
// row of step n + 1, which is the row of step n + 1 - R, so after a run ending
// at cur_step the row of cur_step - R no longer holds that step's spikes.)
static int step_row(scaffold_env *env, int step, int *row)
{
    const int R = env->plan.R;
    if (step < 0 || step >= env->cur_step || step <= env->cur_step - R)
        return fail(SCAFFOLD_ERR_INVALID_ARG, "step %d is not among the last %d steps (next step %lld)", step, R - 1,
                    (long long)env->cur_step);
    *row = step % R;
    return SCAFFOLD_OK;
}

int scaffold_get_spike_row(scaffold_env *env, int step, uint8_t *spikes)
{
    clear_error();
    int rc = check_ready(env);
    if (rc) return rc;
    if (!spikes) return fail(SCAFFOLD_ERR_INVALID_ARG, "null output");
    int row = 0;
    rc = step_row(env, step, &row);
    if (rc) return rc;
    const Plan &P = env->plan;
    env->tmp_words.resize(P.Wr);
    rc = engine_get_ring_row(env->eng, row, env->tmp_words.data());
    if (rc) return rc;
    for (int r = 0; r < P.total_ref; ++r) {
        int d = P.ref_to_dev[r];
        spikes[r] = (uint8_t)((env->tmp_words[d >> 6] >> (d & 63)) & 1u);
    }
    return SCAFFOLD_OK;
}

int scaffold_get_counts(scaffold_env *env, int step, int32_t *counts, uint64_t *events)
{
    clear_error();
    int rc = check_ready(env);
    if (rc) return rc;
    int row = 0;
    rc = step_row(env, step, &row);
    if (rc) return rc;
    unsigned long long ev = 0;
    rc = engine_get_counts(env->eng, row, counts, &ev);
    if (events) *events = ev;
    return rc;
}

int scaffold_get_totals(scaffold_env *env, uint64_t *spikes, uint64_t *events)
{
    clear_error();
    int rc = check_ready(env);
    if (rc) return rc;
    std::vector<unsigned long long> t((size_t)env->n_types + 1, 0ull);
    rc = engine_get_totals(env->eng, t.data());
    if (rc) return rc;
    if (spikes)
        for (int p = 0; p < env->n_types; ++p) spikes[p] = t[p];
    if (events) *events = t[env->n_types];
    return SCAFFOLD_OK;
}

int scaffold_get_fired(scaffold_env *env, int step, int32_t *ids, int32_t cap, int32_t *n)
{
    clear_error();
    int rc = check_ready(env);
    if (rc) return rc;
    if (!n || (cap > 0 && !ids)) return fail(SCAFFOLD_ERR_INVALID_ARG, "null output");
    int row = 0;
    rc = step_row(env, step, &row);
    if (rc) return rc;
    return engine_get_fired(env->eng, row, ids, cap, n);
}

int scaffold_get_fired_noalloc(scaffold_env *env, int step, int32_t *ids, int32_t cap, int32_t *n)
{
    clear_error();
    int rc = check_ready(env);
    if (rc) return rc;
    if (!n || (cap > 0 && !ids)) return fail(SCAFFOLD_ERR_INVALID_ARG, "null output");
    int row = 0;
    rc = step_row(env, step, &row);
    if (rc) return rc;
    return engine_get_fired_noalloc(env->eng, row, ids, cap, n);
}

// The spike lines of steps [step0, step0 + n) in the reference's text format.
// output == 0: the lines print_spike_train writes for the populations simulated
// on the device (GPUScaffoldSimulation.cu:76-84): "%f\t%d\t%d\n" = (step as ms,
// population-local id, ReverseNeuronTypeID[internal index]); the printer's table
// is ReverseNeuronTypeID (:179-183, :429), the inverse of NeuronTypeID, so the
// last column is the population's NeuronType -- the `type` it was registered
// with.  OUTPUT populations are skipped: the reference simulates them on the
// host, and their device bytes stay zero.
// output == 1: the lines host_update_lif_ch writes for the OUTPUT populations
// (Simulation_header.cu:456): (step as ms, global id, NeuronType), through the
// same inverse table (:539).
static int write_spike_lines(scaffold_env *env, int step0, int n, const char *filename, int output)
{
    int rc = check_ready(env);
    if (rc) return rc;
    if (!filename || n < 0) return fail(SCAFFOLD_ERR_INVALID_ARG, "bad arguments");
    for (int s = step0; s < step0 + n; ++s) {  // every step queryable before anything is written
        int row = 0;
        if ((rc = step_row(env, s, &row))) return rc;
    }
    FILE *fp = fopen(filename, "a");
    if (!fp) return fail(SCAFFOLD_ERR_IO, "cannot open %s", filename);
    const Plan &P = env->plan;
    env->tmp_words.resize(P.Wr);
    for (int s = step0; s < step0 + n; ++s) {
        int row = 0;
        rc = step_row(env, s, &row);
        if (!rc) rc = engine_get_ring_row(env->eng, row, env->tmp_words.data());
        if (rc) { fclose(fp); return rc; }
        const float t = (float)s;  // CTYPE t = (n + step), :78
        for (int p = 0; p < (int)P.pops.size(); ++p) {
            const ScPop &sp = P.pops[p];
            if ((sp.kind == SCAFFOLD_OUTPUT) != (output != 0)) continue;
            for (int i = 0; i < sp.num; ++i) {
                int d = sp.dev_base + i;
                if ((env->tmp_words[d >> 6] >> (d & 63)) & 1u)
                    fprintf(fp, "%f\t%d\t%d\n", t, output ? sp.ref_base + i : i, env->pops[p].type);
            }
        }
    }
    fclose(fp);
    return SCAFFOLD_OK;
}

int scaffold_write_spike_text(scaffold_env *env, int step0, int n, const char *filename)
{
    clear_error();
    return write_spike_lines(env, step0, n, filename, 0);
}

int scaffold_write_output_spike_text(scaffold_env *env, int step0, int n, const char *filename)
{
    clear_error();
    return write_spike_lines(env, step0, n, filename, 1);
}

int scaffold_record_spikes(scaffold_env *env, const char *dir)
{
    clear_error();
    int rc = check_ready(env);
    if (rc) return rc;
    if (env->in_batches > 0) return fail(SCAFFOLD_ERR_STATE, "environment is part of a batch");
    if (env->rec) {  // stop the current recording: its windows are written first
        rc = engine_record_sync(env->eng, env->rec);
        if (!rc) rc = engine_sync(env->eng);
        const int frc = recorder_finish(env->rec);
        recorder_destroy(env->rec);
        env->rec = nullptr;
        if (rc || frc) return rc ? rc : frc;
    }
    if (!dir) return SCAFFOLD_OK;
    const Plan &P = env->plan;
    std::vector<int32_t> types;
    for (size_t p = 0; p < P.pops.size(); ++p) types.push_back(env->pops[p].type);
    return recorder_create(&env->rec, dir, env->T_cpu_sim, env->T_print, P.R, P.Wr, P.pops, types);
}

int scaffold_step_timer(scaffold_env *env, int on, double *ms, int64_t *launches)
{
    clear_error();
    int rc = check_ready(env);
    if (rc) return rc;
    if (on < -1 || on > 1000) return fail(SCAFFOLD_ERR_INVALID_ARG, "on is -1, 0, or n >= 1 (one run in n timed)");
    return engine_timer(env->eng, on, ms, launches);
}

int scaffold_time_steps(scaffold_env *env, int n_steps, double *ms_total, double *ms_step_kernel_mean)
{
    clear_error();
    int rc = check_ready(env);
    if (rc) return rc;
    if (n_steps <= 0) return fail(SCAFFOLD_ERR_INVALID_ARG, "n_steps must be positive");
    if (env->rec && (rc = engine_record_sync(env->eng, env->rec))) return rc;  // these steps are not recorded
    rc = engine_time(env->eng, n_steps, env->opt.graph, env->opt.graph_steps, ms_total, ms_step_kernel_mean);
    if (rc) return rc;
    env->cur_step += n_steps;
    env->draws += n_steps;
    return SCAFFOLD_OK;
}

// One trial engine of the network under the options in env->opt: built,
// initialized like InitializeNetworkEnv, warmed up, timed (and its step kernels
// timed with the step-kernel timer when kernel_ms is given), destroyed.
static int trial_time(scaffold_env *env, int T_print, int n_steps, double *ms, double *kernel_ms = nullptr,
                      int *lane_synapses = nullptr, int *word_groups = nullptr)
{
    Plan P;
    P.R = 2 * T_print;
    int rc = plan_build(env, P);
    if (rc) return rc;
    if (lane_synapses) *lane_synapses = P.lane_synapses;  // the values the options resolve to
    if (word_groups) *word_groups = P.group_min;
    Engine *e = nullptr;
    rc = engine_create(&e, P);
    if (rc) return rc;
    rc = engine_init_params(e, env->opt.neuron_seed, 0);
    if (!rc) rc = engine_init_params(e, env->opt.neuron_seed, 1);
    if (!rc) rc = engine_set_ctl(e, 0, 0);
    if (!rc) rc = engine_run(e, 50, env->opt.graph, env->opt.graph_steps);
    double total = 0.0;
    if (!rc && kernel_ms) rc = engine_timer(e, 1, nullptr, nullptr);
    if (!rc) rc = engine_time(e, n_steps, env->opt.graph, env->opt.graph_steps, &total, nullptr);
    if (!rc && kernel_ms) {
        double kms = 0.0;
        int64_t launches = 0;
        rc = engine_timer(e, 0, &kms, &launches);
        *kernel_ms = kms / (double)std::max<int64_t>(1, launches);
    }
    engine_destroy(e);
    if (!rc) *ms = total / n_steps;
    return rc;
}

// The spike ring of a trial engine: T_print of at least 50 steps and the longest delay.
static int trial_T_print(scaffold_env *env)
{
    float max_delay = 0.f;
    for (auto &p : env->projs)
        if (p.set && p.delay_ms > max_delay) max_delay = p.delay_ms;
    return std::max(50, (int)max_delay);
}

int scaffold_time_layout(scaffold_env *env, int n_steps, double *ms_per_step, double *kernel_ms)
{
    clear_error();
    int rc = check_described(env);
    if (rc) return rc;
    if (n_steps <= 0 || !ms_per_step) return fail(SCAFFOLD_ERR_INVALID_ARG, "bad arguments");
    return trial_time(env, trial_T_print(env), n_steps, ms_per_step, kernel_ms);
}

int scaffold_autotune(scaffold_env *env, int n_steps, scaffold_tune_result *out)
{
    clear_error();
    int rc = check_described(env);
    if (rc) return rc;
    if (!out || n_steps <= 0) return fail(SCAFFOLD_ERR_INVALID_ARG, "bad arguments");
    const int T_print = trial_T_print(env);
    const Options saved = env->opt;
    scaffold_tune_result r{};
    r.ms_per_step = 1e300;
    r.ms_default = -1.0;
    const int fanins[] = {16, 48, 128}, synapses[] = {8, 16, 32, 64};
    // the environment's own options are the first candidate
    r.lane_max_fanin = saved.lane_max_fanin;
    r.lane_synapses = saved.lane_synapses;
    r.graph_steps = saved.graph_steps;
    r.index_bits = saved.index_bits;
    r.stage_spikes = saved.stage_spikes;
    r.word_groups = saved.word_groups;
    // the winner so far as options, with one field changed
    auto trial = [&](auto &&change, double *ms) {
        env->opt = saved;
        env->opt.lane_max_fanin = r.lane_max_fanin;
        env->opt.lane_synapses = r.lane_synapses;
        env->opt.graph_steps = r.graph_steps;
        env->opt.index_bits = r.index_bits;
        env->opt.stage_spikes = r.stage_spikes;
        env->opt.word_groups = r.word_groups;
        change(env->opt);
        const int trc = trial_time(env, T_print, n_steps, ms);
        env->opt = saved;
        if (!trc) ++r.n_tried;
        return trc;
    };
    // the default trial reports the team width its options resolve to (0 picks by
    // network size), so that the result can be applied to other networks as is
    env->opt = saved;
    if ((rc = trial_time(env, T_print, n_steps, &r.ms_default, nullptr, &r.lane_synapses, &r.word_groups))) {
        env->opt = saved;
        return rc;
    }
    ++r.n_tried;
    r.ms_per_step = r.ms_default;
    for (int lf : fanins)
        for (int ls : synapses) {
            double ms = 0.0;
            if ((rc = trial([&](Options &o) { o.lane_max_fanin = lf; o.lane_synapses = ls; o.graph_steps = 25; }, &ms)))
                return rc;
            if (ms < r.ms_per_step) {
                r.ms_per_step = ms;
                r.lane_max_fanin = lf;
                r.lane_synapses = ls;
                r.graph_steps = 25;
            }
        }
    double ms = 0.0;
    const int other_bits = r.index_bits == 16 ? 32 : 16;
    if ((rc = trial([&](Options &o) { o.index_bits = other_bits; }, &ms))) return rc;
    if (ms < r.ms_per_step) { r.ms_per_step = ms; r.index_bits = other_bits; }
    const int other_stage = r.stage_spikes ? 0 : 1;
    if ((rc = trial([&](Options &o) { o.stage_spikes = other_stage; }, &ms))) return rc;
    if (ms < r.ms_per_step) { r.ms_per_step = ms; r.stage_spikes = other_stage; }
    if ((rc = trial([&](Options &o) { o.graph_steps = 50; }, &ms))) return rc;
    if (ms < r.ms_per_step) { r.ms_per_step = ms; r.graph_steps = 50; }
    // word groups pay or cost with the firing rates; off, or on at the threshold
    // the network's size picks when the environment had them off
    int other_groups = 0;
    if (r.word_groups == 0) {
        env->opt = saved;
        env->opt.word_groups = -1;
        Plan Q;
        Q.R = 2 * T_print;
        if ((rc = plan_build(env, Q))) { env->opt = saved; return rc; }
        other_groups = Q.group_min;
        env->opt = saved;
    }
    if (other_groups != r.word_groups) {
        if ((rc = trial([&](Options &o) { o.word_groups = other_groups; }, &ms))) return rc;
        if (ms < r.ms_per_step) { r.ms_per_step = ms; r.word_groups = other_groups; }
    } else {
        ++r.n_tried;  // the same layout: nothing to time
    }
    env->opt = saved;
    env->opt.lane_max_fanin = r.lane_max_fanin;
    env->opt.lane_synapses = r.lane_synapses;
    env->opt.graph_steps = r.graph_steps;
    env->opt.index_bits = r.index_bits;
    env->opt.stage_spikes = r.stage_spikes;
    env->opt.word_groups = r.word_groups;
    *out = r;
    return SCAFFOLD_OK;
}

int scaffold_encode_multipliers(int precision, const float *in, int64_t n, float *out, float *scale,
                                float *sum_scale)
{
    clear_error();
    if (precision < SCAFFOLD_WEIGHTS_F32 || precision > SCAFFOLD_WEIGHTS_FP8)
        return fail(SCAFFOLD_ERR_INVALID_ARG, "precision %d: use SCAFFOLD_WEIGHTS_*", precision);
    if (n < 0 || (n > 0 && (!in || !out))) return fail(SCAFFOLD_ERR_INVALID_ARG, "bad arguments");
    for (int64_t e = 0; e < n; ++e)
        if (!std::isfinite(in[e])) return fail(SCAFFOLD_ERR_INVALID_ARG, "multiplier %lld is not finite", (long long)e);
    if (n > 0 && out != in) memmove(out, in, sizeof(float) * (size_t)n);
    const int wp = precision == SCAFFOLD_WEIGHTS_BF16 ? SC_W_BF16 : precision == SCAFFOLD_WEIGHTS_FP8 ? SC_W_FP8 : SC_W_F32;
    float sc = 1.f, ss = 1.f;
    encode_values(wp, out, n, &sc, &ss);
    if (scale) *scale = sc;
    if (sum_scale) *sum_scale = ss;
    return SCAFFOLD_OK;
}

int scaffold_copy_probe(int64_t bytes, int iters, double *ms)
{
    clear_error();
    if (bytes <= 0 || iters <= 0 || !ms) return fail(SCAFFOLD_ERR_INVALID_ARG, "bad arguments");
    return engine_copy_probe(bytes, iters, ms);
}

// ---------------------------------------------------------------- batching
}  // extern "C"

struct scaffold_batch {
    std::vector<scaffold_env *> envs;
    sc::Batch *b = nullptr;
    int graph_steps = 10;
};

extern "C" {

int scaffold_batch_create(scaffold_batch **batch)
{
    clear_error();
    if (!batch) return fail(SCAFFOLD_ERR_INVALID_ARG, "null batch pointer");
    *batch = new (std::nothrow) scaffold_batch();
    return *batch ? SCAFFOLD_OK : fail(SCAFFOLD_ERR_CAPACITY, "out of host memory");
}

int scaffold_batch_add(scaffold_batch *batch, scaffold_env *env)
{
    clear_error();
    if (!batch) return fail(SCAFFOLD_ERR_INVALID_ARG, "null batch");
    if (batch->b) return fail(SCAFFOLD_ERR_STATE, "batch already built");
    int rc = check_ready(env);
    if (rc) return rc;
    for (auto *e : batch->envs)
        if (e == env) return fail(SCAFFOLD_ERR_INVALID_ARG, "environment added twice");
    if (!env->plast.empty())
        return fail(SCAFFOLD_ERR_UNSUPPORTED, "environments with plasticity rules cannot join a batch");
    if (env->rec) return fail(SCAFFOLD_ERR_STATE, "a recording environment cannot join a batch");
    batch->envs.push_back(env);
    env->in_batches++;
    return SCAFFOLD_OK;
}

int scaffold_batch_build(scaffold_batch *batch)
{
    clear_error();
    if (!batch || batch->envs.empty()) return fail(SCAFFOLD_ERR_INVALID_ARG, "empty batch");
    if (batch->b) return fail(SCAFFOLD_ERR_STATE, "batch already built");
    std::vector<Engine *> engines;
    std::vector<const Plan *> plans;
    for (auto *e : batch->envs) {
        engines.push_back(e->eng);
        plans.push_back(&e->plan);
    }
    batch->graph_steps = batch->envs[0]->opt.graph_steps;
    return batch_build(&batch->b, engines, plans);
}

static int batch_advance(scaffold_batch *batch, int n_steps, double *ms)
{
    if (!batch || !batch->b) return fail(SCAFFOLD_ERR_STATE, "batch not built");
    if (n_steps < 0) return fail(SCAFFOLD_ERR_INVALID_ARG, "negative step count");
    for (auto *e : batch->envs) {
        int rc = check_ready(e);
        if (rc) return rc;
    }
    int rc = batch_run(batch->b, n_steps, batch->graph_steps, ms);
    if (rc) return rc;
    for (auto *e : batch->envs) {
        e->cur_step += n_steps;
        e->draws += n_steps;
    }
    return SCAFFOLD_OK;
}

int scaffold_batch_run(scaffold_batch *batch, int n_steps)
{
    clear_error();
    return batch_advance(batch, n_steps, nullptr);
}

int scaffold_batch_time(scaffold_batch *batch, int n_steps, double *ms_total)
{
    clear_error();
    return batch_advance(batch, n_steps, ms_total);
}

int scaffold_batch_destroy(scaffold_batch *batch)
{
    clear_error();
    if (!batch) return fail(SCAFFOLD_ERR_INVALID_ARG, "null batch");
    batch_destroy(batch->b);
    for (auto *e : batch->envs) e->in_batches--;
    delete batch;
    return SCAFFOLD_OK;
}

}  // extern "C"
