// engine.hip -- gfx950 kernels and device-side engine of the exchange-step hot path.
//
// One exchange step (1 ms = 10 Euler sub-steps of 0.1 ms) is two kernels:
//   k_gather   one wavefront per task:
//                STIM  64 INPUT neurons: counter-based Philox4x32-10 draws, stimulus
//                      rule, ballot -> spike word;
//                LANE  64 post neurons, one lane each, over the sliced-ELL streams of
//                      every lane-mode projection; groups without row-mode input
//                      run the LIF update right here (ballot -> spike word);
//                ROW   one chunk of one long row (coalesced col/weight streams),
//                      wavefront reduction -> f64 partial slot.
//   k_finalize one wavefront per group that has row-mode input: combines its
//              partial slots in projection order, LIF update, ballot -> spike word.
// The step reads only ring rows of earlier steps (every delay >= 1) and writes
// the row of the current step, so the kernels of one step have no intra-kernel
// dependencies; the kernel boundary orders consecutive steps.  Runs of steps are
// captured once as a hipGraph and replayed; a device-side step counter
// (ScCtl) advanced by a one-thread kernel at the end of each run gives every
// replay its step numbers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "model.hpp"

namespace sc {

#define HIP_TRY(x)                                                                                   \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess)                                                                        \
            return fail(SCAFFOLD_ERR_HIP, "%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

// ------------------------------------------------------------------ device math
// Philox4x32-10 (Salmon et al., SC'11), counter (c0,c1,c2,c3), key (k0,k1).
__device__ __forceinline__ void philox10(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3,
                                         uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// u32 -> (0,1]: 2^-32 + v * 2^-32 in float.
__device__ __forceinline__ float u01(uint32_t v)
{
    const float inv = 2.3283064365386963e-10f;
    return inv + (float)v * inv;
}

__device__ __forceinline__ uint32_t spike_bit(const uint32_t *__restrict__ row32, uint32_t c)
{
    return (row32[c >> 5] >> (c & 31)) & 1u;
}

template <int WP>
__device__ __forceinline__ double wload(const void *__restrict__ val, int64_t e)
{
    if constexpr (WP == SC_W_F32) {
        return (double)static_cast<const float *>(val)[e];
    } else if constexpr (WP == SC_W_BF16) {
        return (double)__uint_as_float((uint32_t) static_cast<const uint16_t *>(val)[e] << 16);
    } else if constexpr (WP == SC_W_FP8) {
        const uint32_t c = static_cast<const uint8_t *>(val)[e];
        const uint32_t ex = (c >> 3) & 15u, mt = c & 7u;
        float v = ex ? __uint_as_float(((ex + 120u) << 23) | (mt << 20)) : (float)mt * 0x1p-9f;
        return (double)((c & 0x80u) ? -v : v);
    } else {
        return 1.0;
    }
}

__device__ __forceinline__ double wave_sum_f64(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ int ring_row(int64_t n, int R)
{
    int64_t r = n % R;
    return (int)(r < 0 ? r + R : r);
}

// One exchange step of the conductance LIF neuron, update_lif_ch
// (Simulation_header.cu:315-377), every expression promoted as the reference
// writes it; the library is compiled with -ffp-contract=off so no product is
// fused into a following sum.
__device__ __forceinline__ int lif_step(const ScPop &P, float &u, float &g0, float &g1, int32_t &r,
                                        float dge, float dgi)
{
    const int cn = P.c_num;
    const float inj0 = P.gw[0] * (P.exc_ch[0] ? dge : dgi);
    const float inj1 = P.gw[1] * (P.exc_ch[1] ? dge : dgi);
    int spk = 0;
    for (int k = 0; k < kStepMax; ++k) {
        float Isyn = 0.f;
        if (cn > 0) Isyn = Isyn + g0 * (u - P.E[0]);
        if (cn > 1) Isyn = Isyn + g1 * (u - P.E[1]);
        const float leak = P.gL * (u - P.El);
        const float inner = (P.Ie - leak) - Isyn;
        const float du = (float)(P.dtc * (double)inner);
        if (cn > 0) {
            const double dec = (-P.dtt[0]) * (double)g0;
            g0 = (float)((double)g0 + (k == 0 ? dec + (double)inj0 : dec));
        }
        if (cn > 1) {
            const double dec = (-P.dtt[1]) * (double)g1;
            g1 = (float)((double)g1 + (k == 0 ? dec + (double)inj1 : dec));
        }
        const int fired = u > P.Vth;
        spk |= fired;
        r = fired ? P.ref_step : (r > 0 ? r - 1 : 0);
        u = r ? P.Vr : u + du;
    }
    return spk;
}

// Spike word, per-population count and synaptic events of one group.
__device__ __forceinline__ void emit_group(const ScInstance &I, int gid, int pop, int d, int lane, int spk,
                                           int row_n)
{
    const unsigned long long word = __ballot(spk != 0);
    if (lane == 0) I.ring[(int64_t)row_n * I.Wr + gid] = word;
    if (word) {
        const unsigned long long ev = wave_sum_u64(spk ? (unsigned long long)I.outdeg[d] : 0ull);
        if (lane == 0) {
            atomicAdd(&I.counts[(int64_t)row_n * I.n_pops + pop], (int)__popcll(word));
            atomicAdd(&I.events[row_n], ev);
        }
    }
}

__device__ __forceinline__ void finalize_group(const ScInstance &I, int gid, const ScGroup &G, const ScPop &P,
                                               int lane, float dge, float dgi, int row_n)
{
    const int d = gid * SC_WAVE + lane;
    int spk = 0;
    if (lane < G.nvalid) {
        float u = I.V[d];
        float2 gg = reinterpret_cast<const float2 *>(I.g)[d];
        int32_t r = I.refr[d];
        spk = lif_step(P, u, gg.x, gg.y, r, dge, dgi);
        I.V[d] = u;
        reinterpret_cast<float2 *>(I.g)[d] = gg;
        I.refr[d] = r;
    }
    emit_group(I, gid, G.pop, d, lane, spk, row_n);
}

// Sliced-ELL row of `lane` in post group gi.
template <int WP>
__device__ __forceinline__ double lane_sum(const ScProj &pr, const uint32_t *__restrict__ row32, int gi, int lane)
{
    const int64_t off = pr.sell_off[gi] + lane;
    const int w = pr.sell_w[gi];
    const uint32_t *__restrict__ col = pr.col;
    double S = 0.0;
    for (int k = 0; k < w; ++k) {
        const int64_t e = off + (int64_t)k * SC_WAVE;
        if (spike_bit(row32, col[e])) S += wload<WP>(pr.val, e);
    }
    return S;
}

// One chunk of a CSR row, lanes striding over it.
template <int WP>
__device__ __forceinline__ double chunk_sum(const ScProj &pr, const uint32_t *__restrict__ row32, int64_t e0,
                                            int len, int lane)
{
    const uint32_t *__restrict__ col = pr.col + e0;
    double S = 0.0;
    for (int i = lane; i < len; i += SC_WAVE)
        if (spike_bit(row32, col[i])) S += wload<WP>(pr.val, e0 + i);
    return S;
}

template <int MODE>
__device__ __forceinline__ double proj_sum(const ScProj &pr, const uint32_t *row32, int gi, int64_t e0, int len,
                                           int lane)
{
    switch (pr.wprec) {
    case SC_W_F32: return MODE == SC_MODE_LANE ? lane_sum<SC_W_F32>(pr, row32, gi, lane) : chunk_sum<SC_W_F32>(pr, row32, e0, len, lane);
    case SC_W_BF16: return MODE == SC_MODE_LANE ? lane_sum<SC_W_BF16>(pr, row32, gi, lane) : chunk_sum<SC_W_BF16>(pr, row32, e0, len, lane);
    case SC_W_FP8: return MODE == SC_MODE_LANE ? lane_sum<SC_W_FP8>(pr, row32, gi, lane) : chunk_sum<SC_W_FP8>(pr, row32, e0, len, lane);
    default: return MODE == SC_MODE_LANE ? lane_sum<SC_W_UNIT>(pr, row32, gi, lane) : chunk_sum<SC_W_UNIT>(pr, row32, e0, len, lane);
    }
}

// Whether one stimulus entry fires in exchange step n (draw index m):
// InputStimulation (UserInputFunctions.cu:94-118) with the rule of ScStim.
__device__ int stim_fires(const ScStim &S, int entry, int64_t n, int64_t m)
{
    float r[12];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        const uint64_t blk = (uint64_t)(3 * m + b);
        uint32_t c0 = (uint32_t)blk, c1 = (uint32_t)(blk >> 32), c2 = (uint32_t)entry, c3 = 0u;
        philox10(c0, c1, c2, c3, S.key0, S.key1);
        r[4 * b + 0] = u01(c0);
        r[4 * b + 1] = u01(c1);
        r[4 * b + 2] = u01(c2);
        r[4 * b + 3] = u01(c3);
    }
    int spk = 0;
    const double nf = (double)(float)n;
#pragma unroll
    for (int k = 0; k < kStepMax; ++k) {
        const float t = (float)(nf + (double)(float)k * kDT);
        const double td = (double)t;
        const double rk = (double)r[k];
        int f = 0;
        if (S.poisson_on <= td && td < S.poisson_off && rk < S.p_thr) f = 1;
        if (S.burst_on <= td && td < S.burst_off && rk < S.b_thr) f = 1;
        if (S.periodic_hz != 0.f && S.periodic_on <= td && (S.periodic_incl ? td <= S.periodic_off : td < S.periodic_off)) {
            // PeriodicFiring (UserInputFunctions.cu:8-10)
            const float dd = t - S.periodic_start;
            const int a = (int)(((double)dd - kDT) * (double)S.periodic_hz / kSec);
            const float fp = dd * S.periodic_hz;
            const int b = (int)((double)fp / kSec);
            if (a != b) f = 1;
        }
        spk |= f;
    }
    return spk;
}

// ---------------------------------------------------------------------- kernels
__global__ void __launch_bounds__(256) k_gather(const ScTask *__restrict__ tasks, int n_tasks,
                                                const ScInstance *__restrict__ insts, int k)
{
    const int lane = threadIdx.x & (SC_WAVE - 1);
    const int wid = blockIdx.x * (blockDim.x / SC_WAVE) + (threadIdx.x / SC_WAVE);
    if (wid >= n_tasks) return;
    const ScTask t = tasks[wid];
    const ScInstance &I = insts[t.inst];
    const int64_t n = I.ctl->n + k;
    const int R = I.R;
    const int row_n = ring_row(n, R);
    if (t.kind == SC_TASK_ROW) {
        const ScProj &pr = I.projs[t.proj];
        const uint32_t *row32 = reinterpret_cast<const uint32_t *>(I.ring + (int64_t)ring_row(n - pr.delay, R) * I.Wr);
        double S = proj_sum<SC_MODE_ROW>(pr, row32, 0, t.e0, t.len, lane);
        S = wave_sum_f64(S);
        if (lane == 0) I.partial[pr.slot + (int64_t)t.post * pr.maxch + t.chunk_idx] = S;
    } else if (t.kind == SC_TASK_LANE) {
        const ScGroup G = I.groups[t.group];
        const ScPop &P = I.pops[G.pop];
        const int gi = t.group - P.group_base;
        const int j = G.first_local + lane;
        float dge = 0.f, dgi = 0.f;
        for (int q = 0; q < P.n_in; ++q) {
            const ScProj &pr = I.projs[P.in_proj[q]];
            if (pr.mode != SC_MODE_LANE) continue;
            const uint32_t *row32 =
                reinterpret_cast<const uint32_t *>(I.ring + (int64_t)ring_row(n - pr.delay, R) * I.Wr);
            const double S = proj_sum<SC_MODE_LANE>(pr, row32, gi, 0, 0, lane);
            if (G.self_final) {
                // dg[post] += S * w_bar (Simulation_header.cu:98), projections in index order
                const float Sf = (float)(S * (double)pr.scale);
                const float add = Sf * pr.W;
                if (pr.exc) dge = dge + add;
                else dgi = dgi + add;
            } else if (lane < G.nvalid) {
                I.partial[pr.slot + j] = S;
            }
        }
        if (G.self_final) finalize_group(I, t.group, G, P, lane, dge, dgi, row_n);
    } else {
        const ScGroup G = I.groups[t.group];
        const int d = t.group * SC_WAVE + lane;
        int spk = 0;
        if (lane < G.nvalid) {
            const int s = I.stim_s[d];
            if (s >= 0) spk = stim_fires(I.stims[s], I.stim_e[d], n, I.ctl->m + k);
        }
        emit_group(I, t.group, G.pop, d, lane, spk, row_n);
    }
}

__global__ void __launch_bounds__(256) k_finalize(const int32_t *__restrict__ fin, int n_fin,
                                                  const ScInstance *__restrict__ insts, int n_inst, int k)
{
    if (blockIdx.x == 0) {
        // clear the count row of the next step (its gather kernel starts after this one ends)
        for (int idx = threadIdx.x; idx < n_inst * (SC_MAX_POPS + 1); idx += blockDim.x) {
            const int i = idx / (SC_MAX_POPS + 1), p = idx % (SC_MAX_POPS + 1);
            const ScInstance &I = insts[i];
            const int row1 = ring_row(I.ctl->n + k + 1, I.R);
            if (p < I.n_pops) I.counts[(int64_t)row1 * I.n_pops + p] = 0;
            else if (p == SC_MAX_POPS) I.events[row1] = 0ull;
        }
    }
    const int lane = threadIdx.x & (SC_WAVE - 1);
    const int wid = blockIdx.x * (blockDim.x / SC_WAVE) + (threadIdx.x / SC_WAVE);
    if (wid >= n_fin) return;
    const ScInstance &I = insts[fin[2 * wid]];
    const int gid = fin[2 * wid + 1];
    const int64_t n = I.ctl->n + k;
    const ScGroup G = I.groups[gid];
    const ScPop &P = I.pops[G.pop];
    const int j = G.first_local + lane;
    const bool valid = lane < G.nvalid;
    float dge = 0.f, dgi = 0.f;
    for (int q = 0; q < P.n_in; ++q) {
        const ScProj &pr = I.projs[P.in_proj[q]];
        double S = 0.0;
        if (valid) {
            if (pr.mode == SC_MODE_LANE) {
                S = I.partial[pr.slot + j];
            } else {
                const int64_t len = pr.row_ptr[j + 1] - pr.row_ptr[j];
                const int nch = (int)((len + pr.chunk - 1) / pr.chunk);
                const double *ps = I.partial + pr.slot + (int64_t)j * pr.maxch;
                for (int c = 0; c < nch; ++c) S += ps[c];
            }
        }
        const float Sf = (float)(S * (double)pr.scale);
        const float add = Sf * pr.W;
        if (pr.exc) dge = dge + add;
        else dgi = dgi + add;
    }
    finalize_group(I, gid, G, P, lane, dge, dgi, ring_row(n, I.R));
}

__global__ void k_advance(const ScInstance *__restrict__ insts, int n_inst, int K)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_inst) {
        insts[i].ctl->n += K;
        insts[i].ctl->m += K;
    }
}

__global__ void k_set_ctl(ScCtl *ctl, int64_t n, int64_t m)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl->n = n;
        ctl->m = m;
    }
}

// InitParams (init_params.cu:343-352): u = Vr + (Vth - Vr) * U with U the word
// `off` of the Philox stream (seed, subsequence = reference global id); refractory
// counters and ring row 0 cleared.
__global__ void k_init_params(const ScInstance *__restrict__ ip, uint32_t k0, uint32_t k1, uint64_t off)
{
    const ScInstance &I = *ip;
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d < I.Wr) I.ring[d] = 0ull;
    if (d >= I.Ndev) return;
    const ScGroup &G = I.groups[d / SC_WAVE];
    const int lane = d & (SC_WAVE - 1);
    if (lane >= G.nvalid) return;
    const ScPop &P = I.pops[G.pop];
    const uint32_t ref = (uint32_t)(P.ref_base + G.first_local + lane);
    const uint64_t blk = off >> 2;
    uint32_t c0 = (uint32_t)blk, c1 = (uint32_t)(blk >> 32), c2 = ref, c3 = 0u;
    philox10(c0, c1, c2, c3, k0, k1);
    const uint32_t w = (off & 3) == 0 ? c0 : (off & 3) == 1 ? c1 : (off & 3) == 2 ? c2 : c3;
    const float prod = P.Vr_span * u01(w);
    I.V[d] = P.Vr + prod;
    I.refr[d] = 0;
}

// Fired-neuron list of one ring row: the spike words are the ballots of spike
// detection; one block popcounts them, prefix-sums the counts and writes the
// reference global ids in ascending order.
__global__ void __launch_bounds__(1024) k_compact(const uint64_t *__restrict__ row, int n_groups,
                                                  const ScGroup *__restrict__ groups,
                                                  const ScPop *__restrict__ pops, int32_t *__restrict__ out,
                                                  int cap, int32_t *__restrict__ count)
{
    __shared__ int scan[1024];
    const int t = threadIdx.x;
    const int per = (n_groups + 1023) / 1024;
    const int g0 = min(n_groups, t * per), g1 = min(n_groups, g0 + per);
    int c = 0;
    for (int g = g0; g < g1; ++g) c += __popcll(row[g]);
    scan[t] = c;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int v = t >= o ? scan[t - o] : 0;
        __syncthreads();
        scan[t] += v;
        __syncthreads();
    }
    int pos = scan[t] - c;
    if (t == 1023) *count = scan[1023];
    for (int g = g0; g < g1; ++g) {
        unsigned long long w = row[g];
        const ScGroup G = groups[g];
        const int base = pops[G.pop].ref_base + G.first_local;
        while (w) {
            const int b = __ffsll((long long)w) - 1;
            w &= w - 1;
            if (pos < cap) out[pos] = base + b;
            ++pos;
        }
    }
}

__global__ void k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ----------------------------------------------------------------------- engine
struct Engine {
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    int64_t bytes = 0;
    ScInstance hi{};
    ScInstance *di = nullptr;
    ScTask *tasks = nullptr;
    int n_tasks = 0;
    int32_t *fin = nullptr;
    int n_fin = 0;
    ScStim *stims = nullptr;
    int n_stims = 0;
    int n_pops = 0;
    int32_t *cnt_buf = nullptr;  // compaction count
    std::vector<int32_t> dev_to_ref;
    int Ndev = 0, total_ref = 0;
    hipGraphExec_t gexec = nullptr;
    int gexec_steps = 0;
};

template <typename T>
static int dalloc(Engine *e, T **p, size_t count)
{
    size_t b = sizeof(T) * (count ? count : 1);
    void *q = nullptr;
    HIP_TRY(hipMalloc(&q, b));
    HIP_TRY(hipMemset(q, 0, b));
    e->allocs.push_back(q);
    e->bytes += (int64_t)b;
    *p = static_cast<T *>(q);
    return SCAFFOLD_OK;
}

template <typename T>
static int dupload(Engine *e, T **p, const std::vector<T> &v)
{
    int rc = dalloc(e, p, v.size());
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

int engine_create(Engine **out, Plan &P)
{
    int rc = check_device();
    if (rc) return rc;
    Engine *e = new Engine();
    *out = nullptr;
    auto bail = [&](int code) { engine_destroy(e); return code; };
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(SCAFFOLD_ERR_HIP, "hipStreamCreate failed"));
    ScInstance &I = e->hi;
    I.R = P.R;
    I.Wr = P.Wr;
    I.n_groups = P.n_groups;
    I.n_pops = (int)P.pops.size();
    I.n_projs = (int)P.projs.size();
    I.n_stims = (int)P.stims.size();
    I.Ndev = P.Ndev;
    e->n_pops = I.n_pops;
    e->Ndev = P.Ndev;
    e->total_ref = P.total_ref;
    e->dev_to_ref = P.dev_to_ref;
    if ((rc = dalloc(e, &I.V, P.Ndev))) return bail(rc);
    if ((rc = dalloc(e, &I.g, 2 * (size_t)P.Ndev))) return bail(rc);
    if ((rc = dalloc(e, &I.refr, P.Ndev))) return bail(rc);
    if ((rc = dalloc(e, &I.ring, (size_t)P.R * P.Wr))) return bail(rc);
    if ((rc = dalloc(e, &I.partial, (size_t)P.n_slots))) return bail(rc);
    if ((rc = dalloc(e, &I.counts, (size_t)P.R * I.n_pops))) return bail(rc);
    if ((rc = dalloc(e, &I.events, (size_t)P.R))) return bail(rc);
    ScGroup *groups;
    if ((rc = dupload(e, &groups, P.groups))) return bail(rc);
    I.groups = groups;
    ScPop *pops;
    if ((rc = dupload(e, &pops, P.pops))) return bail(rc);
    I.pops = pops;
    // projection streams
    std::vector<ScProj> projs = P.projs;
    for (size_t pi = 0; pi < projs.size(); ++pi) {
        ScProj &sp = projs[pi];
        if (sp.mode < 0) continue;
        HostProjArrays &A = P.parr[pi];
        uint32_t *col;
        if ((rc = dupload(e, &col, A.col))) return bail(rc);
        sp.col = col;
        if (!A.val.empty()) {
            uint8_t *val;
            if ((rc = dupload(e, &val, A.val))) return bail(rc);
            sp.val = val;
        }
        if (!A.row_ptr.empty()) {
            int64_t *rp;
            if ((rc = dupload(e, &rp, A.row_ptr))) return bail(rc);
            sp.row_ptr = rp;
        }
        if (!A.sell_off.empty()) {
            int64_t *so;
            int32_t *sw;
            if ((rc = dupload(e, &so, A.sell_off))) return bail(rc);
            if ((rc = dupload(e, &sw, A.sell_w))) return bail(rc);
            sp.sell_off = so;
            sp.sell_w = sw;
        }
        // the column and weight streams now live on the device only
        std::vector<uint32_t>().swap(A.col);
        std::vector<uint8_t>().swap(A.val);
    }
    ScProj *dprojs;
    if ((rc = dupload(e, &dprojs, projs))) return bail(rc);
    I.projs = dprojs;
    if ((rc = dupload(e, &e->stims, P.stims))) return bail(rc);
    I.stims = e->stims;
    e->n_stims = (int)P.stims.size();
    int32_t *ss, *se;
    if ((rc = dupload(e, &ss, P.stim_s))) return bail(rc);
    if ((rc = dupload(e, &se, P.stim_e))) return bail(rc);
    I.stim_s = ss;
    I.stim_e = se;
    uint32_t *od;
    if ((rc = dupload(e, &od, P.outdeg))) return bail(rc);
    I.outdeg = od;
    ScCtl *ctl;
    if ((rc = dalloc(e, &ctl, 1))) return bail(rc);
    I.ctl = ctl;
    if ((rc = dupload(e, &e->tasks, P.tasks))) return bail(rc);
    e->n_tasks = (int)P.tasks.size();
    std::vector<int32_t> fin;
    for (int g : P.fin_groups) { fin.push_back(0); fin.push_back(g); }
    if ((rc = dupload(e, &e->fin, fin))) return bail(rc);
    e->n_fin = (int)P.fin_groups.size();
    if ((rc = dalloc(e, &e->cnt_buf, 1))) return bail(rc);
    std::vector<ScInstance> one{I};
    if ((rc = dupload(e, &e->di, one))) return bail(rc);
    *out = e;
    return SCAFFOLD_OK;
}

void engine_destroy(Engine *e)
{
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    for (void *p : e->allocs) (void)hipFree(p);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

int64_t engine_device_bytes(Engine *e) { return e ? e->bytes : 0; }

int engine_sync(Engine *e)
{
    HIP_TRY(hipStreamSynchronize(e->stream));
    return SCAFFOLD_OK;
}

int engine_init_params(Engine *e, uint64_t seed, int64_t call_index)
{
    const int n = std::max(e->Ndev, e->hi.Wr);
    k_init_params<<<(n + 255) / 256, 256, 0, e->stream>>>(e->di, (uint32_t)seed, (uint32_t)(seed >> 32),
                                                           (uint64_t)call_index);
    HIP_TRY(hipGetLastError());
    return engine_sync(e);
}

int engine_zero_g(Engine *e)
{
    HIP_TRY(hipMemsetAsync(e->hi.g, 0, sizeof(float) * 2 * (size_t)e->Ndev, e->stream));
    return engine_sync(e);
}

int engine_set_ctl(Engine *e, int64_t n, int64_t m)
{
    k_set_ctl<<<1, 64, 0, e->stream>>>(e->hi.ctl, n, m);
    HIP_TRY(hipGetLastError());
    return engine_sync(e);
}

int engine_clear_counts_row(Engine *e, int row)
{
    HIP_TRY(hipMemsetAsync(e->hi.counts + (size_t)row * e->n_pops, 0, sizeof(int32_t) * e->n_pops, e->stream));
    HIP_TRY(hipMemsetAsync(e->hi.events + row, 0, sizeof(unsigned long long), e->stream));
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

static inline int blocks_for(int waves) { return std::max(1, (waves + 3) / 4); }

static void launch_step(hipStream_t s, const ScTask *tasks, int n_tasks, const int32_t *fin, int n_fin,
                        const ScInstance *insts, int n_inst, int k)
{
    if (n_tasks > 0) k_gather<<<blocks_for(n_tasks), 256, 0, s>>>(tasks, n_tasks, insts, k);
    k_finalize<<<blocks_for(n_fin), 256, 0, s>>>(fin, n_fin, insts, n_inst, k);
}

static int run_steps(hipStream_t s, hipGraphExec_t *gexec, int *gexec_steps, const ScTask *tasks, int n_tasks,
                     const int32_t *fin, int n_fin, const ScInstance *insts, int n_inst, int n_steps, int use_graph,
                     int graph_steps)
{
    int done = 0;
    if (use_graph && n_steps >= graph_steps) {
        if (!*gexec || *gexec_steps != graph_steps) {
            if (*gexec) {
                (void)hipGraphExecDestroy(*gexec);
                *gexec = nullptr;
            }
            hipGraph_t g;
            HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < graph_steps; ++k) launch_step(s, tasks, n_tasks, fin, n_fin, insts, n_inst, k);
            k_advance<<<(n_inst + 63) / 64, 64, 0, s>>>(insts, n_inst, graph_steps);
            HIP_TRY(hipStreamEndCapture(s, &g));
            hipError_t err = hipGraphInstantiate(gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (err != hipSuccess) return fail(SCAFFOLD_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(err));
            *gexec_steps = graph_steps;
        }
        while (n_steps - done >= graph_steps) {
            HIP_TRY(hipGraphLaunch(*gexec, s));
            done += graph_steps;
        }
    }
    const int rest = n_steps - done;
    if (rest > 0) {
        for (int k = 0; k < rest; ++k) launch_step(s, tasks, n_tasks, fin, n_fin, insts, n_inst, k);
        k_advance<<<(n_inst + 63) / 64, 64, 0, s>>>(insts, n_inst, rest);
        HIP_TRY(hipGetLastError());
    }
    return SCAFFOLD_OK;
}

int engine_run(Engine *e, int n_steps, int use_graph, int graph_steps)
{
    int rc = run_steps(e->stream, &e->gexec, &e->gexec_steps, e->tasks, e->n_tasks, e->fin, e->n_fin, e->di, 1,
                       n_steps, use_graph, graph_steps);
    if (rc) return rc;
    return engine_sync(e);
}

int engine_time(Engine *e, int n_steps, int use_graph, int graph_steps, double *ms_total, double *ms_kernel_mean)
{
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    if (!ms_kernel_mean) {
        HIP_TRY(hipEventRecord(a, e->stream));
        int rc = run_steps(e->stream, &e->gexec, &e->gexec_steps, e->tasks, e->n_tasks, e->fin, e->n_fin, e->di, 1,
                           n_steps, use_graph, graph_steps);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(b, e->stream));
        HIP_TRY(hipEventSynchronize(b));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, a, b));
        if (ms_total) *ms_total = ms;
    } else {
        // direct launches with events around every gather kernel
        std::vector<hipEvent_t> ev(2 * (size_t)n_steps);
        for (auto &x : ev) HIP_TRY(hipEventCreate(&x));
        HIP_TRY(hipEventRecord(a, e->stream));
        for (int k = 0; k < n_steps; ++k) {
            HIP_TRY(hipEventRecord(ev[2 * k], e->stream));
            if (e->n_tasks > 0) k_gather<<<blocks_for(e->n_tasks), 256, 0, e->stream>>>(e->tasks, e->n_tasks, e->di, k);
            HIP_TRY(hipEventRecord(ev[2 * k + 1], e->stream));
            k_finalize<<<blocks_for(e->n_fin), 256, 0, e->stream>>>(e->fin, e->n_fin, e->di, 1, k);
        }
        k_advance<<<1, 64, 0, e->stream>>>(e->di, 1, n_steps);
        HIP_TRY(hipEventRecord(b, e->stream));
        HIP_TRY(hipEventSynchronize(b));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, a, b));
        if (ms_total) *ms_total = ms;
        double sum = 0.0;
        for (int k = 0; k < n_steps; ++k) {
            float x = 0.f;
            HIP_TRY(hipEventElapsedTime(&x, ev[2 * k], ev[2 * k + 1]));
            sum += x;
        }
        *ms_kernel_mean = sum / n_steps;
        for (auto &x : ev) (void)hipEventDestroy(x);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return SCAFFOLD_OK;
}

int engine_get_state(Engine *e, float *V, float *g, int32_t *refr)
{
    HIP_TRY(hipStreamSynchronize(e->stream));
    const size_t N = (size_t)e->Ndev;
    std::vector<float> hv(N), hg(2 * N);
    std::vector<int32_t> hr(N);
    HIP_TRY(hipMemcpy(hv.data(), e->hi.V, sizeof(float) * N, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hg.data(), e->hi.g, sizeof(float) * 2 * N, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hr.data(), e->hi.refr, sizeof(int32_t) * N, hipMemcpyDeviceToHost));
    for (size_t d = 0; d < N; ++d) {
        const int r = e->dev_to_ref[d];
        if (r < 0) continue;
        if (V) V[r] = hv[d];
        if (g) { g[2 * (size_t)r] = hg[2 * d]; g[2 * (size_t)r + 1] = hg[2 * d + 1]; }
        if (refr) refr[r] = hr[d];
    }
    return SCAFFOLD_OK;
}

int engine_set_state(Engine *e, const float *V, const float *g, const int32_t *refr)
{
    HIP_TRY(hipStreamSynchronize(e->stream));
    const size_t N = (size_t)e->Ndev;
    std::vector<float> hv(N, 0.f), hg(2 * N, 0.f);
    std::vector<int32_t> hr(N, 0);
    HIP_TRY(hipMemcpy(hv.data(), e->hi.V, sizeof(float) * N, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hg.data(), e->hi.g, sizeof(float) * 2 * N, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hr.data(), e->hi.refr, sizeof(int32_t) * N, hipMemcpyDeviceToHost));
    for (size_t d = 0; d < N; ++d) {
        const int r = e->dev_to_ref[d];
        if (r < 0) continue;
        if (V) hv[d] = V[r];
        if (g) { hg[2 * d] = g[2 * (size_t)r]; hg[2 * d + 1] = g[2 * (size_t)r + 1]; }
        if (refr) hr[d] = refr[r];
    }
    HIP_TRY(hipMemcpy(e->hi.V, hv.data(), sizeof(float) * N, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(e->hi.g, hg.data(), sizeof(float) * 2 * N, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(e->hi.refr, hr.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice));
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

int engine_get_fired(Engine *e, int row, int32_t *ids, int32_t cap, int32_t *n)
{
    int32_t *dout = nullptr;
    const int c = std::max(cap, 1);
    HIP_TRY(hipMalloc(&dout, sizeof(int32_t) * c));
    k_compact<<<1, 1024, 0, e->stream>>>(e->hi.ring + (size_t)row * e->hi.Wr, e->hi.n_groups, e->hi.groups,
                                         e->hi.pops, dout, cap, e->cnt_buf);
    hipError_t err = hipGetLastError();
    if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
    if (err == hipSuccess) err = hipMemcpy(n, e->cnt_buf, sizeof(int32_t), hipMemcpyDeviceToHost);
    if (err == hipSuccess && cap > 0)
        err = hipMemcpy(ids, dout, sizeof(int32_t) * std::min(cap, *n), hipMemcpyDeviceToHost);
    (void)hipFree(dout);
    if (err != hipSuccess) return fail(SCAFFOLD_ERR_HIP, "fired-list compaction: %s", hipGetErrorString(err));
    return SCAFFOLD_OK;
}

int engine_copy_probe(int64_t bytes, int iters, double *ms)
{
    int rc = check_device();
    if (rc) return rc;
    const int64_t n = std::max<int64_t>(1, bytes / 16);
    uint4 *a = nullptr, *b = nullptr;
    HIP_TRY(hipMalloc(&a, n * 16));
    HIP_TRY(hipMalloc(&b, n * 16));
    HIP_TRY(hipMemset(a, 1, n * 16));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 256 * 64);
    k_copy<<<grid, 256>>>(a, b, n);
    HIP_TRY(hipEventRecord(e0));
    for (int i = 0; i < iters; ++i) k_copy<<<grid, 256>>>(a, b, n);
    HIP_TRY(hipEventRecord(e1));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    *ms = t / iters;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return SCAFFOLD_OK;
}


// ---------------------------------------------------------------------- batching
// Independent instances of uneven size advance together: one work list spanning
// all instances (each task names its instance), one step = the same two kernels.
struct Batch {
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    ScInstance *di = nullptr;
    int n_inst = 0;
    ScTask *tasks = nullptr;
    int n_tasks = 0;
    int32_t *fin = nullptr;
    int n_fin = 0;
    hipGraphExec_t gexec = nullptr;
    int gexec_steps = 0;
};

void batch_destroy(Batch *b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    for (void *p : b->allocs) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int batch_build(Batch **out, std::vector<Engine *> &engines, const std::vector<const Plan *> &plans)
{
    *out = nullptr;
    Batch *b = new Batch();
    auto bail = [&](int code) { batch_destroy(b); return code; };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(SCAFFOLD_ERR_HIP, "hipStreamCreate failed"));
    std::vector<ScInstance> insts;
    std::vector<ScTask> tasks;
    std::vector<int32_t> fin;
    for (size_t i = 0; i < engines.size(); ++i) insts.push_back(engines[i]->hi);
    // heavier kinds first across all instances: row chunks, lane groups, stimulus groups
    for (int kind : {SC_TASK_ROW, SC_TASK_LANE, SC_TASK_STIM})
        for (size_t i = 0; i < plans.size(); ++i)
            for (const ScTask &t : plans[i]->tasks)
                if (t.kind == kind) {
                    ScTask u = t;
                    u.inst = (int32_t)i;
                    tasks.push_back(u);
                }
    for (size_t i = 0; i < plans.size(); ++i)
        for (int g : plans[i]->fin_groups) { fin.push_back((int32_t)i); fin.push_back(g); }
    auto up = [&](auto **dst, const auto &v) -> int {
        using T = typename std::remove_reference<decltype(v)>::type::value_type;
        void *q = nullptr;
        const size_t bytes = sizeof(T) * std::max<size_t>(v.size(), 1);
        HIP_TRY(hipMalloc(&q, bytes));
        b->allocs.push_back(q);
        if (!v.empty()) HIP_TRY(hipMemcpy(q, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
        *dst = static_cast<T *>(q);
        return SCAFFOLD_OK;
    };
    int rc;
    if ((rc = up(&b->di, insts))) return bail(rc);
    if ((rc = up(&b->tasks, tasks))) return bail(rc);
    if ((rc = up(&b->fin, fin))) return bail(rc);
    b->n_inst = (int)insts.size();
    b->n_tasks = (int)tasks.size();
    b->n_fin = (int)fin.size() / 2;
    *out = b;
    return SCAFFOLD_OK;
}

int batch_run(Batch *b, int n_steps, int graph_steps, double *ms)
{
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, b->stream));
    int rc = run_steps(b->stream, &b->gexec, &b->gexec_steps, b->tasks, b->n_tasks, b->fin, b->n_fin, b->di,
                       b->n_inst, n_steps, 1, graph_steps);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(e1, b->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return SCAFFOLD_OK;
}

}  // namespace sc
