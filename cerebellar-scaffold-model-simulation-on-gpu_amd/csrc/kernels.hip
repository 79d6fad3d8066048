// kernels.hip -- gfx950 kernels of the exchange-step hot path and their launchers.
//
// One exchange step (1 ms = 10 Euler sub-steps of 0.1 ms) is one launch of k_step
// (1024-thread blocks, each holding tasks of one kind):
//   LANE  a wavefront per 64 post neurons of a lane-mode population (short rows):
//         one lane per neuron over the sliced-ELL slices of its projections, LIF
//         update, spike word = wavefront ballot, plain store;
//   TEAM  1..16 wavefronts per post neuron of a team-mode population (long rows):
//         the team strides over the neuron's merged CSR row and accumulates the
//         exact f64 sum of every incoming projection in LDS; one lane (or, for
//         blocks of one-wavefront teams, the first wavefront for all of them) runs
//         the LIF update and sets the neuron's bit in the spike word with an
//         integer atomicOr;
//   STIM  a wavefront per 64 INPUT neurons: counter-based Philox4x32-10 draws and
//         the stimulus rule, spike word = ballot;
//   HIST  (with plasticity rules) a wavefront per 64 presynaptic neurons of a
//         rule: spike counts over the rule's window, for k_plastic, which runs
//         after the step and updates the rule's multipliers.
// A step reads ring rows of earlier steps only (every delay >= 1) and writes the
// row of the current step, so blocks never wait on each other; the kernel
// boundary orders consecutive steps.  k_run_end closes a run of steps: it
// popcounts the run's spike words into per-population counts, sums the
// out-degrees of the spiking neurons (synaptic events), gathers the output spike
// train (unless k_out_train, launched just after the last step, already has: a
// directly launched call then returns on an event after the train, before the
// statistics) and moves the device step counters.
//
// Blocks whose gathers read a bounded range of presynaptic neurons first copy
// those spike words ("windows", planned per block by the engine) into LDS with
// direct global-to-LDS loads, so their spike-word reads are LDS reads instead of
// scattered memory gathers; the copies overlap the block's first descriptor,
// state and column loads.
//
// Latency: at the reference size the work per step is tens of MB, so the step
// time is set by chains of dependent memory accesses and by instruction fetch;
// at larger sizes by the number of load instructions and scattered cache lines.
// Every task is a descriptor with absolute pointers (one dependent load to reach
// the data); columns and multipliers come in 16- or 8-byte vectors (4 or 8
// entries per lane per load); the gathers issue a whole batch of column,
// multiplier and spike-word loads before consuming any of them; and the code a
// block runs is kept small (loops that are serial anyway -- Euler sub-steps,
// Philox rounds -- are not unrolled), because a block executes its path once per
// step and a long straight-line path stalls on instruction fetch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>

#include "kernels.hpp"

namespace sc {

// ------------------------------------------------------------------ device math
// Philox4x32-10 round (Salmon et al., SC'11) on counter (c0,c1,c2,c3) with key (k0,k1).
__device__ __forceinline__ void philox_round(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3, uint32_t k0,
                                             uint32_t k1)
{
    // one 32 x 32 -> 64-bit product each (v_mad_u64_u32: both halves in one instruction)
    const uint64_t p0 = (uint64_t)c0 * 0xD2511F53u, p1 = (uint64_t)c2 * 0xCD9E8D57u;
    const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
}

__device__ __forceinline__ void philox10(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3, uint32_t k0,
                                         uint32_t k1)
{
#pragma unroll 1
    for (int r = 0; r < 10; ++r) {
        philox_round(c0, c1, c2, c3, k0, k1);
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

// Pointers into device memory reached through descriptors are generic pointers
// to the compiler; casting them to the global address space makes it emit
// global loads (vector memory only) instead of flat loads, which would also hold
// the LDS counter and serialise LDS waits behind memory loads.
template <typename T>
using gptr = const __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ gptr<T> gbl(const T *p)
{
    return (gptr<T>)p;
}

// A load of the synapse streams (columns, headers, multipliers).  NT marks it
// nontemporal: on networks whose streams exceed the L2, the streamed lines then
// do not evict the descriptors and spike rows the latency chains wait on.
template <bool NT, typename T>
__device__ __forceinline__ T ldg(gptr<T> p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

// 4 and 2 dwords loaded by one 16- / 8-byte load.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// State record of device neuron d: one 16-byte load, one 16-byte store.
__device__ __forceinline__ void load_state(const ScState *st, int d, float &u, float2 &g, int32_t &r)
{
    const u32x4 x = gbl(reinterpret_cast<const u32x4 *>(st))[d];
    u = __uint_as_float(x.x);
    g = make_float2(__uint_as_float(x.y), __uint_as_float(x.z));
    r = (int32_t)x.w;
}
__device__ __forceinline__ void store_state(ScState *st, int d, float u, float2 g, int32_t r)
{
    u32x4 x;
    x.x = __float_as_uint(u);
    x.y = __float_as_uint(g.x);
    x.z = __float_as_uint(g.y);
    x.w = (uint32_t)r;
    reinterpret_cast<u32x4 *>(st)[d] = x;
}

// One direct global-to-LDS load: lane l copies the 16 bytes at src (its own
// address) to LDS at dst + 16 * l, without passing through registers; dst is
// wave-uniform.  The copy counts as an outstanding vector-memory load of the
// wavefront until it has been written to LDS.
__device__ __forceinline__ void lds_load16(const u32x4 *src, u32x4 *dst)
{
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)src,
                                     (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
}

// Wave-uniform value in a scalar register (descriptor indices derived from the
// wavefront index), so descriptor fields are fetched with scalar loads and
// branches on them are scalar branches.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
template <typename T>
__device__ __forceinline__ T *uniform_ptr(T *p)
{
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return reinterpret_cast<T *>(((uint64_t)hi << 32) | lo);
}

// Multipliers are loaded as raw bits, 4 or 8 entries per load, and decoded only
// after a batch of loads has arrived.
// Raw bits of the 4 multipliers of vector v (entries 4v .. 4v+3) in storage WP,
// one load: 16 bytes (F32), 8 (BF16) or 4 (FP8).
template <int WP>
struct Raw4 {
    uint32_t w[WP == SC_W_F32 ? 4 : WP == SC_W_BF16 ? 2 : 1];
};
template <int WP, bool NT>
__device__ __forceinline__ Raw4<WP> wraw4(const void *val, int64_t v)
{
    Raw4<WP> r;
    if constexpr (WP == SC_W_F32) {
        const u32x4 x = ldg<NT>(gbl(static_cast<const u32x4 *>(val)) + v);
        r.w[0] = x.x; r.w[1] = x.y; r.w[2] = x.z; r.w[3] = x.w;
    } else if constexpr (WP == SC_W_BF16) {
        const u32x2 x = ldg<NT>(gbl(static_cast<const u32x2 *>(val)) + v);
        r.w[0] = x.x; r.w[1] = x.y;
    } else if constexpr (WP == SC_W_FP8) {
        r.w[0] = ldg<NT>(gbl(static_cast<const uint32_t *>(val)) + v);
    } else {
        r.w[0] = 0u;
    }
    return r;
}
// Raw bits of element i of a Raw4.
template <int WP>
__device__ __forceinline__ uint32_t wsel4(const Raw4<WP> &r, int i)
{
    if constexpr (WP == SC_W_F32) return r.w[i];
    else if constexpr (WP == SC_W_BF16) return (r.w[i >> 1] >> (16 * (i & 1))) & 0xFFFFu;
    else if constexpr (WP == SC_W_FP8) return (r.w[0] >> (8 * i)) & 0xFFu;
    else return 0u;
}

// Value of element i (a constant after unrolling) of a Raw4.
template <int WP>
__device__ __forceinline__ float wval4(const Raw4<WP> &r, int i);

// Multiplier value of raw bits in storage WP (every conversion to float is exact).
template <int WP>
__device__ __forceinline__ float wdec(uint32_t c)
{
    if constexpr (WP == SC_W_F32) {
        return __uint_as_float(c);
    } else if constexpr (WP == SC_W_BF16) {
        return __uint_as_float(c << 16);
    } else if constexpr (WP == SC_W_FP8) {
        // OCP e4m3 ((1 + m/8) * 2^(e-7), subnormal m * 2^-9) by gfx950's conversion
        // instruction, which decodes every code the planner stores (no NaN) exactly
        return __builtin_amdgcn_cvt_f32_fp8((int)c, 0);
    } else {
        return 1.f;
    }
}

template <int WP>
__device__ __forceinline__ float wval4(const Raw4<WP> &r, int i)
{
    if constexpr (WP == SC_W_FP8) {
        switch (i) {  // the byte select of the conversion is an immediate
        case 0: return __builtin_amdgcn_cvt_f32_fp8((int)r.w[0], 0);
        case 1: return __builtin_amdgcn_cvt_f32_fp8((int)r.w[0], 1);
        case 2: return __builtin_amdgcn_cvt_f32_fp8((int)r.w[0], 2);
        default: return __builtin_amdgcn_cvt_f32_fp8((int)r.w[0], 3);
        }
    } else {
        return wdec<WP>(wsel4<WP>(r, i));
    }
}

// Raw bits of the 8 multipliers of chunk v (entries 8v .. 8v+7): 32 bytes (F32,
// two loads), 16 (BF16) or 8 (FP8).
template <int WP>
struct Raw8 {
    uint32_t w[WP == SC_W_F32 ? 8 : WP == SC_W_BF16 ? 4 : 2];
};
template <int WP, bool NT>
__device__ __forceinline__ Raw8<WP> wraw8(const void *val, int64_t v)
{
    Raw8<WP> r;
    if constexpr (WP == SC_W_F32) {
        const u32x4 x = ldg<NT>(gbl(static_cast<const u32x4 *>(val)) + 2 * v);
        const u32x4 y = ldg<NT>(gbl(static_cast<const u32x4 *>(val)) + 2 * v + 1);
        r.w[0] = x.x; r.w[1] = x.y; r.w[2] = x.z; r.w[3] = x.w;
        r.w[4] = y.x; r.w[5] = y.y; r.w[6] = y.z; r.w[7] = y.w;
    } else if constexpr (WP == SC_W_BF16) {
        const u32x4 x = ldg<NT>(gbl(static_cast<const u32x4 *>(val)) + v);
        r.w[0] = x.x; r.w[1] = x.y; r.w[2] = x.z; r.w[3] = x.w;
    } else if constexpr (WP == SC_W_FP8) {
        const u32x2 x = ldg<NT>(gbl(static_cast<const u32x2 *>(val)) + v);
        r.w[0] = x.x; r.w[1] = x.y;
    } else {
        r.w[0] = r.w[1] = 0u;
    }
    return r;
}
// Value of element i (a constant after unrolling) of a Raw8.
template <int WP>
__device__ __forceinline__ float wval8(const Raw8<WP> &r, int i)
{
    if constexpr (WP == SC_W_FP8) {
        const int x = (int)r.w[i >> 2];
        switch (i & 3) {  // the byte select of the conversion is an immediate
        case 0: return __builtin_amdgcn_cvt_f32_fp8(x, 0);
        case 1: return __builtin_amdgcn_cvt_f32_fp8(x, 1);
        case 2: return __builtin_amdgcn_cvt_f32_fp8(x, 2);
        default: return __builtin_amdgcn_cvt_f32_fp8(x, 3);
        }
    } else if constexpr (WP == SC_W_BF16) {
        return __uint_as_float(((r.w[i >> 1] >> (16 * (i & 1))) & 0xFFFFu) << 16);
    } else if constexpr (WP == SC_W_F32) {
        return __uint_as_float(r.w[i]);
    } else {
        return 1.f;
    }
}

// Ring row of launch offset k of a run (the run's first row ctl->row < R, k < R).
__device__ __forceinline__ int run_row(const ScInstance &I, int k)
{
    const int r = I.ctl->row + k;
    return r >= I.R ? r - I.R : r;
}

// Ring row of the step `back` steps before the step whose row is `row` (back < R).
__device__ __forceinline__ int row_back(int row, int back, int R)
{
    const int r = row - back;
    return r < 0 ? r + R : r;
}

// One exchange step of the conductance LIF neuron, update_lif_ch
// (Simulation_header.cu:315-377), every expression promoted as the reference
// writes it; the library is compiled with -ffp-contract=off so no product is
// fused into a following sum.  The sub-step loop is a serial dependence chain, so
// it is kept rolled.
// The leading words of ScPop, up to and including ref_step: all the neuron update
// reads.  Team blocks copy them into LDS when a task starts.
#define SC_POP_WORDS ((int)(offsetof(ScPop, ref_step) / 4 + 1))

// One Euler sub-step with CN channels; INJ adds the step's synaptic input (the
// first sub-step only).
template <int CN, bool INJ>
__device__ __forceinline__ int lif_substep(float &u, float &g0, float &g1, int32_t &r, double dtc, double dtt0,
                                           double dtt1, float inj0, float inj1, float E0, float E1, float El, float gL,
                                           float Ie, float Vth, float Vr, int ref_step)
{
    float Isyn = 0.f;
    if (CN > 0) Isyn = Isyn + g0 * (u - E0);
    if (CN > 1) Isyn = Isyn + g1 * (u - E1);
    const float leak = gL * (u - El);
    const float inner = (Ie - leak) - Isyn;
    const float du = (float)(dtc * (double)inner);
    if (CN > 0) {
        const double dec = (-dtt0) * (double)g0;
        g0 = (float)((double)g0 + (INJ ? dec + (double)inj0 : dec));
    }
    if (CN > 1) {
        const double dec = (-dtt1) * (double)g1;
        g1 = (float)((double)g1 + (INJ ? dec + (double)inj1 : dec));
    }
    const int fired = u > Vth;
    r = fired ? ref_step : (r > 0 ? r - 1 : 0);
    u = r ? Vr : u + du;
    return fired;
}

template <int CN>
__device__ __forceinline__ int lif_steps(float &u, float &g0, float &g1, int32_t &r, double dtc, double dtt0,
                                         double dtt1, float inj0, float inj1, float E0, float E1, float El, float gL,
                                         float Ie, float Vth, float Vr, int ref_step)
{
    int spk = lif_substep<CN, true>(u, g0, g1, r, dtc, dtt0, dtt1, inj0, inj1, E0, E1, El, gL, Ie, Vth, Vr, ref_step);
#pragma unroll 1
    for (int k = 1; k < kStepMax; ++k)
        spk |= lif_substep<CN, false>(u, g0, g1, r, dtc, dtt0, dtt1, inj0, inj1, E0, E1, El, gL, Ie, Vth, Vr,
                                      ref_step);
    return spk;
}

template <typename PopT>
__device__ __forceinline__ int lif_step(const PopT &P, float &u, float &g0, float &g1, int32_t &r, float dge,
                                        float dgi)
{
    const int cn = P.c_num;
    const float inj0 = P.gw[0] * (P.exc_ch[0] ? dge : dgi);
    const float inj1 = P.gw[1] * (P.exc_ch[1] ? dge : dgi);
    const double dtc = P.dtc, dtt0 = P.dtt[0], dtt1 = P.dtt[1];
    const float E0 = P.E[0], E1 = P.E[1], El = P.El, gL = P.gL, Ie = P.Ie, Vth = P.Vth, Vr = P.Vr;
    const int ref_step = P.ref_step;
    if (cn == 2) return lif_steps<2>(u, g0, g1, r, dtc, dtt0, dtt1, inj0, inj1, E0, E1, El, gL, Ie, Vth, Vr, ref_step);
    if (cn == 1) return lif_steps<1>(u, g0, g1, r, dtc, dtt0, dtt1, inj0, inj1, E0, E1, El, gL, Ie, Vth, Vr, ref_step);
    return lif_steps<0>(u, g0, g1, r, dtc, dtt0, dtt1, inj0, inj1, E0, E1, El, gL, Ie, Vth, Vr, ref_step);
}

// The spike word holding device bit c of a lane slice: from the ring row in
// global memory, or in a staged block from its LDS window (lds: the LDS word that
// stands for word 0 of the row); padding (bit zb) reads LDS word 0, kept zero.
template <bool STAGED>
__device__ __forceinline__ uint32_t spike_word(gptr<uint32_t> row32, const uint32_t *win, int lds, uint32_t c, uint32_t zb)
{
    if constexpr (STAGED) return win[c == zb ? 0 : lds + (int)(c >> 5)];
    else return row32[c >> 5];
}

// 16-bit slices: lane `lane`'s offsets of columns 2p and 2p+1 share the 32-bit
// word p*64 + lane (low half: column 2p), relative to the pair's header (a
// wave-uniform base device bit); SC_PAD16 marks padding, which reads bit zb.
__device__ __forceinline__ uint32_t pair_bit(uint32_t word, int hi, uint32_t base, uint32_t zb)
{
    const uint32_t o = hi ? word >> 16 : word & 0xFFFFu;
    return o == SC_PAD16 ? zb : base + o;
}

// Entries k >= k0 of one sliced-ELL row (slices wider than the first vector of
// 4 columns), 8 columns a batch: two vectors of columns (32-bit: one 16-byte
// load each; 16-bit: four column pairs) and two vectors of multipliers.  Vector
// indices past the slice are clamped to its last vector and the entries past
// its width masked out of the sum.  k0 is a multiple of 4.
template <int WP, bool IDX16, bool STAGED, bool NT>
__device__ __forceinline__ double slice_tail(const ScSlice &sl, gptr<uint32_t> row32, const uint32_t *win, uint32_t zb,
                                             int k0, int lane)
{
    double S = 0.0;
    const int width = sl.width, nv = (width + 3) >> 2;
    const bool unit = WP == SC_W_UNIT || !sl.val;
#pragma unroll 1
    for (int k = k0; k < width; k += 8) {
        uint32_t c[8], w[8];
        Raw4<WP> raw[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int v = min((k >> 2) + h, nv - 1);
            if (!unit) raw[h] = wraw4<WP, NT>(sl.val, (int64_t)v * SC_WAVE + lane);
            else raw[h].w[0] = 0u;
        }
        if constexpr (IDX16) {
            // pairs 2g, 2g+1 of a lane in one 8-byte load; pair headers (stored for
            // the even-padded pair count) in one load, read back with readlane
            gptr<u32x2> col = gbl(static_cast<const u32x2 *>(sl.col));
            const uint32_t hv = ldg<NT>(gbl(sl.hdr) + min((k >> 1) + min(lane, 3), 2 * nv - 1));
            uint32_t pw[4];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const u32x2 x = ldg<NT>(col + (min((k >> 2) + h, nv - 1) * SC_WAVE + lane));
                pw[2 * h] = x.x;
                pw[2 * h + 1] = x.y;
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) c[j] = pair_bit(pw[j >> 1], j & 1, __builtin_amdgcn_readlane(hv, j >> 1), zb);
        } else {
            gptr<u32x4> col = gbl(static_cast<const u32x4 *>(sl.col));
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const u32x4 x = ldg<NT>(col + (min((k >> 2) + h, nv - 1) * SC_WAVE + lane));
                c[4 * h] = x.x; c[4 * h + 1] = x.y; c[4 * h + 2] = x.z; c[4 * h + 3] = x.w;
            }
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) w[j] = spike_word<STAGED>(row32, win, sl.lds, c[j], zb);
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (k + j < width && ((w[j] >> (c[j] & 31)) & 1u))
                S += unit ? 1.0 : (double)wval4<WP>(raw[j >> 2], j & 3);
    }
    return S;
}

// One 64-neuron group of a lane-mode population: the first 4 columns of every
// slice are loaded in one batch (one vector of columns and one of multipliers per
// slice, then all spike words), the neuron state alongside, then the LIF update.
// In a staged block the window copies are issued (stage()) once the first loads
// are under way.
// Slice widths and pointers are wave-uniform; slices without multipliers (unit
// weights) skip the multiplier loads.
template <int WP, bool IDX16, bool STAGED, bool NT, typename Stage>
__device__ __forceinline__ void lane_group(const ScLaneTask &t, const ScInstance &I, const ScPop &P, int row_n,
                                           const Stage &stage, int lane)
{
    const uint32_t zb = (uint32_t)(I.Wr - 1) * SC_WAVE;
    const int d = t.group * SC_WAVE + lane;
    const bool valid = lane < t.nvalid;
    float u = 0.f;
    float2 gg = {0.f, 0.f};
    int32_t rf = 0;
    if (valid) load_state(I.st, d, u, gg, rf);
    uint32_t c[SC_MAX_SLICES][4], sw[SC_MAX_SLICES][4];
    Raw4<WP> raw[SC_MAX_SLICES];
    uint32_t pw[SC_MAX_SLICES][2];  // 16-bit form: offset pairs
    gptr<uint32_t> row[SC_MAX_SLICES];
    const int n_sl = t.n_sl;
    uint32_t hv = 0u;  // 16-bit form: lane 2s+p holds the header of pair p of slice s
    if constexpr (IDX16) {
        const int hs = lane >> 1, hp = lane & 1;
        const uint32_t *h = nullptr;
        int w = 0;
#pragma unroll
        for (int q = 0; q < SC_MAX_SLICES; ++q)
            if (hs == q && q < n_sl) {
                h = t.sl[q].hdr;
                w = t.sl[q].width;
            }
        if (2 * hp < w) hv = ldg<NT>(gbl(h) + hp);
    }
#pragma unroll
    for (int s = 0; s < SC_MAX_SLICES; ++s) {
        const bool on = s < n_sl;
        const int width = on ? t.sl[s].width : 0;
        // wave-uniform (one instance per task): held in scalar registers
        const int rb = uniform(on ? row_back(row_n, t.sl[s].delay, I.R) : 0);
        row[s] = gbl(reinterpret_cast<const uint32_t *>(I.ring + (int64_t)rb * I.Wr));
        const void *val = t.sl[s].val;
        raw[s].w[0] = 0u;
        if (WP != SC_W_UNIT && val && width > 0) raw[s] = wraw4<WP, NT>(val, lane);
        if constexpr (IDX16) {
            pw[s][0] = pw[s][1] = 0xFFFFFFFFu;
            if (width > 0) {
                const u32x2 x = ldg<NT>(gbl(static_cast<const u32x2 *>(t.sl[s].col)) + lane);
                pw[s][0] = x.x;
                pw[s][1] = x.y;
            }
        } else {
            c[s][0] = c[s][1] = c[s][2] = c[s][3] = zb;
            if (width > 0) {
                const u32x4 x = ldg<NT>(gbl(static_cast<const u32x4 *>(t.sl[s].col)) + lane);
                c[s][0] = x.x; c[s][1] = x.y; c[s][2] = x.z; c[s][3] = x.w;
            }
        }
    }
    // staged block: the window copies are issued now, with the first loads in
    // flight, and are complete past the block's barrier, which every wavefront
    // of the block reaches once
    const uint32_t *win = nullptr;
    if constexpr (STAGED) {
        win = stage();
        __syncthreads();
    }
#pragma unroll
    for (int s = 0; s < SC_MAX_SLICES; ++s)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if constexpr (IDX16)
                c[s][k] = pair_bit(pw[s][k >> 1], k & 1, __builtin_amdgcn_readlane(hv, 2 * s + (k >> 1)), zb);
            sw[s][k] = spike_word<STAGED>(row[s], win, t.sl[s].lds, c[s][k], zb);
        }
    // row sums of the first 4 columns of every slice, then the wider slices' tails
    // (so the head batch's registers are free while a tail runs)
    double S[SC_MAX_SLICES];
#pragma unroll
    for (int s = 0; s < SC_MAX_SLICES; ++s) {
        const bool unit = WP == SC_W_UNIT || !t.sl[s].val;
        const int width = s < n_sl ? t.sl[s].width : 0;
        S[s] = 0.0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (k < width && ((sw[s][k] >> (c[s][k] & 31)) & 1u))
                S[s] += unit ? 1.0 : (double)wval4<WP>(raw[s], k);
    }
    float dge = 0.f, dgi = 0.f;
#pragma unroll
    for (int s = 0; s < SC_MAX_SLICES; ++s) {
        if (s >= n_sl) break;
        if (t.sl[s].width > 4) S[s] += slice_tail<WP, IDX16, STAGED, NT>(t.sl[s], row[s], win, zb, 4, lane);
        // dg[post] += S * w_bar (Simulation_header.cu:98), projections in index order
        const float Sf = (float)(S[s] * (double)t.sl[s].scale);
        const float add = Sf * t.sl[s].W;
        if (t.sl[s].exc) dge = dge + add;
        else dgi = dgi + add;
    }
    int spk = 0;
    if (valid) {
        spk = lif_step(P, u, gg.x, gg.y, rf, dge, dgi);
        store_state(I.st, d, u, gg, rf);
    }
    const unsigned long long word = __ballot(spk != 0);
    if (lane == 0) I.ring[(int64_t)row_n * I.Wr + t.group] = word;
}

// Team gather over a merged row: thread `tid` of a team of TT threads handles
// entries tid, tid+TT, ... in batches of U (all column and multiplier loads of a
// batch, then all spike words).  Indices past the row end are clamped to its last
// entry, so every load of a batch is unconditional and in bounds (the clamped
// entries are masked out when summing).  A column word carries the projection
// index in its top bits; rowoff[q] is the offset (in 32-bit words) of that
// projection's ring row.  Spiking entries add their multiplier to the team's f64
// LDS accumulator of the projection: the sums are exact, so the order of the
// additions cannot change them.
//
// In the 16-bit form the row is a sequence of 8-entry chunks, each with one
// header word (projection slot and base bit) and 8 16-bit offsets from the base;
// padding entries (SC_PAD16) read the ring row's zero word (bit zb) or, staged,
// LDS word 0.
// The wave-uniform fields of a team-mode neuron's descriptor, loaded once when
// the block starts (with scalar loads) and kept in registers across the block's
// barriers.
struct TeamHead {
    const void *col;
    const uint32_t *hdr;
    const void *val;
    const uint32_t *grp;
    int32_t inst, group, bit, pop, n_in, wprec, len, n_grp;
};
__device__ __forceinline__ TeamHead team_head(const ScTeamTask *T)
{
    return TeamHead{T->col, T->hdr, T->val, T->grp, T->inst, T->group, T->bit, T->pop, T->n_in, T->wprec, T->len,
                    T->n_grp};
}

// One multiplier (index i of a row's array) in storage WP.
template <int WP>
__device__ __forceinline__ float wone(const void *val, uint32_t i)
{
    if constexpr (WP == SC_W_F32) return wdec<WP>(gbl(static_cast<const uint32_t *>(val))[i]);
    else if constexpr (WP == SC_W_BF16) return wdec<WP>(gbl(static_cast<const uint16_t *>(val))[i]);
    else if constexpr (WP == SC_W_FP8) return wdec<WP>(gbl(static_cast<const uint8_t *>(val))[i]);
    else return 1.f;
}

// Word groups of a team (thread tid of TT), SC_GROUP_BATCH groups per lane a
// batch: their 16-byte records (indices past the end clamped to the last group,
// so every load is unconditional; the extra copies are skipped below), then each
// group's spike word -- both halves from LDS in a staged block unless its
// projection was left in memory (gmask), each half read only where the mask has
// bits -- then one multiplier per set bit of (spikes & mask), at the group's
// first index plus the number of mask bits below it.  Spiking entries add to the
// team's f64 LDS accumulator of the projection, as in the merged rows.
#define SC_GROUP_BATCH 1
#define SC_GROUP_LOADS 4
template <int WP, bool STAGED, bool NT>
__device__ __forceinline__ void team_groups(const TeamHead &T, gptr<uint32_t> ring32, const uint32_t *win,
                                            const uint32_t *rowoff, uint32_t gmask, int tid, int TT, double *acc)
{
    constexpr int GB = SC_GROUP_BATCH;
    gptr<u32x4> grp = gbl(reinterpret_cast<const u32x4 *>(T.grp));
    const void *val = T.val;
    const int n = T.n_grp;
    for (int b = tid; b < n; b += TT * GB) {
        u32x4 g[GB];
#pragma unroll
        for (int u = 0; u < GB; ++u) g[u] = ldg<NT>(grp + min(b + u * TT, n - 1));
        uint32_t lo[GB], hi[GB];
#pragma unroll
        for (int u = 0; u < GB; ++u) {
            const uint32_t q = g[u].x >> SC_Q_SHIFT, wi = rowoff[q] + 2u * (g[u].x & SC_BIT_MASK);
            lo[u] = hi[u] = 0u;
            if (!STAGED || (gmask != 0u && ((gmask >> q) & 1u))) {
                if (g[u].z) lo[u] = ring32[wi];
                if (g[u].w) hi[u] = ring32[wi + 1];
            } else {
                if (g[u].z) lo[u] = win[wi];
                if (g[u].w) hi[u] = win[wi + 1];
            }
        }
#pragma unroll
        for (int u = 0; u < GB; ++u) {
            if (b + u * TT >= n) break;
            const uint64_t mask = ((uint64_t)g[u].w << 32) | g[u].z;
            uint64_t act = (((uint64_t)hi[u] << 32) | lo[u]) & mask;
            double *a = &acc[g[u].x >> SC_Q_SHIFT];
            // the multipliers of up to SC_GROUP_LOADS spiking neurons are loaded
            // together (a slot without one reloads the group's first multiplier,
            // which exists), then added
            while (act) {
                uint32_t idx[SC_GROUP_LOADS];
                bool on[SC_GROUP_LOADS];
#pragma unroll
                for (int k = 0; k < SC_GROUP_LOADS; ++k) {
                    on[k] = act != 0ull;
                    const int bit = on[k] ? __ffsll((long long)act) - 1 : 0;
                    idx[k] = g[u].y + (on[k] ? (uint32_t)__popcll(mask & ((1ull << bit) - 1ull)) : 0u);
                    act &= act - 1;
                }
                float wv[SC_GROUP_LOADS];
#pragma unroll
                for (int k = 0; k < SC_GROUP_LOADS; ++k) wv[k] = wone<WP>(val, idx[k]);
#pragma unroll
                for (int k = 0; k < SC_GROUP_LOADS; ++k)
                    if (on[k]) atomicAdd(a, (double)wv[k]);
            }
        }
    }
}

template <int WP, int U, bool IDX16, bool STAGED, bool NT>
__device__ __forceinline__ void team_gather(const TeamHead &T, gptr<uint32_t> ring32, const uint32_t *win,
                                            const uint32_t *rowoff, uint32_t gmask, uint32_t zb, int tid, int TT,
                                            double *acc)
{
    const int len = T.len;
    const void *val = T.val;
    if constexpr (IDX16) {
        // 16-bit form: T.len chunks of 8 entries; a lane takes one chunk a batch
        // (one 16-byte load of offsets, its header word and its multipliers)
        gptr<u32x4> col = gbl(static_cast<const u32x4 *>(T.col));
        gptr<uint32_t> hdr = gbl(T.hdr);
        for (int b = tid; b < len; b += TT) {
            const u32x4 x = ldg<NT>(col + b);
            const uint32_t h = ldg<NT>(hdr + b);
            const Raw8<WP> raw = wraw8<WP, NT>(val, b);
            const uint32_t q = h >> SC_Q_SHIFT, base = h & SC_BIT_MASK, ro = rowoff[q];
            uint32_t bit[8], s[8];
            bool pad[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const uint32_t o = (x[i >> 1] >> (16 * (i & 1))) & 0xFFFFu;
                pad[i] = o == SC_PAD16;
                bit[i] = pad[i] ? zb : base + o;  // padding reads the zero word
            }
            // a staged block reads its chunk's words from the LDS window, unless the
            // chunk's projection was left in memory (gmask; one projection per chunk)
            if (!STAGED || (gmask != 0u && ((gmask >> q) & 1u))) {
#pragma unroll
                for (int i = 0; i < 8; ++i) s[i] = ring32[ro + (bit[i] >> 5)];
            } else {
#pragma unroll
                for (int i = 0; i < 8; ++i) s[i] = win[pad[i] ? 0u : ro + (bit[i] >> 5)];
            }
#pragma unroll
            for (int i = 0; i < 8; ++i)
                if ((s[i] >> (bit[i] & 31)) & 1u) atomicAdd(&acc[h >> SC_Q_SHIFT], (double)wval8<WP>(raw, i));
        }
    } else {
        // 32-bit form: rows padded to whole 16-byte vectors; each lane loads 4
        // consecutive entries (columns and multipliers) per load
        constexpr int V = U / 4;  // vectors per batch
        gptr<u32x4> col = gbl(static_cast<const u32x4 *>(T.col));
        const int nv = (len + 3) >> 2;
        for (int b = tid; b < nv; b += TT * V) {
            uint32_t c[V][4], s[V][4];
            Raw4<WP> raw[V];
#pragma unroll
            for (int u = 0; u < V; ++u) {
                const int v = min(b + u * TT, nv - 1);
                const u32x4 x = ldg<NT>(col + v);
                c[u][0] = x.x; c[u][1] = x.y; c[u][2] = x.z; c[u][3] = x.w;
                raw[u] = wraw4<WP, NT>(val, v);
            }
#pragma unroll
            for (int u = 0; u < V; ++u)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const uint32_t wi = rowoff[c[u][i] >> SC_Q_SHIFT] + ((c[u][i] & SC_BIT_MASK) >> 5);
                    s[u][i] = STAGED ? win[wi] : ring32[wi];
                }
#pragma unroll
            for (int u = 0; u < V; ++u)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (4 * (b + u * TT) + i < len && ((s[u][i] >> (c[u][i] & 31)) & 1u))
                        atomicAdd(&acc[c[u][i] >> SC_Q_SHIFT], (double)wval4<WP>(raw[u], i));
        }
    }
}

// Whether one stimulus entry fires in exchange step n (draw index m):
// InputStimulation (UserInputFunctions.cu:94-118) with the rule of ScStim.  The
// three Philox blocks of the step run interleaved, round by round.
// Cosine of a float argument, rounded once to float.  Cody-Waite reduction by
// pi/2 = HI + LO (HI has 33 significant bits, so k*HI and x - k*HI are exact for
// |k| < 2^20; pi/2 - HI - LO < 4e-27), then the Taylor series of cos and sin on
// |r| <= pi/4 (truncation below 1e-17), in double with basic operations only.
// The evaluation order is fixed, so every IEEE machine gives the same bits.
__device__ __forceinline__ float stim_cosf(float x)
{
    const double HI = 0x1.921fb544p+0, LO = 0x1.0b4611a626331p-34, TWO_OVER_PI = 0x1.45f306dc9c883p-1;
    const double xd = (double)x;
    const double k = rint(xd * TWO_OVER_PI);
    const double r = (xd - k * HI) - k * LO;
    const double z = r * r;
    const double c =
        1.0 + z * (-1.0 / 2 + z * (1.0 / 24 + z * (-1.0 / 720 + z * (1.0 / 40320 + z * (-1.0 / 3628800 +
              z * (1.0 / 479001600 + z * (-1.0 / 87178291200.0 + z * (1.0 / 20922789888000.0))))))));
    const double sn =
        r + r * z * (-1.0 / 6 + z * (1.0 / 120 + z * (-1.0 / 5040 + z * (1.0 / 362880 + z * (-1.0 / 39916800 +
              z * (1.0 / 6227020800.0 + z * (-1.0 / 1307674368000.0 + z * (1.0 / 355687428096000.0))))))));
    const int q = (int)((long long)k & 3);
    const double v = q == 0 ? c : q == 1 ? -sn : q == 2 ? -c : sn;
    return (float)v;
}

// Poisson threshold of the sinusoidal component at sub-step time t: the rate of
// SinusoidalOscillation (UserInputFunctions.cu:12-14), max/2*(1 - cosf(phase)),
// in float, with the phase 2*pi*(t/sec)/osci + shift evaluated in double and
// rounded to float for cosf; then PoissonProcess's freq/sec*DT in double.
__device__ __forceinline__ double stim_sine_thr(float t, float max, float cycle, float shift)
{
    const double phase = 0x1.921fb54442d18p+2 * ((double)t / kSec) / (double)cycle + (double)shift;
    const float rate = max / 2.f * (1.f - stim_cosf((float)phase));
    return (double)rate / kSec * kDT;
}

__device__ __forceinline__ int stim_fires(const ScStim &S, int entry, int64_t n, int64_t m)
{
    const uint64_t blk = (uint64_t)(3 * m);
    uint32_t a0 = (uint32_t)blk, a1 = (uint32_t)(blk >> 32), a2 = (uint32_t)entry, a3 = 0u;
    uint32_t b0 = (uint32_t)(blk + 1), b1 = (uint32_t)((blk + 1) >> 32), b2 = a2, b3 = 0u;
    uint32_t c0 = (uint32_t)(blk + 2), c1 = (uint32_t)((blk + 2) >> 32), c2 = a2, c3 = 0u;
    uint32_t k0 = S.key0, k1 = S.key1;
#pragma unroll 1
    for (int r = 0; r < 10; ++r) {
        philox_round(a0, a1, a2, a3, k0, k1);
        philox_round(b0, b1, b2, b3, k0, k1);
        philox_round(c0, c1, c2, c3, k0, k1);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    int spk = 0;
    const double nf = (double)(float)n;
#pragma unroll 1
    for (int k = 0; k < kStepMax; ++k) {
        const float t = (float)(nf + (double)(float)k * kDT);
        const double td = (double)t;
        // draw k is word k of (a0..a3, b0..b3, c0, c1): the words move down one
        // place a sub-step, so the loop indexes no array (which would live in
        // scratch memory)
        const double rk = (double)u01(a0);
        a0 = a1; a1 = a2; a2 = a3; a3 = b0; b0 = b1; b1 = b2; b2 = b3; b3 = c0; c0 = c1;
        int f = 0;
        if (S.poisson_on <= td && td < S.poisson_off && rk < S.p_thr) f = 1;
        if (S.burst_on <= td && td < S.burst_off && rk < S.b_thr) f = 1;
        if (S.periodic_hz != 0.f && S.periodic_on <= td &&
            (S.periodic_incl ? td <= S.periodic_off : td < S.periodic_off)) {
            // PeriodicFiring (UserInputFunctions.cu:8-10)
            const float dd = t - S.periodic_start;
            const int a = (int)(((double)dd - kDT) * (double)S.periodic_hz / kSec);
            const float fp = dd * S.periodic_hz;
            const int b = (int)((double)fp / kSec);
            if (a != b) f = 1;
        }
        if (S.sine_max != 0.f && S.sine_on <= td && td < S.sine_off &&
            rk < stim_sine_thr(t, S.sine_max, S.sine_cycle, S.sine_shift))
            f = 1;
        spk |= f;
    }
    return spk;
}

// One lane task: storage form and multiplier type of its slices.
template <bool IDX16, bool STAGED, bool NT, typename Stage>
__device__ __forceinline__ void lane_dispatch(const ScLaneTask &t, const ScInstance &I, const ScPop &P, int row_n,
                                              const Stage &stage, int lane)
{
    if (IDX16 && t.idx16) {
        switch (t.wp) {
        case SC_W_F32: lane_group<SC_W_F32, true, STAGED, NT>(t, I, P, row_n, stage, lane); break;
        case SC_W_BF16: lane_group<SC_W_BF16, true, STAGED, NT>(t, I, P, row_n, stage, lane); break;
        case SC_W_FP8: lane_group<SC_W_FP8, true, STAGED, NT>(t, I, P, row_n, stage, lane); break;
        default: lane_group<SC_W_UNIT, true, STAGED, NT>(t, I, P, row_n, stage, lane); break;
        }
    } else {
        switch (t.wp) {
        case SC_W_F32: lane_group<SC_W_F32, false, STAGED, NT>(t, I, P, row_n, stage, lane); break;
        case SC_W_BF16: lane_group<SC_W_BF16, false, STAGED, NT>(t, I, P, row_n, stage, lane); break;
        case SC_W_FP8: lane_group<SC_W_FP8, false, STAGED, NT>(t, I, P, row_n, stage, lane); break;
        default: lane_group<SC_W_UNIT, false, STAGED, NT>(t, I, P, row_n, stage, lane); break;
        }
    }
}

// ---------------------------------------------------------------------- kernels
// SINGLE: one instance passed by value (kernel argument); otherwise the task's
// instance is read from the instance table.
template <bool SINGLE>
__device__ __forceinline__ const ScInstance &inst_of(const ScInstance &I0, const ScInstance *__restrict__ insts,
                                                     int i)
{
    if constexpr (SINGLE) return I0;
    else return insts[i];
}

// Gather of one team: storage form and multiplier type of its row.
template <bool IDX16, bool STAGED, bool NT>
__device__ __forceinline__ void team_dispatch(const TeamHead &T, gptr<uint32_t> ring32, const uint32_t *win,
                                              const uint32_t *rowoff, uint32_t gm, uint32_t zb, int tid, int TT,
                                              double *acc)
{
    if (T.n_grp > 0) {
        switch (T.wprec) {
        case SC_W_F32: team_groups<SC_W_F32, STAGED, NT>(T, ring32, win, rowoff, gm, tid, TT, acc); break;
        case SC_W_BF16: team_groups<SC_W_BF16, STAGED, NT>(T, ring32, win, rowoff, gm, tid, TT, acc); break;
        case SC_W_FP8: team_groups<SC_W_FP8, STAGED, NT>(T, ring32, win, rowoff, gm, tid, TT, acc); break;
        default: team_groups<SC_W_UNIT, STAGED, NT>(T, ring32, win, rowoff, gm, tid, TT, acc); break;
        }
    }
    if (IDX16 && T.hdr) {
        switch (T.wprec) {
        case SC_W_F32: team_gather<SC_W_F32, 8, true, STAGED, NT>(T, ring32, win, rowoff, gm, zb, tid, TT, acc); break;
        case SC_W_BF16: team_gather<SC_W_BF16, 8, true, STAGED, NT>(T, ring32, win, rowoff, gm, zb, tid, TT, acc); break;
        case SC_W_FP8: team_gather<SC_W_FP8, 8, true, STAGED, NT>(T, ring32, win, rowoff, gm, zb, tid, TT, acc); break;
        default: team_gather<SC_W_UNIT, 8, true, STAGED, NT>(T, ring32, win, rowoff, gm, zb, tid, TT, acc); break;
        }
    } else {
        switch (T.wprec) {
        case SC_W_F32: team_gather<SC_W_F32, 8, false, STAGED, NT>(T, ring32, win, rowoff, gm, zb, tid, TT, acc); break;
        case SC_W_BF16: team_gather<SC_W_BF16, 8, false, STAGED, NT>(T, ring32, win, rowoff, gm, zb, tid, TT, acc); break;
        case SC_W_FP8: team_gather<SC_W_FP8, 8, false, STAGED, NT>(T, ring32, win, rowoff, gm, zb, tid, TT, acc); break;
        default: team_gather<SC_W_UNIT, 8, false, STAGED, NT>(T, ring32, win, rowoff, gm, zb, tid, TT, acc); break;
        }
    }
}

// Team-mode neurons.  Teams of one wavefront come 16 to a block, wider teams
// fewer.  Before the gather, each team's first wavefront copies the population
// parameters and the per-projection factors into LDS, so nothing after the gather
// waits on global memory but the state store and the spike bit.  After the gather
// that wavefront turns its row sums into dg and, in blocks of several teams,
// parks dg and the neuron state in LDS; then the block's first wavefront runs the
// neuron updates of all its teams side by side (one lane per team) instead of one
// lane in each of up to 16 wavefronts, which would make the SIMDs issue the long
// update 16 times.
template <bool SINGLE, bool IDX16, bool NT, typename Stage>
__device__ __forceinline__ void do_team(const ScInstance &I0, const ScInstance *__restrict__ insts, int t, int count,
                                        const TeamHead &h, const ScTeamIn &pin, int row_single, int k,
                                        const Stage &stage)
{
    constexpr int NTEAM = SC_BLOCK / SC_WAVE;
    __shared__ double acc[NTEAM][SC_MAX_IN];
    __shared__ uint32_t rowoff[NTEAM][SC_MAX_IN];
    constexpr int PW = (SC_POP_WORDS + 1) / 2 * 2;
    __shared__ alignas(8) uint32_t par[NTEAM][PW];    // population parameters (8-byte aligned rows)
    __shared__ float fac[NTEAM][3][SC_MAX_IN];  // per projection: scale, W, excitatory flag (as bits)
    __shared__ float sv[NTEAM][10];  // u, g_exc, g_inh, dge, dgi; refr, device index, ring row, Wr, R (as bits)
    __shared__ void *ptr[NTEAM][2];  // instance arrays (state, ring) of each team's network
    const int wave = uniform(threadIdx.x / SC_WAVE), lane = threadIdx.x & (SC_WAVE - 1);
    const int team = wave / t, tw = wave - team * t;
    const bool active = team < count;
    const bool gather_updates = count > 1;
    const ScInstance &I = inst_of<SINGLE>(I0, insts, h.inst);
    int row_n = row_single;
    if (!SINGLE && active) row_n = uniform(run_row(I, k));
    float u = 0.f;
    float2 gg = {0.f, 0.f};
    int32_t rf = 0;
    const int d = h.group * SC_WAVE + h.bit;
    // the neuron state and the population parameters are loaded (the state while
    // the gather runs), then the block's window copies are issued, so both are in
    // flight together
    uint32_t pv = 0u;
    if (active && tw == 0) {
        if (lane < SC_POP_WORDS) pv = gbl(reinterpret_cast<const uint32_t *>(&I.pops[h.pop]))[lane];
        if (lane == 0) load_state(I.st, d, u, gg, rf);
    }
    const uint32_t *win = stage();
    if (active && tw == 0) {
        if (lane < SC_MAX_IN) {
            acc[team][lane] = 0.0;
            // word offset of each projection's ring row: in global memory, or in a
            // staged block the LDS word that stands for word 0 of the row
            rowoff[team][lane] = lane >= h.n_in ? 0u
                                 : (win && pin.lds != SC_LDS_GLOBAL)
                                     ? (uint32_t)pin.lds
                                     : (uint32_t)row_back(row_n, pin.delay, I.R) * (uint32_t)(2 * I.Wr);
            if (lane < h.n_in) {  // the factors that turn row sums into dg
                fac[team][0][lane] = pin.scale;
                fac[team][1][lane] = pin.W;
                fac[team][2][lane] = __int_as_float(pin.exc);
            }
        }
        if (lane < SC_POP_WORDS) par[team][lane] = pv;
    }
    // projections of a staged block left in memory (their windows did not fit)
    const uint32_t gmask = win ? (uint32_t)__ballot(lane < h.n_in && pin.lds == SC_LDS_GLOBAL) : 0u;
    __syncthreads();  // staged block: the windows are complete past this barrier
    if (active) {
        // the team's network is the same for the whole wavefront
        gptr<uint32_t> ring32 = gbl(reinterpret_cast<const uint32_t *>(uniform_ptr(I.ring)));
        const int tid = tw * SC_WAVE + lane, TT = t * SC_WAVE;
        const uint32_t zb = (uint32_t)uniform((I.Wr - 1) * SC_WAVE);
        if (win) team_dispatch<IDX16, true, NT>(h, ring32, win, rowoff[team], gmask, zb, tid, TT, acc[team]);
        else team_dispatch<IDX16, false, NT>(h, ring32, win, rowoff[team], 0u, zb, tid, TT, acc[team]);
    }
    __syncthreads();
    float dge = 0.f, dgi = 0.f;
    if (active && tw == 0) {
        // dg[post] += S * w_bar (Simulation_header.cu:98): lane j forms the term of
        // projection j, then the terms are summed in projection index order
        const int n_in = h.n_in;
        float add = 0.f;
        int ex = 0;
        if (lane < n_in) {
            add = (float)(acc[team][lane] * (double)fac[team][0][lane]) * fac[team][1][lane];
            ex = __float_as_int(fac[team][2][lane]);
        }
        const uint64_t em = __ballot(ex != 0);
        for (int q = 0; q < n_in; ++q) {
            const float a = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(add), q));
            if ((em >> q) & 1) dge = dge + a;
            else dgi = dgi + a;
        }
        if (gather_updates && lane == 0) {
            sv[team][0] = u;
            sv[team][1] = gg.x;
            sv[team][2] = gg.y;
            sv[team][3] = dge;
            sv[team][4] = dgi;
            sv[team][5] = __int_as_float(rf);
            sv[team][6] = __int_as_float(d);
            if (!SINGLE) {  // teams of a batch block may belong to different instances
                sv[team][7] = __int_as_float(row_n);
                sv[team][8] = __int_as_float(I.Wr);
                sv[team][9] = __int_as_float(I.R);
                ptr[team][0] = I.st;
                ptr[team][1] = I.ring;
            }
        }
    }
    int j = team;  // the team whose update this lane runs
    bool upd = active && tw == 0 && lane == 0;
    int dj = d;
    ScState *ST = I.st;
    uint64_t *ring = I.ring;
    int Wr = I.Wr, R = I.R, rown = row_n;
    if (gather_updates) {
        __syncthreads();
        j = lane;
        upd = wave == 0 && lane < count;
        if (upd) {
            u = sv[j][0];
            gg = make_float2(sv[j][1], sv[j][2]);
            dge = sv[j][3];
            dgi = sv[j][4];
            rf = __float_as_int(sv[j][5]);
            dj = __float_as_int(sv[j][6]);
            if (!SINGLE) {
                rown = __float_as_int(sv[j][7]);
                Wr = __float_as_int(sv[j][8]);
                R = __float_as_int(sv[j][9]);
                ST = static_cast<ScState *>(ptr[j][0]);
                ring = static_cast<uint64_t *>(ptr[j][1]);
            }
        }
    }
    if (upd) {
        const int spk = lif_step(*reinterpret_cast<const ScPop *>(par[j]), u, gg.x, gg.y, rf, dge, dgi);
        const int group = dj / SC_WAVE, bit = dj & (SC_WAVE - 1);
        store_state(ST, dj, u, gg, rf);
        if (spk) atomicOr(&ring[(int64_t)rown * Wr + group], 1ull << bit);
        // the word of the next step starts empty (that row is not read during this step)
        if (bit == 0) {
            const int row1 = rown + 1 == R ? 0 : rown + 1;
            ring[(int64_t)row1 * Wr + group] = 0ull;
        }
    }
}

// Window counts of the plasticity rules for step n: per neuron of a rule's
// presynaptic population, one byte of the rule's `hist` holds its spikes in rows
// n - delay - s (s < window) and, in bit 7, its spike at n - delay.  Those rows are
// complete before step n starts (every delay is at least one step).  Wavefront
// `item` of the class takes SC_HIST_GROUPS 64-neuron groups of one rule, one lane
// per neuron of each.  The count is updated from the previous step's,
//   count(n) = count(n - 1) + [spike at n - delay] - [spike at n - delay - window],
// which reads two ring words per group (the ring holds delay + window + 2 rows, so
// the row leaving the window is still intact).  In the first step after the ring
// changed outside the step sequence (ctl->recount, cleared by the k_plastic launch
// that follows the step) the whole window is counted again, 16 rows in flight.
__device__ __forceinline__ void plastic_hist(const ScInstance &I, const ScPlastic *__restrict__ rules, int item,
                                             int row_n, int lane)
{
    const ScPlastic *P = rules;
    for (;;) {  // the rule of this item
        const int nw = (P->pre_groups + SC_HIST_GROUPS - 1) / SC_HIST_GROUPS;
        if (item < nw) break;
        item -= nw;
        ++P;
    }
    const int R = I.R, window = P->window;
    const int pre_t = row_back(row_n, P->delay, R);
    const int g0 = item * SC_HIST_GROUPS, g1 = min(P->pre_groups, g0 + SC_HIST_GROUPS);
    const uint64_t *ring = I.ring + P->pre_group0;
    uint8_t *hist = P->hist;
    if (uniform(I.ctl->recount)) {
        for (int g = g0; g < g1; ++g) {
            const uint64_t *col = ring + g;
            int cnt = 0, s0 = 0;
            for (int b = 0; b < window; b += 16) {
                uint64_t w[16];
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    w[u] = b + u < window ? col[(int64_t)row_back(pre_t, b + u, R) * I.Wr] : 0ull;
#pragma unroll
                for (int u = 0; u < 16; ++u) cnt += (int)((w[u] >> lane) & 1ull);
                if (b == 0) s0 = (int)(w[0] >> lane) & 1;
            }
            hist[(int64_t)g * SC_WAVE + lane] = (uint8_t)(cnt | (s0 << 7));
        }
        return;
    }
    const int64_t r_in = (int64_t)pre_t * I.Wr, r_out = (int64_t)row_back(pre_t, window, R) * I.Wr;
    uint64_t a[SC_HIST_GROUPS], b[SC_HIST_GROUPS];
    uint32_t old[SC_HIST_GROUPS];
#pragma unroll
    for (int u = 0; u < SC_HIST_GROUPS; ++u) {
        const int g = g0 + u;
        const bool v = g < g1;
        a[u] = v ? ring[r_in + g] : 0ull;
        b[u] = v ? ring[r_out + g] : 0ull;
        old[u] = v ? hist[(int64_t)g * SC_WAVE + lane] : 0u;
    }
#pragma unroll
    for (int u = 0; u < SC_HIST_GROUPS; ++u) {
        const int g = g0 + u;
        if (g >= g1) break;
        const uint32_t in = (uint32_t)(a[u] >> lane) & 1u, out = (uint32_t)(b[u] >> lane) & 1u;
        hist[(int64_t)g * SC_WAVE + lane] = (uint8_t)((((old[u] & 127u) + in - out) & 127u) | (in << 7));
    }
}

template <bool SINGLE>
__device__ __forceinline__ const uint32_t *stage_windows(uint32_t sd, uint32_t *win_lds, const ScInstance &I0,
                                                         const ScInstance *__restrict__ insts, int row_single,
                                                         int k, int wave, int lane)
{
    constexpr int W0 = (int)(offsetof(ScStage, w) / 4), WS = (int)(sizeof(ScWin) / 4);
    constexpr int F_INST = (int)(offsetof(ScWin, inst) / 4), F_DELAY = (int)(offsetof(ScWin, delay) / 4);
    constexpr int F_W0 = (int)(offsetof(ScWin, w0) / 4), F_NW = (int)(offsetof(ScWin, nw) / 4);
    constexpr int F_OFF = (int)(offsetof(ScWin, off) / 4);
    const int nwin = __builtin_amdgcn_readlane((int)sd, (int)(offsetof(ScStage, n) / 4));
    if (nwin <= 0) return nullptr;
    u32x4 *dst = reinterpret_cast<u32x4 *>(win_lds);
    for (int w = 0; w < nwin; ++w) {
        const int base = W0 + w * WS;
        const int inst = __builtin_amdgcn_readlane((int)sd, base + F_INST);
        const int delay = __builtin_amdgcn_readlane((int)sd, base + F_DELAY);
        const int w0 = __builtin_amdgcn_readlane((int)sd, base + F_W0);
        const int n4 = __builtin_amdgcn_readlane((int)sd, base + F_NW) >> 2;
        const int o4 = __builtin_amdgcn_readlane((int)sd, base + F_OFF) >> 2;
        const ScInstance &I = inst_of<SINGLE>(I0, insts, inst);
        const int rn = SINGLE ? row_single : run_row(I, k);
        const u32x4 *src = reinterpret_cast<const u32x4 *>(
            reinterpret_cast<const uint32_t *>(I.ring + (int64_t)row_back(rn, delay, I.R) * I.Wr) + w0);
        for (int cb = wave * SC_WAVE; cb < n4; cb += SC_BLOCK)
            if (cb + lane < n4) lds_load16(src + cb + lane, dst + o4 + cb);
    }
    if (threadIdx.x == 0) dst[0] = u32x4{0u, 0u, 0u, 0u};
    return win_lds;
}

// The block layout is a kernel argument: block b belongs to the first class c with
// b < cls_end[c]; classes 0..4 are team neurons of width 16, 8, 4, 2, 1
// wavefronts, class 5 lane groups, class 6 stimulus groups.
template <bool SINGLE, bool IDX16, bool NT>
__device__ __forceinline__ void step_body(const ScInstance &I0, const ScInstance *__restrict__ insts, const ScLayout &L,
                                          const ScLaneTask *__restrict__ lanes, const ScStimTask *__restrict__ stt,
                                          const ScTeamTask *__restrict__ teams, const ScStage *__restrict__ stages,
                                          const ScPlastic *__restrict__ plastic, int k)
{
    // the step counter load is issued first; it overlaps the descriptor loads
    int row_single = 0;
    if (SINGLE) {
        row_single = I0.ctl->row + k;
        if (row_single >= I0.R) row_single -= I0.R;
    }
    const int b = blockIdx.x;
    int c = 0;
#pragma unroll
    for (int j = 0; j < SC_CLASSES - 1; ++j) c += (b >= L.cls_end[j]) ? 1 : 0;
    const int bi = b - (c ? L.cls_end[c - 1] : 0);
    const int wave = uniform(threadIdx.x / SC_WAVE), lane = threadIdx.x & (SC_WAVE - 1);
    // staged block: its spike-bitmap windows are copied into LDS, so that its
    // gathers read LDS instead of scattering over the ring rows in memory.  The
    // block's window table arrives with one load (a dword per lane); the copies
    // are direct global-to-LDS loads, issued once the block's first descriptor
    // loads are under way and waited for only at the block's first barrier, so
    // the copies overlap the loads of the neuron state, the population
    // parameters and (lane blocks) the first columns.  Stimulus blocks are never
    // staged.
    __shared__ __attribute__((aligned(16))) uint32_t win_lds[SC_WIN_WORDS];
    uint32_t sd = 0u;
    if (stages && lane < SC_STAGE_DWORDS) sd = gbl(reinterpret_cast<const uint32_t *>(stages + b))[lane];
    const auto stage = [&]() { return stage_windows<SINGLE>(sd, win_lds, I0, insts, row_single, k, wave, lane); };
    if (c < SC_CLASS_LANE) {
        const int t = SC_MAX_TEAM >> c, per = 1 << c;
        const int first = L.cls_task0[c] + bi * per, count = min(per, L.cls_ntask[c] - bi * per);
        // wavefronts past the block's teams read the last team's descriptor (unused)
        const ScTeamTask *T = &teams[first + min(wave / t, count - 1)];
        const TeamHead h = team_head(T);
        ScTeamIn pin{};
        if (lane < SC_MAX_IN) pin = T->in[lane];
        do_team<SINGLE, IDX16, NT>(I0, insts, t, count, h, pin, row_single, k, stage);
        return;
    }
    if (c == SC_CLASS_HIST) {  // one instance (the batch holds no plasticity rules)
        const int item = uniform(bi * (SC_BLOCK / SC_WAVE) + wave);
        if (item < L.cls_ntask[c]) plastic_hist(I0, plastic, item, row_single, lane);
        return;
    }
    const int per_block = c == SC_CLASS_LANE ? SC_LANE_PER_BLOCK : SC_STIM_PER_BLOCK;
    const int idx = uniform(bi * per_block + wave);
    const bool staged = c == SC_CLASS_LANE && __builtin_amdgcn_readlane((int)sd, (int)(offsetof(ScStage, n) / 4)) > 0;
    if (wave >= per_block || idx >= L.cls_ntask[c]) {
        if (staged) {  // the wavefront's share of the window copies, complete before it ends
            stage();
            __syncthreads();
        }
        return;
    }
    if (c == SC_CLASS_LANE) {
        const ScLaneTask &t = lanes[idx];
        const ScInstance &I = inst_of<SINGLE>(I0, insts, t.inst);
        const int row_n = SINGLE ? row_single : uniform(run_row(I, k));
        const ScPop &P = I.pops[t.pop];
        if (staged) lane_dispatch<IDX16, true, NT>(t, I, P, row_n, stage, lane);
        else lane_dispatch<IDX16, false, NT>(t, I, P, row_n, stage, lane);
    } else {  // stimulus blocks (never staged)
        const ScStimTask t = stt[idx];
        const ScInstance &I = inst_of<SINGLE>(I0, insts, t.inst);
        const int row_n = SINGLE ? row_single : run_row(I, k);
        const int d = t.group * SC_WAVE + lane;
        int spk = 0;
        if (lane < t.nvalid) {
            const int s = I.stim_s[d];
            if (s >= 0) spk = stim_fires(I.stims[s], I.stim_e[d], I.ctl->n + k, I.ctl->m + k);
        }
        const unsigned long long word = __ballot(spk != 0);
        if (lane == 0) I.ring[(int64_t)row_n * I.Wr + t.group] = word;
    }
}

// One network: 64 VGPRs at most, so two 1024-thread blocks share a CU and one
// block's serial phases (descriptor loads, barriers, the neuron update) overlap
// the other's gathers.  IDX16: the network holds 16-bit column offsets (the
// instantiation without them carries none of their code or registers).  NT: the
// synapse streams are loaded nontemporal (networks whose streams exceed the L2).
template <bool IDX16, bool NT>
__global__ void __launch_bounds__(SC_BLOCK) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_step(ScInstance I0, ScLayout L, const ScLaneTask *__restrict__ lanes, const ScStimTask *__restrict__ stt,
       const ScTeamTask *__restrict__ teams, const ScStage *__restrict__ stages,
       const ScPlastic *__restrict__ plastic, int k)
{
    step_body<true, IDX16, NT>(I0, nullptr, L, lanes, stt, teams, stages, plastic, k);
}

// Many instances: every task names its instance in the instance table.  Held to
// 64 VGPRs as well (some registers spill to scratch, more in the 16-bit
// instantiation): the doubled occupancy outweighs the spill traffic.
template <bool IDX16, bool NT>
__global__ void __launch_bounds__(SC_BLOCK) __attribute__((amdgpu_waves_per_eu(8, 8))) k_step_batch(const ScInstance *__restrict__ insts, ScLayout L,
                                                         const ScLaneTask *__restrict__ lanes,
                                                         const ScStimTask *__restrict__ stt,
                                                         const ScTeamTask *__restrict__ teams,
                                                         const ScStage *__restrict__ stages, int k)
{
    step_body<false, IDX16, NT>(ScInstance{}, insts, L, lanes, stt, teams, stages, nullptr, k);
}

// Per-population spike counts and synaptic events of K consecutive steps.  The
// ring row of each (instance, step) is cut into chunks of SC_STATS_CHUNK spike
// words, one block each: popcounts per population and the out-degrees of the set
// bits, summed in LDS, then added with integer atomics to the row's counts and
// events (zeroed first by k_stats_zero) and to the running totals.  Integer sums:
// the result does not depend on the order of the blocks.  With one chunk per row
// (nchunk 1) the block writes the row's counts and events directly, unzeroed.
#define SC_STATS_CHUNK 2048
template <bool SINGLE>
__global__ void __launch_bounds__(256) k_stats_zero(ScInstance I0, const ScInstance *__restrict__ insts, int K)
{
    const int inst = blockIdx.x / K, kk = blockIdx.x - inst * K;
    const ScInstance &I = inst_of<SINGLE>(I0, insts, inst);
    const int row = run_row(I, kk);
    if ((int)threadIdx.x < I.n_pops) I.counts[(int64_t)row * I.n_pops + threadIdx.x] = 0;
    if (threadIdx.x == 0) I.events[row] = 0ull;
}

// One statistics block (block `b` of K x instances x nchunk; blockDim = SC_STATS_CHUNK / 2).
template <bool SINGLE>
__device__ __forceinline__ void stats_block(const ScInstance &I0, const ScInstance *__restrict__ insts, int K,
                                            int nchunk, int b)
{
    __shared__ unsigned long long acc[SC_MAX_POPS + 1];
    const int chunk = b % nchunk, rest = b / nchunk;
    const int inst = rest / K, kk = rest - inst * K;
    const ScInstance &I = inst_of<SINGLE>(I0, insts, inst);
    const int g0 = chunk * SC_STATS_CHUNK;
    if (chunk > 0 && g0 >= I.n_groups) return;  // the whole block: this instance has fewer groups
    const int g1 = min(I.n_groups, g0 + SC_STATS_CHUNK);
    if (threadIdx.x <= SC_MAX_POPS) acc[threadIdx.x] = 0ull;
    __syncthreads();
    const int row = run_row(I, kk);
    const uint64_t *__restrict__ w = I.ring + (int64_t)row * I.Wr;
    unsigned long long ev = 0ull;
    // popcount into the population's counter, then the out-degrees of the set
    // bits, 8 loads in flight at a time
    const auto word_stats = [&](uint64_t word, int pop, int g) {
        if (!word) return;
        atomicAdd(&acc[pop], (unsigned long long)__popcll(word));
        const uint32_t *od = I.outdeg + (int64_t)g * SC_WAVE;
        while (word) {
            uint32_t v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                v[u] = 0u;
                if (word) {
                    v[u] = od[__ffsll((long long)word) - 1];
                    word &= word - 1;
                }
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) ev += v[u];
        }
    };
    // two words per thread (one pass over a chunk of 2048 words with 1024
    // threads), their population ids loaded alongside
    for (int g = g0 + threadIdx.x; g < g1; g += 2 * blockDim.x) {
        const int gb = g + blockDim.x;
        const uint64_t wa = w[g];
        const int pa = I.groups[g].pop;
        uint64_t wb = 0ull;
        int pb = 0;
        if (gb < g1) {
            wb = w[gb];
            pb = I.groups[gb].pop;
        }
        word_stats(wa, pa, g);
        word_stats(wb, pb, gb);
    }
    if (ev) atomicAdd(&acc[SC_MAX_POPS], ev);
    __syncthreads();
    const int p = threadIdx.x;
    if (p < I.n_pops) {
        if (nchunk == 1) I.counts[(int64_t)row * I.n_pops + p] = (int32_t)acc[p];
        else if (acc[p]) atomicAdd(&I.counts[(int64_t)row * I.n_pops + p], (int32_t)acc[p]);
        if (acc[p]) atomicAdd(&I.totals[p], acc[p]);
    }
    if (p == 0) {
        if (nchunk == 1) I.events[row] = acc[SC_MAX_POPS];
        else if (acc[SC_MAX_POPS]) atomicAdd(&I.events[row], acc[SC_MAX_POPS]);
        if (acc[SC_MAX_POPS]) atomicAdd(&I.totals[I.n_pops], acc[SC_MAX_POPS]);
    }
}

// The end of a run of K steps in one launch: blocks [0, n_stats) are the
// statistics blocks above; the blocks after them (one instance, O.out set) write
// the output spike train of the run's last O.T steps: byte (s, c) is the spike of
// output neuron c (device bit O.bits[c]) at step s, row-major by step, as
// loop_n_steps hands it over (GPUScaffoldSimulation.cu:777).  Every block reads the
// step counters (the run's ring rows); once it no longer does, it counts itself in
// *done, and the last block to arrive advances the counters of every instance by K
// (n, m, ring row; the window-count recount flag cleared) and resets *done for the
// next launch.
template <bool SINGLE>
__global__ void __launch_bounds__(SC_STATS_CHUNK / 2) k_run_end(ScInstance I0, const ScInstance *__restrict__ insts,
                                                                 int n_inst, int K, int nchunk, int n_stats, ScOut O,
                                                                 unsigned int *__restrict__ done)
{
    __shared__ int last;
    const int b = blockIdx.x;
    if (b < n_stats) {
        stats_block<SINGLE>(I0, insts, K, nchunk, b);
    } else {
        const int i = (b - n_stats) * (int)blockDim.x + (int)threadIdx.x;
        if (i < O.T * O.n_out) {
            const int st = i / O.n_out, c = i - st * O.n_out;
            const int row = run_row(I0, K - O.T + st);
            const int bit = O.bits[c];
            O.out[i] = (uint8_t)((I0.ring[(int64_t)row * I0.Wr + (bit >> 6)] >> (bit & 63)) & 1ull);
        }
    }
    __syncthreads();  // every read of the counters in this block is done
    if (threadIdx.x == 0) {
        __threadfence();
        last = atomicAdd(done, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    for (int i = threadIdx.x; i < n_inst; i += blockDim.x) {
        const ScInstance &I = inst_of<SINGLE>(I0, insts, i);
        ScCtl *c = I.ctl;
        c->n += K;
        c->m += K;
        c->row = (c->row + K) % I.R;
        c->recount = 0;
    }
    if (threadIdx.x == 0) *done = 0u;
}

__global__ void k_set_ctl(ScCtl *ctl, int64_t n, int64_t m, int row)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl->n = n;
        ctl->m = m;
        ctl->row = row;
        ctl->recount = 1;
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
    if (d == 0) I.ctl->recount = 1;  // ring row 0 changed
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
    I.st[d].V = P.Vr + prod;
    I.st[d].refr = 0;
}

// Conductances of every device neuron set to zero (the other fields kept).
__global__ void k_zero_g(ScState *st, int n)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d < n) {
        st[d].g_exc = 0.f;
        st[d].g_inh = 0.f;
    }
}

// One plasticity rule after exchange step n (launch offset k): Hebb_plasticity /
// Teacher_plasticity (SynapticPlasticity.cu:38-121), one wavefront per chunk of
// synapses of one post row, SC_PLASTIC_CHUNK / 64 synapses per lane.  The loads
// form four dependent rounds: the chunk's descriptor with the ring position, the
// trigger's spike word, every lane's synapse records (lanes past the chunk's end
// load its last record and store nothing), then the window counts and
// multipliers.  dw is summed in float in the
// reference's order: A0; A1_J if the presynaptic neuron spiked at n - delay; A1_I
// if the row's trigger spiked -- the postsynaptic neuron at n (Hebb) or the
// teacher at n - t_delay (Teacher); with that spike, A2_JI once for every spike
// of the presynaptic neuron at n - delay - s, s < window (all A2_JI terms are the
// same constant, so their count fixes the float sum).  Then w + dw is clamped
// (Hebb to [0, 1], Teacher to >= 0) and rounded to a multiple of 2^grid_exp,
// which keeps every row sum of the projection exact in f64 (so the order of the
// gather's additions cannot change it).
__global__ void __launch_bounds__(256) k_plastic(ScInstance I, ScPlastic P, int k)
{
    const int c = uniform((int)blockIdx.x * 4 + (int)(threadIdx.x / SC_WAVE));
    const int lane = threadIdx.x & (SC_WAVE - 1);
    if (c < P.n_chunks) {
        const int32_t trig = uniform(P.ch_trig[c]);
        const uint32_t b0 = (uint32_t)uniform((int)P.ch_begin[c]), b1 = (uint32_t)uniform((int)P.ch_begin[c + 1]);
        const int R = I.R;
        const int row = (I.ctl->row + k) % R;
        int s_post = 0;
        if (trig >= 0) {
            const int r = P.rule == SC_STDP_HEBB ? row : row_back(row, P.t_delay, R);
            s_post = (int)((I.ring[(int64_t)r * I.Wr + (trig >> 6)] >> (trig & 63)) & 1ull);
        }
        s_post = uniform(s_post);
        if (!P.skip_quiet || s_post) {
            constexpr int U = SC_PLASTIC_CHUNK / SC_WAVE;
            uint32_t vi[U], pr[U], h[U];
            float w[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t i = min(b0 + lane + u * SC_WAVE, b1 - 1u);  // a chunk holds >= 1 synapse
                vi[u] = P.vidx[i];
                pr[u] = P.pre[i];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                h[u] = P.hist[pr[u]];
                w[u] = P.val[vi[u]];
            }
            // every lane computes (its loads are then needed on every lane and stay
            // ahead of the stores together); lanes past the chunk's end store nothing
#pragma unroll
            for (int u = 0; u < U; ++u) {
                // coefficient c for this synapse: a constant, or the weight-dependent
                // value (float)(ac + aa * (ab - (double)w)) of the multiplier before the update
                const auto coef = [&](int c) {
                    return ((P.affine >> c) & 1) ? (float)(P.ac[c] + P.aa[c] * (P.ab[c] - (double)w[u])) : P.coef[c];
                };
                float dw = 0.f;
                if (P.present & 1) dw += coef(0);
                if ((P.present & 4) && (h[u] >> 7)) dw += coef(2);
                if ((P.present & 2) && s_post) dw += coef(1);
                if ((P.present & 8) && s_post) {
                    const float a2 = coef(3);  // the same value for every s of the window (a function of w only)
                    for (int q = 0; q < (int)(h[u] & 127u); ++q) dw += a2;
                }
                float nw = w[u] + dw;
                if (P.rule == SC_STDP_HEBB) nw = nw < 0.f ? 0.f : nw > 1.f ? 1.f : nw;
                else nw = nw < 0.f ? 0.f : nw;
                w[u] = ldexpf(rintf(ldexpf(nw, -P.grid_exp)), P.grid_exp);
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (b0 + lane + u * SC_WAVE < b1) P.val[vi[u]] = w[u];
        }
    }
    // the window counts of this step are complete (the step kernel ran before this
    // launch): the next step updates them from these
    if (c == 0 && lane == 0) I.ctl->recount = 0;
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

// Output spike train of T consecutive steps, the first in ring row row0: byte
// (s, c) is the spike of output neuron c (device bit bits[c]) at step s, row-major
// by step, as loop_n_steps hands it over (GPUScaffoldSimulation.cu:777).
__global__ void __launch_bounds__(256) k_out_train(const uint64_t *__restrict__ ring, int Wr, int R, int row0, int T,
                                                   const int32_t *__restrict__ bits, int n_out,
                                                   uint8_t *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= T * n_out) return;
    const int s = i / n_out, c = i - s * n_out;
    int row = row0 + s;
    if (row >= R) row -= R;
    const int b = bits[c];
    out[i] = (uint8_t)((ring[(int64_t)row * Wr + (b >> 6)] >> (b & 63)) & 1ull);
}

__global__ void k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// --------------------------------------------------------------------- launchers
template <bool NT>
static void launch_step_nt(hipStream_t s, const ScInstance *single, const ScInstance *insts, const DevWork &w, int k)
{
    if (single && w.idx16)
        k_step<true, NT><<<w.n_blocks, SC_BLOCK, 0, s>>>(*single, w.layout, w.lanes, w.stims, w.teams, w.stages,
                                                         w.d_plastic, k);
    else if (single)
        k_step<false, NT><<<w.n_blocks, SC_BLOCK, 0, s>>>(*single, w.layout, w.lanes, w.stims, w.teams, w.stages,
                                                          w.d_plastic, k);
    else if (w.idx16)
        k_step_batch<true, NT><<<w.n_blocks, SC_BLOCK, 0, s>>>(insts, w.layout, w.lanes, w.stims, w.teams, w.stages, k);
    else
        k_step_batch<false, NT><<<w.n_blocks, SC_BLOCK, 0, s>>>(insts, w.layout, w.lanes, w.stims, w.teams, w.stages, k);
}

void launch_step(hipStream_t s, const ScInstance *single, const ScInstance *insts, const DevWork &w, int k)
{
    if (w.n_blocks <= 0) return;
    if (w.nontemporal) launch_step_nt<true>(s, single, insts, w, k);
    else launch_step_nt<false>(s, single, insts, w, k);
}

void launch_run_end(hipStream_t s, const ScInstance *single, const ScInstance *insts, int n_inst, int K,
                    int max_groups, const ScOut *out, unsigned int *done)
{
    if (K <= 0 || n_inst <= 0) return;
    const int nchunk = std::max(1, (max_groups + SC_STATS_CHUNK - 1) / SC_STATS_CHUNK);
    const int n_stats = K * n_inst * nchunk;
    constexpr int bs = SC_STATS_CHUNK / 2;
    ScOut o{};
    int n_out_blocks = 0;
    if (single && out && out->out && out->T * out->n_out > 0) {
        o = *out;
        n_out_blocks = (o.T * o.n_out + bs - 1) / bs;
    }
    if (single) {
        if (nchunk > 1) k_stats_zero<true><<<K, 256, 0, s>>>(*single, insts, K);
        k_run_end<true><<<n_stats + n_out_blocks, bs, 0, s>>>(*single, insts, 1, K, nchunk, n_stats, o, done);
    } else {
        if (nchunk > 1) k_stats_zero<false><<<K * n_inst, 256, 0, s>>>(ScInstance{}, insts, K);
        k_run_end<false><<<n_stats, bs, 0, s>>>(ScInstance{}, insts, n_inst, K, nchunk, n_stats, o, done);
    }
}

void launch_set_ctl(hipStream_t s, ScCtl *ctl, int64_t n, int64_t m, int row)
{
    k_set_ctl<<<1, 64, 0, s>>>(ctl, n, m, row);
}

void launch_init_params(hipStream_t s, const ScInstance *di, int n_threads, uint64_t seed, uint64_t off)
{
    k_init_params<<<(n_threads + 255) / 256, 256, 0, s>>>(di, (uint32_t)seed, (uint32_t)(seed >> 32), off);
}

void launch_compact(hipStream_t s, const uint64_t *row, int n_groups, const ScGroup *groups, const ScPop *pops,
                    int32_t *out, int cap, int32_t *count)
{
    k_compact<<<1, 1024, 0, s>>>(row, n_groups, groups, pops, out, cap, count);
}

void launch_zero_g(hipStream_t s, ScState *st, int n)
{
    if (n > 0) k_zero_g<<<(n + 255) / 256, 256, 0, s>>>(st, n);
}

void launch_plastic(hipStream_t s, const ScInstance &I, const DevWork &w, int k)
{
    for (const ScPlastic &P : w.plastic) {
        if (P.n_chunks <= 0) continue;
        k_plastic<<<(unsigned)((P.n_chunks + 3) / 4), 256, 0, s>>>(I, P, k);
    }
}

void launch_out_train(hipStream_t s, const uint64_t *ring, int Wr, int R, int row0, int T, const int32_t *bits,
                      int n_out, uint8_t *out)
{
    const int n = T * n_out;
    if (n > 0) k_out_train<<<(n + 255) / 256, 256, 0, s>>>(ring, Wr, R, row0, T, bits, n_out, out);
}

void launch_copy(hipStream_t s, const void *src, void *dst, int64_t n16, int grid)
{
    k_copy<<<grid, 256, 0, s>>>(static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), n16);
}

}  // namespace sc
