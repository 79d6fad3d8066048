// kernels.hpp -- launchers of the gfx950 kernels (kernels.hip) used by the engine.
#pragma once
#include <hip/hip_runtime.h>

#include "model.hpp"

namespace sc {

#define HIP_TRY(x)                                                                                   \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess)                                                                        \
            return fail(SCAFFOLD_ERR_HIP, "%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

// Device work of one step: blocks and the descriptor arrays they index.
struct DevWork {
    ScLayout layout{};
    int n_blocks = 0;
    const ScLaneTask *lanes = nullptr;
    const ScStimTask *stims = nullptr;
    const ScTeamTask *teams = nullptr;
    bool idx16 = false;  // some descriptor holds 16-bit column offsets
    bool nontemporal = false;  // synapse streams loaded as nontemporal (networks beyond the L2)
    const ScStage *stages = nullptr;  // per block: spike windows staged in LDS (nullptr: none)
    int n_staged = 0;                 // blocks with staged windows
    int max_groups = 0;               // most 64-neuron groups of one instance (statistics grid)
    std::vector<ScPlastic> plastic;   // plasticity rules, applied in order after each step
    const ScPlastic *d_plastic = nullptr;  // the rules on the device (window-count blocks of the step kernel)
    unsigned int *end_done = nullptr;      // arrival counter of the run-end kernel (zero between launches)
};

// Output spike train of a run's last T steps: n_out output neurons at device bits
// bits[0..n_out), written T x n_out bytes, row-major by step, to out.
struct ScOut {
    const int32_t *bits = nullptr;
    uint8_t *out = nullptr;
    int T = 0, n_out = 0;
};

// All launchers are asynchronous on stream s.  `single` != nullptr selects the
// kernels that take one instance by value; otherwise instances come from `insts`.
void launch_step(hipStream_t s, const ScInstance *single, const ScInstance *insts, const DevWork &w, int k);
// The end of a run of K steps, one launch (plus a zeroing launch when a ring row
// spans several statistics chunks): per-population counts and synaptic events of
// the run's rows, the output spike train of its last out->T steps when out is set
// (one instance only), then the advance of every instance's step counters by K.
// `done` is a zeroed device counter kept for this launch sequence (the kernel
// leaves it zero again).
void launch_run_end(hipStream_t s, const ScInstance *single, const ScInstance *insts, int n_inst, int K,
                    int max_groups, const ScOut *out, unsigned int *done);
void launch_set_ctl(hipStream_t s, ScCtl *ctl, int64_t n, int64_t m, int row);
void launch_init_params(hipStream_t s, const ScInstance *di, int n_threads, uint64_t seed, uint64_t off);
void launch_compact(hipStream_t s, const uint64_t *row, int n_groups, const ScGroup *groups, const ScPop *pops,
                    int32_t *out, int cap, int32_t *count);
void launch_copy(hipStream_t s, const void *src, void *dst, int64_t n16, int grid);
// Output spike train: T steps from ring row row0 (wrapping at R), one byte per
// (step, output neuron), output neurons at device bits bits[0..n_out).
void launch_out_train(hipStream_t s, const uint64_t *ring, int Wr, int R, int row0, int T, const int32_t *bits,
                      int n_out, uint8_t *out);
void launch_zero_g(hipStream_t s, ScState *st, int n);
// The plasticity rules of one network after step k of a run (single instance).
void launch_plastic(hipStream_t s, const ScInstance &I, const DevWork &w, int k);

}  // namespace sc
