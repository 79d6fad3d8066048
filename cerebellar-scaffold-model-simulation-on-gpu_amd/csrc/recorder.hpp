// recorder.hpp -- the spike files of a recording environment (recorder.cpp).
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include <vector>

#include "device_types.hpp"

namespace sc {

struct Recorder;

// Opens <dir>/spike_dev0_0.dat, spike_dev0_1.dat, spike_host0.dat and
// spike_host1.dat (truncated) and starts the writer thread.  `pops` give the
// bits of each population in a ring row of Wr words (R rows), `pop_type` the
// NeuronType printed for each population.
int recorder_create(Recorder **out, const char *dir, int T, int T_print, int R, int Wr,
                    const std::vector<ScPop> &pops, const std::vector<int32_t> &pop_type);
// Hands over the window of T steps ending at step `end` (its rows are complete
// once the work queued so far on `compute` is done).
int recorder_window(Recorder *r, hipStream_t compute, const uint64_t *ring, int64_t end);
// Orders `compute` after the copies of windows whose rows steps before `seg_end`
// would change.
int recorder_before(Recorder *r, hipStream_t compute, int64_t seg_end);
// Orders `compute` after every copy handed over.
int recorder_wait_all(Recorder *r, hipStream_t compute);
// The step count jumped: waits for the writer and drops an incomplete T_print window.
int recorder_restart(Recorder *r);
// ResetNetworkEnv: waits for the writer, merges the files into spike_dev0.dat and
// spike_result_trial<trial>.dat, reopens the four files empty.
int recorder_merge(Recorder *r, int trial);
// Waits for the writer; reports a failed write.
int recorder_finish(Recorder *r);
void recorder_destroy(Recorder *r);

}  // namespace sc
