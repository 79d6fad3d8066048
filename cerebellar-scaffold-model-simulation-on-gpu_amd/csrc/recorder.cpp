// recorder.cpp -- spike files written while the simulation runs, as the
// reference's printer threads and host threads write them in its PRINT mode
// (GPUScaffoldSimulation.cu:7-89) and merge them at the end of a trial
// (:818-827).
//
// After every step that completes a window of T_cpu_sim steps the engine calls
// recorder_window(): the window's ring rows are copied into a pinned host slot
// by an asynchronous copy on a stream of the recorder's own, ordered after the
// steps by an event, so the next steps run while it is under way.  A writer
// thread waits for each copy and formats the window's lines:
//   * device populations (every population but the OUTPUT ones), as
//     print_spike_train writes them: "%f\t%d\t%d\n" = (step as ms,
//     population-local id, NeuronType), into spike_dev0_<k>.dat, where printer
//     thread k = (window index / (T_print / T_cpu_sim)) % 2 -- each thread prints
//     whole T_print windows, so a T_print window is written once complete;
//   * OUTPUT populations, as host_update_lif_ch writes them (Simulation_header.cu
//     :456): (step as ms, global id, NeuronType), into spike_host<k>.dat, every
//     T_cpu_sim window.
// recorder_merge() (ResetNetworkEnv) waits for the writer, closes the files and
// concatenates them as the reference's `cat` calls do: spike_dev0.dat =
// spike_dev0_0.dat + spike_dev0_1.dat, spike_result_trial<trial>.dat =
// spike_dev0.dat + spike_host0.dat + spike_host1.dat; then the four files start
// empty for the next trial.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdint>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "model.hpp"
#include "recorder.hpp"

namespace sc {

#define REC_HIP(x)                                                                                   \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return fail(SCAFFOLD_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_));    \
    } while (0)

struct Recorder {
    std::string dir;
    FILE *dev[2] = {nullptr, nullptr}, *host[2] = {nullptr, nullptr};
    int T = 1, T_print = 1, R = 1, Wr = 1;
    hipStream_t copy = nullptr;
    int Q = 0;                              // pinned slots of T rows each
    std::vector<uint64_t *> slot;
    std::vector<hipEvent_t> ready, copied;  // per slot: steps done (compute stream), rows copied
    std::vector<int64_t> slot_end;          // window end held by a slot (-1: free)
    std::deque<int> queue;                  // slots to write, in window order
    std::deque<int64_t> pending;            // windows copied, not yet waited for by the compute stream
    std::mutex mu;
    std::condition_variable cv_work, cv_free, cv_idle;
    bool stop = false, busy = false;
    std::thread writer;
    std::string devbuf;                     // lines of the current T_print window
    // formatting tables, per population: first 64-neuron group, size, kind,
    // global id base, printed NeuronType
    std::vector<int32_t> pop_group, pop_num, pop_kind, pop_ref, pop_type;
    int failed = 0;                         // a file write failed
};

// "%f" of the float step and the two integers, as the reference prints them.
static void put_line(std::string &out, int64_t step, int id, int type)
{
    char b[96];
    const float t = (float)step;
    int n;
    if (step >= 0 && step < (1 << 24))  // exact in float: an integer with six zero decimals
        n = snprintf(b, sizeof b, "%lld.000000\t%d\t%d\n", (long long)step, id, type);
    else
        n = snprintf(b, sizeof b, "%f\t%d\t%d\n", (double)t, id, type);
    out.append(b, (size_t)n);
}

static bool write_all(FILE *f, const std::string &s)
{
    return s.empty() || fwrite(s.data(), 1, s.size(), f) == s.size();
}

// Formats the window held by slot `s` (rows of steps [end - T, end)).
static void write_window(Recorder *r, int s, int64_t end)
{
    const uint64_t *rows = r->slot[(size_t)s];
    const int64_t h = end / r->T - 1;                 // hand-off index in the trial
    const int k = (int)((h / std::max(1, r->T_print / r->T)) % 2);
    if (!r->host[k] || !r->dev[k]) {  // a merge that failed left the files closed
        r->failed = 1;
        return;
    }
    std::string hostbuf;
    const int np = (int)r->pop_kind.size();
    for (int i = 0; i < r->T; ++i) {
        const int64_t step = end - r->T + i;
        const uint64_t *w = rows + (size_t)i * r->Wr;
        for (int p = 0; p < np; ++p) {  // populations in index order, neurons in id order
            const bool out = r->pop_kind[p] == SCAFFOLD_OUTPUT;
            std::string &buf = out ? hostbuf : r->devbuf;
            const int g1 = r->pop_group[p] + (r->pop_num[p] + 63) / 64;
            for (int g = r->pop_group[p]; g < g1; ++g)
                for (uint64_t word = w[g]; word; word &= word - 1) {
                    const int local = (g - r->pop_group[p]) * 64 + __builtin_ctzll(word);
                    if (local >= r->pop_num[p]) break;  // padding lanes of the last group
                    put_line(buf, step, out ? r->pop_ref[p] + local : local, r->pop_type[p]);
                }
        }
    }
    bool ok = write_all(r->host[k], hostbuf);
    if (end % r->T_print == 0) {                      // a printer thread prints whole T_print windows
        ok = write_all(r->dev[k], r->devbuf) && ok;
        r->devbuf.clear();
    }
    if (!ok) r->failed = 1;
}

static void writer_loop(Recorder *r)
{
    std::unique_lock<std::mutex> lock(r->mu);
    for (;;) {
        r->cv_work.wait(lock, [&] { return r->stop || !r->queue.empty(); });
        if (r->queue.empty()) {
            if (r->stop) return;
            continue;
        }
        const int s = r->queue.front();
        r->queue.pop_front();
        r->busy = true;
        const int64_t end = r->slot_end[(size_t)s];
        lock.unlock();
        if (hipEventSynchronize(r->copied[(size_t)s]) == hipSuccess) write_window(r, s, end);
        else r->failed = 1;  // reported by the next merge, restart or finish
        lock.lock();
        r->slot_end[(size_t)s] = -1;
        r->busy = false;
        r->cv_free.notify_all();
        if (r->queue.empty()) r->cv_idle.notify_all();
    }
}

static int open_files(Recorder *r)
{
    const char *names[4] = {"spike_dev0_0.dat", "spike_dev0_1.dat", "spike_host0.dat", "spike_host1.dat"};
    FILE **fp[4] = {&r->dev[0], &r->dev[1], &r->host[0], &r->host[1]};
    for (int i = 0; i < 4; ++i) {
        const std::string path = r->dir + "/" + names[i];
        *fp[i] = fopen(path.c_str(), "w");
        if (!*fp[i]) return fail(SCAFFOLD_ERR_IO, "cannot open %s for writing", path.c_str());
    }
    return SCAFFOLD_OK;
}

static void close_files(Recorder *r)
{
    for (FILE **f : {&r->dev[0], &r->dev[1], &r->host[0], &r->host[1]})
        if (*f) {
            fclose(*f);
            *f = nullptr;
        }
}

// Waits until the writer has written every window handed to it.
static void drain(Recorder *r)
{
    std::unique_lock<std::mutex> lock(r->mu);
    r->cv_idle.wait(lock, [&] { return r->queue.empty() && !r->busy; });
}

void recorder_destroy(Recorder *r)
{
    if (!r) return;
    if (r->writer.joinable()) {
        drain(r);
        {
            std::lock_guard<std::mutex> lock(r->mu);
            r->stop = true;
        }
        r->cv_work.notify_all();
        r->writer.join();
    }
    close_files(r);
    if (r->copy) (void)hipStreamSynchronize(r->copy);
    for (hipEvent_t e : r->ready) (void)hipEventDestroy(e);
    for (hipEvent_t e : r->copied) (void)hipEventDestroy(e);
    for (uint64_t *p : r->slot) (void)hipHostFree(p);
    if (r->copy) (void)hipStreamDestroy(r->copy);
    delete r;
}

int recorder_create(Recorder **out, const char *dir, int T, int T_print, int R, int Wr,
                    const std::vector<ScPop> &pops, const std::vector<int32_t> &pop_type)
{
    *out = nullptr;
    Recorder *r = new Recorder();
    auto bail = [&](int rc) { recorder_destroy(r); return rc; };
    r->dir = dir;
    r->T = T;
    r->T_print = T_print;
    r->R = R;
    r->Wr = Wr;
    for (size_t p = 0; p < pops.size(); ++p) {
        if (pops[p].group_base < 0 || (int64_t)pops[p].group_base * 64 + pops[p].num > (int64_t)Wr * 64)
            return bail(fail(SCAFFOLD_ERR_STATE, "population %zu outside the spike ring", p));
        r->pop_group.push_back(pops[p].group_base);
        r->pop_num.push_back(pops[p].num);
        r->pop_kind.push_back(pops[p].kind);
        r->pop_ref.push_back(pops[p].ref_base);
        r->pop_type.push_back(pop_type[p]);
    }
    int rc = open_files(r);
    if (rc) return bail(rc);
    if (hipStreamCreateWithFlags(&r->copy, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(SCAFFOLD_ERR_HIP, "hipStreamCreate failed"));
    r->Q = std::max(4, T_print / T + 2);
    for (int s = 0; s < r->Q; ++s) {
        uint64_t *p = nullptr;
        if (hipHostMalloc(reinterpret_cast<void **>(&p), sizeof(uint64_t) * (size_t)T * Wr, hipHostMallocDefault) !=
            hipSuccess)
            return bail(fail(SCAFFOLD_ERR_HIP, "hipHostMalloc of a recording slot failed"));
        r->slot.push_back(p);
        hipEvent_t a, b;
        if (hipEventCreateWithFlags(&a, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&b, hipEventDisableTiming) != hipSuccess)
            return bail(fail(SCAFFOLD_ERR_HIP, "hipEventCreate failed"));
        r->ready.push_back(a);
        r->copied.push_back(b);
        r->slot_end.push_back(-1);
    }
    r->writer = std::thread(writer_loop, r);
    *out = r;
    return SCAFFOLD_OK;
}

int recorder_window(Recorder *r, hipStream_t compute, const uint64_t *ring, int64_t end)
{
    const int s = (int)((end / r->T) % r->Q);
    {
        std::unique_lock<std::mutex> lock(r->mu);  // the slot's previous window is written
        r->cv_free.wait(lock, [&] { return r->slot_end[(size_t)s] < 0; });
    }
    REC_HIP(hipEventRecord(r->ready[(size_t)s], compute));
    REC_HIP(hipStreamWaitEvent(r->copy, r->ready[(size_t)s], 0));
    const int64_t first = end - r->T;
    const int r0 = (int)(((first % r->R) + r->R) % r->R);
    const int n0 = std::min(r->T, r->R - r0);
    const size_t row_bytes = sizeof(uint64_t) * (size_t)r->Wr;
    REC_HIP(hipMemcpyAsync(r->slot[(size_t)s], ring + (size_t)r0 * r->Wr, row_bytes * n0, hipMemcpyDeviceToHost,
                           r->copy));
    if (n0 < r->T)  // the window wraps around the ring
        REC_HIP(hipMemcpyAsync(r->slot[(size_t)s] + (size_t)n0 * r->Wr, ring, row_bytes * (r->T - n0),
                               hipMemcpyDeviceToHost, r->copy));
    REC_HIP(hipEventRecord(r->copied[(size_t)s], r->copy));
    {
        std::lock_guard<std::mutex> lock(r->mu);
        r->slot_end[(size_t)s] = end;
        r->queue.push_back(s);
        r->pending.push_back(end);
    }
    r->cv_work.notify_all();
    return SCAFFOLD_OK;
}

// The compute stream waits for the copy of every window whose rows the steps up
// to `seg_end` would change: the oldest row of window [e - T, e) is first
// changed by step e - T + R - 1 (each step clears the row after its own).
int recorder_before(Recorder *r, hipStream_t compute, int64_t seg_end)
{
    std::lock_guard<std::mutex> lock(r->mu);
    while (!r->pending.empty() && r->pending.front() - r->T + r->R - 1 < seg_end) {
        const int64_t e = r->pending.front();
        r->pending.pop_front();
        const int s = (int)((e / r->T) % r->Q);
        if (r->slot_end[(size_t)s] == e) REC_HIP(hipStreamWaitEvent(compute, r->copied[(size_t)s], 0));
    }
    return SCAFFOLD_OK;
}

int recorder_wait_all(Recorder *r, hipStream_t compute)
{
    return recorder_before(r, compute, INT64_MAX / 2);
}

static int append_file(FILE *dst, const std::string &path)
{
    FILE *f = fopen(path.c_str(), "r");
    if (!f) return fail(SCAFFOLD_ERR_IO, "cannot open %s", path.c_str());
    char buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0)
        if (fwrite(buf, 1, n, dst) != n) {
            fclose(f);
            return fail(SCAFFOLD_ERR_IO, "write error while merging spike files");
        }
    fclose(f);
    return SCAFFOLD_OK;
}

int recorder_merge(Recorder *r, int trial)
{
    drain(r);
    r->pending.clear();
    r->devbuf.clear();  // an incomplete T_print window is not printed (the printer threads print whole windows)
    close_files(r);
    if (r->failed) return fail(SCAFFOLD_ERR_IO, "writing the spike files failed");
    const std::string d = r->dir + "/";
    int rc = SCAFFOLD_OK;
    {
        FILE *f = fopen((d + "spike_dev0.dat").c_str(), "w");
        if (!f) return fail(SCAFFOLD_ERR_IO, "cannot open %sspike_dev0.dat", d.c_str());
        rc = append_file(f, d + "spike_dev0_0.dat");
        if (!rc) rc = append_file(f, d + "spike_dev0_1.dat");
        fclose(f);
    }
    if (!rc) {
        char name[64];
        snprintf(name, sizeof name, "spike_result_trial%d.dat", trial);
        FILE *f = fopen((d + name).c_str(), "w");
        if (!f) return fail(SCAFFOLD_ERR_IO, "cannot open %s%s", d.c_str(), name);
        rc = append_file(f, d + "spike_dev0.dat");
        if (!rc) rc = append_file(f, d + "spike_host0.dat");
        if (!rc) rc = append_file(f, d + "spike_host1.dat");
        fclose(f);
    }
    if (rc) return rc;
    return open_files(r);
}

int recorder_restart(Recorder *r)
{
    drain(r);
    r->pending.clear();
    r->devbuf.clear();
    return r->failed ? fail(SCAFFOLD_ERR_IO, "writing the spike files failed") : SCAFFOLD_OK;
}

int recorder_finish(Recorder *r)
{
    drain(r);
    return r->failed ? fail(SCAFFOLD_ERR_IO, "writing the spike files failed") : SCAFFOLD_OK;
}

}  // namespace sc
