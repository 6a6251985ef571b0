// Fused execution of one KernelBundle as a single kernel: the program format
// (fused_types.hpp), the compiled interpreter kernel (fused.hip) and the
// kernels compiled at run time for one program shape (jit.cpp).
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "fused_types.hpp"

namespace sdb {

// Launches the program over rows [0, n) of the segment the pointers address.
void launch_fused(const FProgram &p, int64_t n, hipStream_t s);


// 4-row chunks per thread per loop iteration of a compiled kernel for
// launches too small to tune.
constexpr int F_DEFAULT_UNROLL = 2;

// Runtime-compiled fused kernels.  jit_available() is false when the hipRTC
// library cannot be loaded (the reason is in jit_unavailable_reason()).
bool jit_available();
std::string jit_unavailable_reason();
// How one fused launch runs: on the interpreter kernel (while the shape's
// compiled kernels are still being built), on an explicitly configured
// compiled kernel, on the autotuner's current choice, or as one per-launch
// tuning trial.
struct FusedLaunch {
    enum Mode { INTERPRET, EXPLICIT, TUNED, TRIAL } mode = INTERPRET;
    int cand = 0;            // TUNED / TRIAL: autotuner candidate
    int unroll = 0, block = 256; // EXPLICIT
    bool queue = false;      // queued row tail
    std::pair<int, std::string> key; // TRIAL: the tuning entry it reports to
};
// Decides how the program's next launch over n rows runs, doing first
// whatever that needs: queuing the shape's kernels for the background
// compiler, loading compiled kernels, and the autotuner's one-launch trials
// (on a scratch copy of the program's accumulators, ordered on `s` after the
// work that produces the program's inputs).  `unroll` = 4-row chunks per
// thread per iteration (1, 2 or 4, optionally | SDB_UNROLL_QUEUED: compiled
// now if needed), or 0 for the autotuner.  wait = true never returns
// INTERPRET: it waits for the background compiler instead.
FusedLaunch prepare_fused_jit(const FProgram &p, int64_t n, hipStream_t s, int unroll, bool wait);
// Launches the program as prepared (a TRIAL also times itself and reports).
void run_fused_jit(const FusedLaunch &L, const FProgram &p, int64_t n, hipStream_t s);
// prepare_fused_jit (without waiting) followed by run_fused_jit.
void launch_fused_jit(const FProgram &p, int64_t n, hipStream_t s, int unroll);
// Waits until the background compiler has built every kernel queued so far
// (true), or until timeout_s seconds have passed (false).
bool jit_wait_idle(double timeout_s);
// Stops the background compiler (queued compilations dropped, the one in
// progress finished); call before the process exits while kernels may still
// be compiling.
void jit_shutdown();
// Program shapes whose kernels are still compiling, or whose launches over
// at least the tuning size have not yet settled on a configuration, plus
// batched-kernel libraries still tuning.
int jit_pending();
// One line per tuned program shape: the chosen configuration and the best
// measured ns/row of every candidate.
std::string jit_tuning_report();
// Generated source for the program shape (exposed for tests and debugging).
std::string jit_source(const FProgram &p, int unroll, int block = 256, bool queue = true);
// Number of distinct kernels compiled so far in this process.
int jit_kernels_compiled();
// Device memory for one batched launch (mask tables and their descriptors),
// valid until the launch has completed.
typedef void *(*BatchAlloc)(void *ctx, size_t bytes);

// Runs the fact-table programs of several pipelines over rows [0, n) in one
// launch (one scan of their columns).  unroll: 1, 2 or 4 (0: default).
// scope: identity of the fact table, whose batches share one library of
// compiled shapes.  Probes on fact key columns that every row of a program
// must pass are evaluated once per row for all programs through shared mask
// tables (FMaskCol); masks = false keeps every probe in its program.  The
// launch's device block (arguments, programs, mask tables) comes from
// `alloc`, or, with batched graphs on, from a block kept per library and
// launch shape whose launch is a replayed hipGraph.
void launch_fused_batch(const std::vector<FProgram> &in, int64_t n, hipStream_t s, int unroll,
                        const void *scope, BatchAlloc alloc, void *alloc_ctx, bool masks = true);
// Batched launches as replayed hipGraphs (1), as plain stream operations
// (0), or as SDB_BATCH_GRAPH says (-1, the default: off).
void set_batch_graphs(int on);
// Graph captures and replays of batched launches so far.
void batch_graph_stats(int64_t *captures, int64_t *replays);
// Whether batched launches on the calling thread may run autotuner trials
// (default on).  Off, a batch takes the choice already tuned for its kind,
// or the default shape: the server's workers run so, keeping trials off
// their requests' latency.
void set_batch_trials(bool on);
// Waves per CU a batched launch on the calling thread may take (0, the
// default: as many as fit).  The server's per-engine workers leave a share
// of every CU free, so that one engine's kernels run beside another's
// batched scan instead of behind it.
void set_batch_waves_per_cu(int cap);
// Builds the mask tables of a batched launch (ops.hip).
void launch_batch_masks(const FMaskCol *cols, int ncols, uint32_t max_len, const FMaskEntry *ents, hipStream_t s);
// Builds the rank dictionaries of some of them, after the tables (ops.hip).
void launch_batch_dict(const FMaskDict *dicts, int ndicts, hipStream_t s);
// Compiles (no device needed) a batched kernel over the coverage programs.
int jit_batch_check(const std::string &arch);
// Compiles (without launching; no device needed) the kernels of programs
// covering every instruction form for `arch`; returns how many compiled.
int jit_check(const std::string &arch);

} // namespace sdb
