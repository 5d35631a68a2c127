/*
 * partconv_amd.h -- C ABI of the MI355X (gfx950) partitioned-convolution engine.
 *
 * Every entry point takes plain pointers and sizes, returns an int status
 * (PC_OK == 0) unless stated, and never exits the process: where the reference
 * calls exit(-1) (config/IR/frameCount errors) these return an error code and
 * pc_last_error() holds the message.
 *
 * Two handle types mirror the reference's two public classes:
 *   pcm_*  PartConvMulti / PartConvMultiRelaxed   (src/partconvMulti.h:52-148)
 *   pcf_*  PartConvFilter                         (src/partconv.h:40-114)
 * The C++ classes in partconv.h / partconvMulti.h and the Python module
 * partconv_amd are thin layers over these functions.
 *
 * Threading: a handle is driven by one host thread at a time (as the reference's
 * run() is called serially from one audio callback).  All GPU work is internal
 * to the engine's HIP streams.
 */
#ifndef PARTCONV_AMD_H
#define PARTCONV_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

enum pc_status {
    PC_OK = 0,
    PC_ERR_ARG = 1,        /* bad argument / null handle */
    PC_ERR_CONFIG = 2,     /* config file missing or malformed */
    PC_ERR_PARTITION = 3,  /* partition rejected by the reference's rules */
    PC_ERR_IO = 4,         /* impulse-response file unreadable */
    PC_ERR_HIP = 5,        /* HIP runtime error (message has the HIP string) */
    PC_ERR_STATE = 6,      /* call out of order (run before setup, ...) */
    PC_ERR_FRAMECOUNT = 7, /* frameCount != buffer_size */
    PC_ERR_TIMEOUT = 8     /* a tick did not complete within the engine's bound */
};

/* Engine layouts for pcm_set_mode(): how instances share kernel launches. */
enum pc_mode {
    PC_BATCHED = 0, /* all instances of a level in one launch per level tick (default) */
    PC_PLAIN = 1    /* one launch per instance per level tick: the per-instance
                       runLevel granularity of the reference's work units */
};

/* Fields of PartConvMulti that the reference exposes publicly
 * (src/partconvMulti.h:64-79). */
typedef struct pc_multi_info {
    int FS;
    int buffer_size;
    int max_block_size;
    int verbosity;
    int numInputChannels;
    int numOutputChannels;
    int numPCs;           /* instances after repeats */
    int num_levels;       /* distinct block sizes (global levels) */
    unsigned long long frameNum;
} pc_multi_info;

/* Per-tick timing collected by the engine (host wall clock, submit -> output
 * visible).  Latencies in microseconds. */
typedef struct pc_stats {
    unsigned long long ticks;
    unsigned long long deadline_misses; /* ticks whose latency exceeded the block period */
    double p50_us, p99_us, max_us, mean_us;
    double period_us;                   /* buffer_size / FS */
    unsigned long long persistent_ticks; /* ticks served by the persistent tick kernel */
    unsigned long long stalls;           /* persistent runs a late level block interrupted (the
                                            engine then waited for the block and went on) */
} pc_stats;

/* One instance of an in-memory setup: the fields of one config line
 * (src/partconvMulti.cpp:73-114) with the IR passed by pointer. */
typedef struct pc_instance {
    int mute;
    int in_ch;
    int out_ch;
    float scale;
    int M;             /* number of levels */
    const int *N;      /* M block sizes, N[0] == buffer_size */
    const int *P;      /* M counts; the last one is recounted from the IR length */
    const float *ir;   /* impulse response */
    int ir_len;
} pc_instance;

typedef struct pc_multi pc_multi;
typedef struct pc_filter pc_filter;

/* Last error message of the calling thread ("" if none). */
const char *pc_last_error(void);

/* End the library's use in this process (no reference counterpart: the
 * reference frees per object in cleanup(), src/partconvMulti.cpp:720-804).
 * Destroys every handle still live -- stopping resident persistent tick
 * kernels and joining the engine's launcher threads -- returns the library's
 * host-writable device memory and twiddle tables, and pairs its hsa_init with
 * hsa_shut_down.  Idempotent.  Runs by itself at process exit (an atexit
 * handler registered after the library's first HIP call, so before the HIP
 * runtime's own teardown; the Python loader also calls it from its atexit
 * hook).  Afterwards *_destroy of a handle it destroyed is a no-op and every
 * *_create fails with PC_ERR_STATE. */
void pc_shutdown(void);

/* Library / device information: writes the HIP device name into buf. */
int pc_device_name(char *buf, int len);

/* readImpulseResponse(Vector&, filename, verbosity) (src/sndtools.cpp:3-47):
 * the first channel of a RIFF/WAVE file as float.  Integer PCM of 8 (unsigned),
 * 16, 24 or 32 bits is scaled by 2^(bits-1); IEEE float of 32 or 64 bits is
 * taken as is; WAVE_FORMAT_EXTENSIBLE resolves to its subformat.  Writes up to
 * cap samples into out (out may be null with cap 0 to query the length);
 * *frames gets the file's frame count and *sample_rate its rate (either may be
 * null).  verbosity > 0 prints the file's description as the reference does.
 * PC_ERR_IO for a missing, unsupported or malformed file (a header whose frame
 * size cannot hold one sample per channel is rejected before any sample is
 * read). */
int pc_read_ir(const char *filename, float *out, long long cap, long long *frames, int *sample_rate,
               int verbosity);

/* Greedy partition of PartConvMax::setup (src/partconv.cpp:712-728): over the
 * candidate block sizes block_sizes[0..levels), P_i = 2*N_{i+1}/N_i - 2, stopping
 * at the first level whose cumulative span reaches filter_length; the last
 * level's count is recounted from filter_length (src/partconv.cpp:169).
 * Writes *M_out levels into N and P (each with room for `levels`). */
int pc_greedy_partition(const int *block_sizes, int levels, int filter_length, int *N, int *P,
                        int *M_out);

/* ----------------------------------------------------------------------------
 * PartConvMulti
 * ------------------------------------------------------------------------- */

/* PartConvMulti(int verbosity) / PartConvMultiRelaxed(int verbosity)
 * (src/partconvMulti.h:57, :145).  relaxed != 0 selects the relaxed schedule:
 * output delayed by buffer_size samples, every level gets a full period. */
pc_multi *pcm_create(int verbosity, int relaxed);

/* Engine layout (PC_BATCHED / PC_PLAIN); call before setup.  No reference
 * counterpart: the reference has one layout (per-instance work units). */
int pcm_set_mode(pc_multi *h, int mode);

/* PartConvMulti::setup(config_file, max_threads_per_level, max_level0_threads,
 * min_work) (src/partconvMulti.cpp:23-718).  Reads the pc.config grammar and the
 * RIFF/WAVE IR files it names (first channel).  The three thread arguments are
 * accepted for API compatibility; work is distributed over HIP streams. */
int pcm_setup(pc_multi *h, const char *config_file, int max_threads_per_level,
              int max_level0_threads, double min_work);

/* In-memory setup: same semantics as a config file with FS, buffer_size, nIn,
 * nOut and numPCs instances (numRepeats = 1), IRs passed by pointer (the
 * reference's in-memory front end is PartConvMax::setup,
 * src/partconv.cpp:653-847). */
int pcm_setup_mem(pc_multi *h, int FS, int buffer_size, int nIn, int nOut, int numPCs,
                  const pc_instance *inst);

/* PartConvMax::setup(inSampleRate, n, m, k, impulses, filterLength, stride,
 * blockSizes, levels, ...) (src/partconv.cpp:653-847): the in-memory front end.
 * Channel modes: n == m == k parallel filters with distinct IRs (instance i:
 * input i -> output i); n == m, k == 1 the same IR on every channel; m == n*k
 * filterbank (IR i on input j -> output i*n + j); otherwise PC_ERR_CONFIG.
 * IR i, tap l is impulses[stride*(filterLength*i + l)].  One greedy partition
 * over blockSizes[0..levels) (P_i = 2 N_{i+1}/N_i - 2 until the IR is
 * covered, last count recounted; src/partconv.cpp:712-728) for every
 * instance; buffer_size = blockSizes[0].  FFTW options and thread counts of
 * the reference signature have no counterpart here. */
int pcm_setup_max(pc_multi *h, int FS, int n, int m, int k, const float *impulses, int filterLength,
                  int stride, const int *blockSizes, int levels);

/* PartConvMax::run(float** output, float** input) (src/partconv.cpp:1219-1303):
 * one tick on per-channel host buffers of buffer_size samples. */
int pcm_run_channels(pc_multi *h, float *const *output, const float *const *input);

/* PartConvMulti::run (src/partconvMulti.cpp:815-928) /
 * PartConvMultiRelaxed::run (src/partconvMulti.cpp:1017-1134): one tick.
 * input: frameCount x numInputChannels interleaved, output: frameCount x
 * numOutputChannels interleaved (host memory).  Blocks until output is ready.
 * The batched engine serves these calls through a persistent tick kernel that
 * stays resident between calls (frames and output through pinned host memory,
 * the call only releases the tick); it is restarted after 1 s without a call,
 * and stopped by run_device, reset, trace, cleanup. */
int pcm_run(pc_multi *h, float *output, const float *input, int frameCount);

/* nframes consecutive pcm_run calls over interleaved host buffers (frame t at
 * input + t*buffer_size*nIn, output + t*buffer_size*nOut).  Convenience, no
 * reference counterpart. */
int pcm_process(pc_multi *h, float *output, const float *input, long long nframes);

/* Host-buffer calls (pcm_run, pcm_run_channels) through a persistent session
 * (1, default) or one launch per tick (0).  In a session on a large-BAR device
 * with nIn * buffer_size <= 16384, each input frame is written into device
 * memory as tagged words (each 8-byte word a sample and its tick), which the
 * kernel polls directly: the frame is its own release.  2 selects a session
 * without them (release word, frame read from pinned host memory).  Identical
 * output in every mode.  No reference counterpart. */
int pcm_set_host_session(pc_multi *h, int on);

/* Device-resident ticks: runs nframes consecutive ticks reading frame t from
 * d_in + t*buffer_size*nIn and writing d_out + t*buffer_size*nOut (device
 * pointers, interleaved).  sync_each_tick:
 *   0  queue every tick, wait once at the end;
 *   1  one tick-kernel launch per tick, wait for each tick's output (real-time
 *      semantics, per-tick latency recorded);
 *   2  as 1, but one persistent tick kernel serves all ticks: the host
 *      releases each tick through pinned memory instead of launching it
 *      (batched engine; falls back to 1 if the kernel cannot be made resident
 *      or a level stream stalls beside it).
 * No reference counterpart (the reference only has the host callback); the
 * tick semantics are those of pcm_run. */
int pcm_run_device(pc_multi *h, float *d_out, const float *d_in, long long nframes,
                   int sync_each_tick);

/* Release window of pcm_run_device's persistent mode (sync_each_tick = 2),
 * 1 .. 16 ticks (default 1): up to this many ticks are released ahead of the
 * oldest tick not yet complete, so the host's completion poll and the next
 * release overlap the device's work on released ticks.  Applies where each
 * instance writes its own output channel (direct mix with per-workgroup
 * completion flags); else 1.  Deadlines are unchanged: a tick is released only
 * once the level blocks it reads first are complete.  Latency counts from each
 * tick's release.  Identical output for every window.  No reference
 * counterpart (the reference's run() serves one tick per call). */
int pcm_set_release_window(pc_multi *h, int ticks);

/* PartConvMulti::reset (src/partconvMulti.cpp:806-813).  Unlike the
 * reference, also rewinds the global level counters and frameNum. */
int pcm_reset(pc_multi *h);

/* PartConvMulti::cleanup (src/partconvMulti.cpp:720-804).  Idempotent. */
int pcm_cleanup(pc_multi *h);

/* Free the handle (runs cleanup if needed). */
void pcm_destroy(pc_multi *h);

int pcm_get_info(const pc_multi *h, pc_multi_info *info);
int pcm_get_stats(const pc_multi *h, pc_stats *st);
int pcm_clear_stats(pc_multi *h);
/* Copies up to cap per-tick latencies (us) of the most recent ticks; returns the
 * number written (or -1). */
long long pcm_get_latencies(const pc_multi *h, double *out_us, long long cap);
/* Instance partition after setup (last count recounted): fills N, P (cap
 * entries each); returns M or -1. */
int pcm_instance_partition(const pc_multi *h, int inst, int *N, int *P, int cap);

/* Kernel probe (instrumentation, no reference counterpart; the reference's
 * nearest is the per-stage Timer of its older engine,
 * src/partconvMulti_cuda.cpp:1679-1818): from now on, time with HIP events
 * stamped by the launch itself (hipExtLaunchKernelGGL) the delay-line
 * partial-sum kernel of global level `global_level` (0 = the tick kernel,
 * which carries level 0) each time it launches, on the stream it runs on.
 * -1 disables. */
int pcm_probe(pc_multi *h, int global_level);
/* Average probed kernel duration (us), number of launches timed, and the
 * algorithmic HBM bytes one launch moves: per job 8N(2P-1) for the partial
 * sum (P-1 delay-line rows + P-1 filter rows read, one row written); per
 * instance 8(N0+1)(2P0+1) + 12*N0 for the tick kernel. */
int pcm_probe_read(pc_multi *h, double *avg_us, long long *launches, double *bytes_per_launch);

/* Persistent-kernel trace (instrumentation, no reference counterpart): for
 * the next persistent run, up to nticks ticks (0 disables), 8 GPU clock
 * values (s_memrealtime, 100 MHz) per tick: release seen / output published
 * by workgroups 0, njobs/2 and njobs-1; completion raised (ticket paths);
 * workgroup 0 ready for the next release.  pcm_trace_read copies 8*n values
 * (n returned). */
int pcm_trace(pc_multi *h, long long nticks);
long long pcm_trace_read(pc_multi *h, unsigned long long *out, long long cap);

/* Level workers of persistent runs (tuning switch, no reference counterpart):
 * mode 1 (default): the blocks of levels with period <= 8 ticks, N <= 2048 and
 * at most 8 bins per thread of the persistent kernel run inside it -- in worker workgroups of their own (one
 * per instance and level) when those fit in half of the device, else in each
 * instance's tick workgroup after its tick ("self mode"); mode 2: self mode
 * always; mode 0: every level block is a host launch on the level's stream.
 * Output is identical in all modes.  pcm_workers_info returns the number of
 * worker workgroups and sets *self_mode (may be null). */
int pcm_set_persistent_workers(pc_multi *h, int mode);

/* Width of the level delay-line-sum kernel (tuning switch, no reference
 * counterpart): 2, 4 or 8 frequency bins per thread (default 2).  Output is
 * identical for every width.  pcm_get_tuning reads the current workers mode
 * and width (the autotuner's knobs, partconv_amd/autotune.py). */
int pcm_set_partial_width(pc_multi *h, int bins_per_thread);
int pcm_get_tuning(const pc_multi *h, int *workers_mode, int *bins_per_thread);

/* Level blocks launched as HIP graphs (no reference counterpart; on != 0):
 * each global level's block kernels (every batch, plus the next block's
 * delay-line sum) form one executable graph, built at first use; per block
 * start the tick-dependent arguments are updated in place
 * (hipGraphExecKernelNodeSetParams) and one hipGraphLaunch on the level's
 * stream replaces the 2-4 kernel launches per batch.  Applies to per-tick and
 * persistent runs (the levels launched from the host).  Output identical.
 * A probed level (pcm_probe) is not timed while graphs are on. */
int pcm_set_level_graphs(pc_multi *h, int on);
int pcm_workers_info(pc_multi *h, int *self_mode);

/* Deadline-handling test hook (no reference counterpart): the next block of
 * global level `global_level` (>= 1) launched from the host waits on its
 * stream behind a host function that sleeps `seconds` (0 < seconds <= 5), so
 * the tick that first reads the block is late.  A persistent run then ends
 * its session after the block is PC_LEVEL_LATE (0.2 s) past due and serves the
 * remaining ticks with per-tick launches that wait for the block on the
 * device: the output is unchanged, the late tick counts as a deadline miss
 * and pc_stats.stalls counts the interruption.  One shot. */
int pcm_stall_level(pc_multi *h, int global_level, double seconds);

/* Where the persistent kernel's release counter lives (no reference
 * counterpart): 1 = device memory written by the host through the PCIe BAR
 * (large-BAR systems; the kernel polls locally), 0 = pinned host memory (the
 * kernel polls across the bus), -1 = not set up. */
int pcm_release_info(const pc_multi *h);

/* ----------------------------------------------------------------------------
 * PartConvFilter (single instance, caller-driven levels)
 * ------------------------------------------------------------------------- */

/* PartConvFilter(int verbosity, int sampleRate) (src/partconv.h:46). */
pc_filter *pcf_create(int verbosity, int sample_rate);

/* PartConvFilter::setup(M, N, num_parts, impulse, fftwOptLevel, wisdom)
 * (src/partconv.cpp:70-306).  impulse == NULL gives an all-ones IR of sum N*P
 * taps (src/partconv.cpp:132-142).  N and num_parts are read, never written
 * (the reference copies them into its members, src/partconv.cpp:83-84, and
 * recounts the last level there, :169); pcf_get_levels returns the recounted
 * partition.  Partition errors are reported before any GPU call.  fftwOptLevel
 * and wisdom are accepted and ignored (no planner). */
int pcf_setup(pc_filter *h, int M, const int *N, const int *num_parts, const float *impulse,
              int impulse_len, int fftwOptLevel, const char *wisdom);

/* PartConvFilter::runLevel(L) (src/partconv.cpp:354-490): enqueues level L's
 * next block on the level's stream and returns. */
int pcf_run_level(pc_filter *h, int L);

/* PartConvFilter::sync_levels / sync_levels_relaxed (src/partconv.cpp:568-603). */
int pcf_sync_levels(pc_filter *h);
int pcf_sync_levels_relaxed(pc_filter *h);
int pcf_start_sync_depth(const pc_filter *h);
int pcf_end_sync_depth(const pc_filter *h);

/* PartConvFilter::input (src/partconv.cpp:605-613): N[0] samples (host). */
int pcf_input(pc_filter *h, const float *input);

/* PartConvFilter::output / output_relaxed (src/partconv.cpp:616-648): N[0]
 * samples (host).  Waits for the levels whose blocks are read. */
int pcf_output(pc_filter *h, float *output);
int pcf_output_relaxed(pc_filter *h, float *output);

/* PartConvFilter::reset (src/partconv.cpp:492-506). */
int pcf_reset(pc_filter *h);

/* PartConvFilter::bench(L, trials, pollute_every, max_time, benchData)
 * (src/partconv.cpp:508-566): mean seconds of one level-L block on the GPU
 * (HIP events), stopping after max_time seconds; benchData[2L], [2L+1] get the
 * percentage of the level's block period.  Resets the filter afterwards.
 * pollute_every is accepted and ignored.  Returns < 0 on error. */
double pcf_bench(pc_filter *h, int L, int trials, int pollute_every, float max_time,
                 double *benchData);

/* PartConvFilter::cleanup (src/partconv.cpp:308-352).  Idempotent. */
int pcf_cleanup(pc_filter *h);
void pcf_destroy(pc_filter *h);

/* Public members M, N[], num_parts[], FS, verbosity (src/partconv.h:66-71). */
int pcf_get_levels(const pc_filter *h, int *M, int *N, int *num_parts, int cap);
int pcf_get_fs(const pc_filter *h);
int pcf_set_fs(pc_filter *h, int fs);

/* ----------------------------------------------------------------------------
 * Offline batched partitioned convolution and its adjoint (differentiable
 * form).  No reference counterpart: the reference has no batch or gradient
 * path.  It uses the reference's partition rules (src/partconv.cpp:86-182) and
 * filter-spectrum definition (src/partconv.cpp:199-230): y[c] = conv(x[c], h)
 * truncated to T samples, with every block of every level of every clip
 * computed at once.  Device pointers; `stream` is a hipStream_t (may be null).
 * ------------------------------------------------------------------------- */
typedef struct pc_conv pc_conv;

/* Partition N[0..M), P[0..M) (last count recounted for `taps`). */
pc_conv *pcc_create(int M, const int *N, const int *P, int taps);
/* Levels after recounting; returns M or -1. */
int pcc_levels(const pc_conv *c, int *N, int *P, int cap);
/* H = r2c([0_N | taps]) / 2N per partition, from h[0..taps) on the device. */
int pcc_set_filter(pc_conv *c, const float *d_h, void *stream);
/* y[C][T] = partitioned convolution of x[C][T] with the filter. */
int pcc_forward(pc_conv *c, const float *d_x, float *d_y, int C, int T, void *stream);
/* pcc_forward without allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory -- code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  Same result as
 * pcc_forward (the same kernels in the same order, bitwise), but every level's
 * block spectra, products and transform staging live in one file-scope device
 * array of 64 MiB per device (held until the process ends) instead of the
 * handle's grown work space; nothing is allocated, freed or synchronised.
 * Limit: 3 * C * ceil(T / N_L) * N_L * 8 bytes <= 64 MiB for every level L
 * (e.g. 16 clips x 131072 samples), else PC_ERR_ARG before any GPU work.
 * No spectra are kept for pcc_backward (it recomputes them from x).  The
 * filter must be set (pcc_set_filter).  Calls that overlap in time on one
 * device share the array and race. */
int pcc_forward_noalloc(pc_conv *c, const float *d_x, float *d_y, int C, int T, void *stream);
/* Adjoint for upstream gradient g[C][T]: dx[C][T] = correlation of g with h,
 * dh[taps] = sum over clips of the correlation of g with x (either output may
 * be null). */
int pcc_backward(pc_conv *c, const float *d_x, const float *d_g, float *d_dx, float *d_dh, int C, int T,
                 void *stream);
/* Fused mean-squared-error loss (no reference counterpart: the reference has
 * no loss or adjoint; for training with the differentiable form, where a
 * separate framework loss would read y and the target and write the gradient
 * in several more passes).  From y and t ([C][T], device): *d_loss =
 * sum((y - t)^2) / (C T) (one device float, a fixed-order sum that does not
 * depend on timing) and d_g[C][T] = (2 / (C T)) (y - t), in one read of y and
 * t and one write of g. */
int pcc_mse(pc_conv *c, const float *d_y, const float *d_t, float *d_g, float *d_loss, int C, int T, void *stream);
/* pcc_backward for the upstream gradient d_g times the device scalar
 * d_scale[0] (a loss's own upstream gradient, read on the device: no host
 * synchronisation), applied inside the gradient transforms.  No reference
 * counterpart; for pcc_mse's gradient. */
int pcc_backward_scaled(pc_conv *c, const float *d_x, const float *d_g, const float *d_scale, float *d_dx,
                        float *d_dh, int C, int T, void *stream);
/* Level-parallel calls (no reference counterpart): with mode
 * PCC_LEVEL_STREAMS_ON pcc_forward and pcc_backward run each level on its own
 * stream, forked from the caller's stream with events and joined back to it,
 * so they behave as one stream for the caller; PCC_LEVEL_STREAMS_CAPTURED
 * (the default) does so only while the caller's stream is being captured
 * into a HIP graph (eager calls are faster on one stream, replayed graphs on
 * several); PCC_LEVEL_STREAMS_OFF keeps every level on the caller's stream.
 * The inverse transforms that add into y or dx run in level order, so results
 * are bitwise the same in every mode.  Each level holds its own work space. */
#define PCC_LEVEL_STREAMS_OFF 0
#define PCC_LEVEL_STREAMS_ON 1
#define PCC_LEVEL_STREAMS_CAPTURED 2
int pcc_set_level_streams(pc_conv *c, int mode);
/* Call graphs (no reference counterpart; on by default): the second
 * pcc_forward (pcc_backward) call with the same pointers, clip count, length
 * and kept-spectra state as the previous eager one captures its launches into
 * a HIP graph on a private stream and launches the graph on the caller's
 * stream; later calls with those arguments replay it (up to 4 graphs per
 * direction, most recently used kept).  A graph is replayed only while every
 * buffer, chunk size and level-stream mode it embeds is unchanged.  Calls
 * while a kernel probe is set or while the caller's stream is itself being
 * captured always launch eagerly.  Results are bitwise the same as eager
 * calls.  on = 0 turns them off and frees the graphs. */
int pcc_set_call_graphs(pc_conv *c, int on);
/* Replays and captures so far of the forward's and the backward's call graphs
 * (any pointer may be null). */
int pcc_call_graph_stats(pc_conv *c, long long *fwd_replays, long long *fwd_captures, long long *bwd_replays,
                         long long *bwd_captures);
/* Memory budget (bytes) for keeping every level's input spectra from forward
 * for backward; above it backward recomputes them from the stored input. */
int pcc_set_budget(pc_conv *c, long long bytes);
/* Number of level spectra recomputed by backward so far. */
long long pcc_recomputed_levels(const pc_conv *c);
/* Autotuner (the reference tunes partitions only, src/tune.py): choose each
 * level's chunk sizes (blocks per workgroup column) of the spectral products
 * and of the dL/dh partial sums by timing candidates with HIP events on the
 * caller's clips d_x [C][T].  Setup-time: allocates and synchronises.  The
 * forward output does not depend on the choice; dL/dh sums are grouped by it
 * (deterministic for a fixed choice).  Candidates within 5% of the fastest
 * count as equal, and then the default (32, 64) wins, else the first in
 * candidate order.  pcc_level_tuning reads the choice. */
int pcc_tune(pc_conv *c, const float *d_x, int C, int T, void *stream);
int pcc_level_tuning(const pc_conv *c, int level, int *mac_mc, int *dh_mc);
/* Apply a chunk-size choice without timing (a cached pcc_tune result):
 * powers of two in [1, 1024]. */
int pcc_set_level_tuning(pc_conv *c, int level, int mac_mc, int dh_mc);
/* Measurement hook, as pcm_probe (no reference counterpart): time with HIP
 * events stamped by the launch itself the forward spectral product (k_off_mac)
 * of level `level` each time it launches (-1 disables); read back the mean
 * duration (us), the launch count and the algorithmic bytes one launch moves
 * (every block spectrum of every clip read once, the level's filter spectra
 * once, the products written once). */
int pcc_probe(pc_conv *c, int level);
/* The same for another kernel of the level (no reference counterpart): kind
 * PCC_PROBE_PRODUCT as pcc_probe; PCC_PROBE_FORWARD_TRANSFORM the r2c of the
 * forward's input windows (every input sample read once, every spectrum
 * written once); PCC_PROBE_INVERSE_TRANSFORM the forward's c2r into y (every
 * spectrum read once, y written once, and read once for levels after the
 * first); PCC_PROBE_DH_PARTIAL the dL/dh partial sums of the backward (G and X
 * read once, the partial sums written once).  Four-step levels are timed
 * from their first kernel's start to their second's end. */
#define PCC_PROBE_PRODUCT 0
#define PCC_PROBE_FORWARD_TRANSFORM 1
#define PCC_PROBE_INVERSE_TRANSFORM 2
#define PCC_PROBE_DH_PARTIAL 3
int pcc_probe_kernel(pc_conv *c, int level, int kind);
int pcc_probe_read(pc_conv *c, double *avg_us, long long *launches, double *bytes_per_launch);
void pcc_destroy(pc_conv *c);

/* ----------------------------------------------------------------------------
 * MIMO throughput form: y[o] = sum_i conv(x[i], h[o][i]) for up to 32 x 32
 * channels over whole device-resident signals.  In the reference this is
 * n_in * n_out PartConvMulti instances summed into the outputs
 * (src/partconvMulti.cpp:906-921); here every frequency bin's complex
 * (n_out x n_in) matrix product over all output blocks runs on the fp32
 * matrix cores.  (The real-time form is pcm_* with one instance per pair.)
 * ------------------------------------------------------------------------- */
typedef struct pc_mimo pc_mimo;
pc_mimo *pcx_create(int M, const int *N, const int *P, int taps, int n_in, int n_out);
/* h: [n_out][n_in][taps] on the device */
int pcx_set_filters(pc_mimo *x, const float *d_h, void *stream);
/* x: [n_in][T], y: [n_out][T] on the device */
int pcx_forward(pc_mimo *x, const float *d_x, float *d_y, int T, void *stream);
void pcx_destroy(pc_mimo *x);

/* ----------------------------------------------------------------------------
 * Real-time MIMO engine (no reference counterpart: the reference runs an
 * n_in x n_out network as one PartConvMulti instance per pair, summed per
 * output, src/partconvMulti.cpp:815-928).  Same output as those pair
 * instances (strict semantics, zero added latency), with the input side
 * shared: one transform and one delay line per input and level, and per
 * frequency bin a complex (n_out x n_in P) matrix-vector product streaming the
 * filter spectra.  Level 0 runs in one persistent kernel per call (each tick
 * released by the host and waited for); levels >= 1 are launched per block.
 * n_in, n_out <= 32; N[0] == buffer_size in 64..1024; block sizes <= 4096;
 * partition rules as PartConvFilter::setup (the last count is recounted).
 * ------------------------------------------------------------------------- */
typedef struct pc_mimo_rt pc_mimo_rt;
pc_mimo_rt *pcxr_create(int FS, int buffer_size, int M, const int *N, const int *P, int taps, int n_in,
                        int n_out);
/* d_h: [n_out][n_in][taps] float32 on the device */
int pcxr_set_filters(pc_mimo_rt *r, const float *d_h);
/* nframes ticks; d_in / d_out: frames of buffer_size x n_in / n_out interleaved
 * samples on the device (PartConvMulti::run's layout, frame after frame) */
int pcxr_run_device(pc_mimo_rt *r, float *d_out, const float *d_in, long long nframes);
/* Release window of pcxr_run_device, 1 .. 3 ticks (default 1; at most the
 * level-0 partition count): up to this many ticks are released ahead of the
 * oldest tick whose outputs are not all complete, so the input transforms of
 * the next ticks overlap the outputs of the current one and the host's round
 * trip.  Deadlines and output unchanged; latency counts from each tick's
 * release.  No reference counterpart. */
int pcxr_set_release_window(pc_mimo_rt *r, int ticks);
/* One tick on host buffers, interleaved as PartConvMulti::run's
 * (src/partconvMulti.cpp:815-928; frameCount == buffer_size, else an error):
 * what a real-time callback calls for a MIMO network. */
int pcxr_run(pc_mimo_rt *r, float *out, const float *in, int frameCount);
int pcxr_reset(pc_mimo_rt *r);
int pcxr_get_stats(const pc_mimo_rt *r, pc_stats *st);
int pcxr_clear_stats(pc_mimo_rt *r);
/* Deadline-handling test hook, as pcm_stall_level: the next block of level
 * `level` (>= 1) waits behind a host function sleeping `seconds` (0, 5].  The
 * engine stops its persistent kernel once the block is 0.2 s late, waits for
 * the block, and serves the remaining ticks with a new persistent kernel
 * (pc_stats.stalls counts it); output unchanged, the late tick a miss. */
int pcxr_stall_level(pc_mimo_rt *r, int level, double seconds);
/* partition after the last count is recounted; returns the level count */
int pcxr_partition(const pc_mimo_rt *r, int *N, int *P, int cap);
/* Time the delay-line-sum kernel of level `level` (>= 1; -1 disables) with HIP
 * events (up to 64 launches); read back the mean duration, launch count and
 * algorithmic bytes per launch (filter rows p >= 1 of every pair, delay-line
 * rows of every input, one S row per output). */
int pcxr_probe(pc_mimo_rt *r, int level);
int pcxr_probe_read(pc_mimo_rt *r, double *avg_us, long long *launches, double *bytes_per_launch);
void pcxr_destroy(pc_mimo_rt *r);

#ifdef __cplusplus
}
#endif

#endif /* PARTCONV_AMD_H */
