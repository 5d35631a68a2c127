/*
 * partconv.h -- C++ drop-in for the reference's PartConvFilter
 * (src/partconv.h:40-114) and PartConvMax (src/partconv.h:117-175),
 * header-only over the C ABI of partconv_amd.h.
 *
 * Same class names, method signatures (including PartConvFilter::setup's
 * Vector* impulse, src/partconv.h:48-53), argument meaning and public fields.
 * Where the reference calls exit(-1), these methods return the nonzero status
 * instead and pc_last_error() has the message.
 *
 * Link with -lpartconv_amd (libpartconv_amd.so, built by partconv_amd/build.py).
 */
#ifndef PARTCONV_AMD_PARTCONV_H
#define PARTCONV_AMD_PARTCONV_H

#include <stdio.h>

#include "buffers.h"
#include "partconv_amd.h"

class PartConvFilter
{
public:
    PartConvFilter(int inVerbosity = 0, int inSampleRate = 44100)
        : M(0), N(nullptr), num_parts(nullptr), FS(inSampleRate), verbosity(inVerbosity),
          h_(pcf_create(inVerbosity, inSampleRate))
    {
    }
    ~PartConvFilter()
    {
        pcf_destroy(h_);
        delete[] N;
        delete[] num_parts;
    }
    PartConvFilter(const PartConvFilter &) = delete;
    PartConvFilter &operator=(const PartConvFilter &) = delete;

    /* PartConvFilter::setup (src/partconv.cpp:70-306): impulse == NULL gives an
     * all-ones IR covering the partition.  The caller's arrays are read, never
     * written; the members N and num_parts hold the partition after the last
     * count is recounted from the IR.  fftwOptLevel and wisdom have no
     * counterpart (no FFT planner) and are ignored.  0 on success, 1 on error. */
    int setup(int M_in, int *N_in, int *num_parts_in, Vector *impulse = NULL, int fftwOptLevel = 1,
              const char *wisdom = NULL)
    {
        return setup_ir(M_in, N_in, num_parts_in, impulse ? (*impulse)() : nullptr,
                        impulse ? impulse->getSize() : 0, fftwOptLevel, wisdom);
    }
    /* The same with the IR as a plain array of impulse_len samples. */
    int setup_ir(int M_in, const int *N_in, const int *num_parts_in, const float *impulse, int impulse_len,
                 int fftwOptLevel = 1, const char *wisdom = NULL)
    {
        pcf_set_fs(h_, FS);
        int rc = pcf_setup(h_, M_in, N_in, num_parts_in, impulse, impulse_len, fftwOptLevel, wisdom);
        if (rc) {
            fprintf(stderr, "%s\n", pc_last_error());
            return 1;
        }
        delete[] N;
        delete[] num_parts;
        M = M_in;
        N = new int[M];
        num_parts = new int[M];
        pcf_get_levels(h_, nullptr, N, num_parts, M);
        return 0;
    }
    double bench(const int L, const int trials = 100, const int pollute_every = 0, const float max_time = 0.1,
                 double *benchData = NULL)
    {
        return pcf_bench(h_, L, trials, pollute_every, max_time, benchData);
    }
    void reset() { pcf_reset(h_); }
    void sync_levels() { pcf_sync_levels(h_); }
    void sync_levels_relaxed() { pcf_sync_levels_relaxed(h_); }
    void input(const float *const input) { pcf_input(h_, input); }
    void output(float *const output) { pcf_output(h_, output); }
    void output_relaxed(float *const output) { pcf_output_relaxed(h_, output); }
    int runLevel(int L) { return pcf_run_level(h_, L); }
    int cleanup() { return pcf_cleanup(h_); }

    int M;
    int *N;
    int *num_parts;
    int FS;
    int verbosity;
    float *output_mix = nullptr; /* kept for source compatibility */
    float *input_mix = nullptr;

private:
    pc_filter *h_;
};

/* PartConvMax (src/partconv.h:117-175): the in-memory front end.  Channel
 * modes from (n, m, k), IRs from one array, one greedy partition over
 * blockSizes, per-channel buffers (float** run). */
class PartConvMax
{
public:
    PartConvMax(int inVerbosity = 0) : verbosity(inVerbosity), h_(pcm_create(inVerbosity, 0)) {}
    ~PartConvMax() { pcm_destroy(h_); }
    PartConvMax(const PartConvMax &) = delete;
    PartConvMax &operator=(const PartConvMax &) = delete;

    int setup(int inSampleRate, int n, int m, int k, float *impulses, int filterLength, int stride, int *blockSizes,
              int levels, int fftwOptLevel = 1, const char *wisdom = NULL, int max_threads_per_level = 1,
              int max_level0_threads = 0)
    {
        (void)fftwOptLevel;
        (void)wisdom;
        (void)max_threads_per_level;
        (void)max_level0_threads;
        int rc = pcm_setup_max(h_, inSampleRate, n, m, k, impulses, filterLength, stride, blockSizes, levels);
        if (rc) {
            fprintf(stderr, "%s\n", pc_last_error());
            return 1;
        }
        pc_multi_info i;
        if (pcm_get_info(h_, &i) == PC_OK) {
            FS = i.FS;
            buffer_size = i.buffer_size;
            max_block_size = i.max_block_size;
            numInputChannels = i.numInputChannels;
            numOutputChannels = i.numOutputChannels;
        }
        return 0;
    }
    int cleanup() { return pcm_cleanup(h_); }
    /* one tick; frameNum counts the ticks that completed */
    int run(float **const output, float **const input)
    {
        int rc = pcm_run_channels(h_, output, input);
        if (rc) fprintf(stderr, "%s\n", pc_last_error());
        else frameNum++;
        return rc;
    }
    void reset()
    {
        pcm_reset(h_);
        frameNum = 0;
    }
    pc_multi *handle() { return h_; }

    float **outbuffers = nullptr; /* kept for source compatibility */
    float **inbuffers = nullptr;
    int FS = 0;
    int buffer_size = 0;
    int max_block_size = 0;
    int verbosity;
    int numInputChannels = 0;
    int numOutputChannels = 0;
    unsigned long frameNum = 0;

private:
    pc_multi *h_;
};

#endif /* PARTCONV_AMD_PARTCONV_H */
