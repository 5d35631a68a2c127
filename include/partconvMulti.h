/*
 * partconvMulti.h -- C++ drop-in for the reference's PartConvMulti and
 * PartConvMultiRelaxed (src/partconvMulti.h:52-148), header-only over the C
 * ABI of partconv_amd.h.  Like the reference's header it includes partconv.h
 * (PartConvFilter, PartConvMax, and Vector through buffers.h).
 *
 * Same class names, method names, argument meaning and public fields.  Where
 * the reference calls exit(-1) (bad config, unreadable IR, frameCount !=
 * buffer_size), these methods return the nonzero status instead and
 * pc_last_error() has the message.
 *
 *   #include "partconvMulti.h"
 *   PartConvMulti pc(1);
 *   pc.setup("pc.config");
 *   ... in the audio callback: pc.run(out, in, frameCount);
 *   pc.cleanup();
 *
 * Link with -lpartconv_amd (libpartconv_amd.so, built by partconv_amd/build.py).
 */
#ifndef PARTCONV_AMD_PARTCONVMULTI_H
#define PARTCONV_AMD_PARTCONVMULTI_H

#include <stdio.h>

#include "partconv.h"
#include "partconv_amd.h"

class Waiter; /* the reference's completion waiter type (src/partconvMulti.h:10) */

class PartConvMulti
{
public:
    PartConvMulti(int inVerbosity = 0) : verbosity(inVerbosity), h_(pcm_create(inVerbosity, 0)) { clear(); }
    virtual ~PartConvMulti() { pcm_destroy(h_); }

    int setup(const char *config_file_name, int max_threads_per_level = 1, int max_level0_threads = 0,
              double min_work = 0)
    {
        int rc = pcm_setup(h_, config_file_name, max_threads_per_level, max_level0_threads, min_work);
        if (rc) fprintf(stderr, "%s\n", pc_last_error());
        refresh();
        return rc;
    }
    int cleanup() { return pcm_cleanup(h_); }
    /* one tick; frameNum counts the ticks that completed */
    virtual int run(float *const output, const float *const input, const int frameCount)
    {
        int rc = pcm_run(h_, output, input, frameCount);
        if (rc) fprintf(stderr, "%s\n", pc_last_error());
        else frameNum++;
        return rc;
    }
    void reset()
    {
        pcm_reset(h_);
        frameNum = 0;
    }
    /* engine layout: PC_BATCHED (default) or PC_PLAIN, before setup */
    int set_mode(int mode) { return pcm_set_mode(h_, mode); }
    pc_multi *handle() { return h_; }

    float *output_mix = nullptr; /* kept for source compatibility; mixing happens on the GPU */
    float *input_mix = nullptr;
    int FS;
    int buffer_size;
    int max_block_size;
    int verbosity;
    int numInputChannels;
    int numOutputChannels;
    unsigned long frameNum;
    Waiter *doneWaiter = nullptr; /* no waiter: run() returns when the tick's output is ready */
    unsigned long lastFrame = 0;

protected:
    PartConvMulti(int inVerbosity, int relaxed) : verbosity(inVerbosity), h_(pcm_create(inVerbosity, relaxed))
    {
        clear();
    }
    void clear()
    {
        FS = buffer_size = max_block_size = numInputChannels = numOutputChannels = 0;
        frameNum = 0;
    }
    void refresh()
    {
        pc_multi_info i;
        if (pcm_get_info(h_, &i) == PC_OK) {
            FS = i.FS;
            buffer_size = i.buffer_size;
            max_block_size = i.max_block_size;
            numInputChannels = i.numInputChannels;
            numOutputChannels = i.numOutputChannels;
        }
    }
    pc_multi *h_;
};

class PartConvMultiRelaxed : public PartConvMulti
{
public:
    PartConvMultiRelaxed(int inVerbosity = 0) : PartConvMulti(inVerbosity, 1) {}
};

/* Real-time MIMO, shared input transforms (no reference counterpart: the
 * reference runs an n_in x n_out matrix as n_in * n_out PartConvMulti
 * instances, instance (o, i) with inCh = i, outCh = o, summed into the
 * outputs).  Each input is transformed once and has one delay line; frames
 * are device-resident and interleaved as run()'s buffers.  Filters: device
 * float32 [n_out][n_in][taps]. */
class PartConvMimo
{
public:
    PartConvMimo() = default;
    PartConvMimo(const PartConvMimo &) = delete;
    PartConvMimo &operator=(const PartConvMimo &) = delete;
    ~PartConvMimo() { cleanup(); }

    int setup(int inSampleRate, int buffer_size, int M, const int *N, const int *P, int taps, int n_in, int n_out)
    {
        cleanup();
        h_ = pcxr_create(inSampleRate, buffer_size, M, N, P, taps, n_in, n_out);
        if (!h_) {
            fprintf(stderr, "%s\n", pc_last_error());
            return 1;
        }
        return 0;
    }
    int set_filters(const float *d_h) { return report(pcxr_set_filters(h_, d_h)); }
    /* nframes ticks, each released by the host and waited for */
    int run_device(float *d_out, const float *d_in, long long nframes)
    {
        return report(pcxr_run_device(h_, d_out, d_in, nframes));
    }
    /* ticks run_device may release ahead of the oldest incomplete one (1 .. 3) */
    int set_release_window(int ticks) { return report(pcxr_set_release_window(h_, ticks)); }
    /* one tick on host buffers (interleaved, frameCount == buffer_size) */
    int run(float *const output, const float *const input, const int frameCount)
    {
        return report(pcxr_run(h_, output, input, frameCount));
    }
    int reset() { return report(pcxr_reset(h_)); }
    int stats(pc_stats *st) const { return pcxr_get_stats(h_, st); }
    int cleanup()
    {
        if (h_) pcxr_destroy(h_);
        h_ = nullptr;
        return 0;
    }
    pc_mimo_rt *handle() { return h_; }

private:
    static int report(int rc)
    {
        if (rc) fprintf(stderr, "%s\n", pc_last_error());
        return rc;
    }
    pc_mimo_rt *h_ = nullptr;
};

#endif /* PARTCONV_AMD_PARTCONVMULTI_H */
