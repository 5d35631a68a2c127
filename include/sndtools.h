/*
 * sndtools.h -- readImpulseResponse(Vector&, filename, verbosity)
 * (src/sndtools.h:16, src/sndtools.cpp:3-47) over the library's RIFF/WAVE
 * reader (pc_read_ir): the first channel of the file as float, integer PCM
 * scaled by 2^(bits-1).  Returns 0, or 1 with a message on stderr.
 */
#ifndef PARTCONV_AMD_SNDTOOLS_H
#define PARTCONV_AMD_SNDTOOLS_H

#include <stdio.h>

#include "buffers.h"
#include "partconv_amd.h"

inline int readImpulseResponse(Vector &v, const char *filename, int verbosity = 0)
{
    long long frames = 0;
    if (pc_read_ir(filename, nullptr, 0, &frames, nullptr, verbosity) != PC_OK || frames > 0x7fffffffll) {
        fprintf(stderr, "%s\n", pc_last_error());
        return 1;
    }
    v.create((int)frames);
    if (frames > 0 && pc_read_ir(filename, v(), frames, &frames, nullptr, 0) != PC_OK) {
        fprintf(stderr, "%s\n", pc_last_error());
        return 1;
    }
    return 0;
}

#endif /* PARTCONV_AMD_SNDTOOLS_H */
