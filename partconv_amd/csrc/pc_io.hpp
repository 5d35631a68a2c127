// pc_io.hpp -- config grammar and WAV reading (host only).
#pragma once
#include <string>
#include <vector>

namespace pcio {

struct ConfigLine {
    std::string ir_file;
    int mute = 0, in_ch = 0, out_ch = 0;
    float scale = 1.f;
    std::vector<int> N, P;
};

struct Config {
    int FS = 0, buffer_size = 0, nIn = 0, nOut = 0, numPCs = 0;
    std::vector<ConfigLine> lines;  // after repeats
};

bool parse_config(const std::string &path, Config &cfg, std::string &err);
struct WavInfo {
    int format_tag = 0;  // 1 integer PCM, 3 IEEE float (WAVE_FORMAT_EXTENSIBLE resolved to its subformat)
    int channels = 0, bits = 0, sample_rate = 0;
};

// First channel of a RIFF/WAVE file as float; false with a message in err.
bool read_wav(const std::string &path, std::vector<float> &out, WavInfo *info, std::string &err);

}  // namespace pcio
