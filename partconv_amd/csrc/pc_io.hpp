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
bool read_wav(const std::string &path, std::vector<float> &out, int *sample_rate, std::string &err);

}  // namespace pcio
