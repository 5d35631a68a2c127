// pc_io.cpp -- host-side file formats: the pc.config grammar and a RIFF/WAVE
// reader for impulse responses (first channel, as the reference's
// readImpulseResponse takes via libsndfile, src/sndtools.cpp:3-47).
#include "pc_io.hpp"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

namespace pcio {

static bool read_file(const std::string &path, std::vector<unsigned char> &buf)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    f.seekg(0, std::ios::end);
    std::streamoff n = f.tellg();
    if (n < 0) return false;
    f.seekg(0, std::ios::beg);
    buf.resize((size_t)n);
    if (n > 0) f.read(reinterpret_cast<char *>(buf.data()), n);
    return (bool)f;
}

static uint32_t le32(const unsigned char *p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
static uint16_t le16(const unsigned char *p) { return (uint16_t)(p[0] | (p[1] << 8)); }

// PCM integer samples scale to [-1, 1) by 2^(bits-1), the convention libsndfile
// applies when reading integer files as float; 8-bit WAVE data is unsigned.
// The header is validated before any sample is read: the sample width must be
// a whole number of bytes the format supports, and one frame (block_align)
// must hold every channel's sample, so each read stays inside the data chunk.
bool read_wav(const std::string &path, std::vector<float> &out, WavInfo *info, std::string &err)
{
    std::vector<unsigned char> b;
    if (!read_file(path, b)) {
        err = "readImpulseResponse() ERROR: Can't open audio file: " + path;
        return false;
    }
    if (b.size() < 12 || memcmp(b.data(), "RIFF", 4) != 0 || memcmp(b.data() + 8, "WAVE", 4) != 0) {
        err = "not a RIFF/WAVE file: " + path;
        return false;
    }
    int fmt_tag = 0, channels = 0, bits = 0, block_align = 0, rate = 0;
    bool have_fmt = false;
    const unsigned char *data = nullptr;
    size_t data_len = 0;
    size_t pos = 12;
    while (pos + 8 <= b.size()) {
        const unsigned char *ck = b.data() + pos;
        uint32_t len = le32(ck + 4);
        size_t body = pos + 8;
        size_t avail = b.size() - body;
        if (memcmp(ck, "fmt ", 4) == 0 && len >= 16 && avail >= 16) {
            have_fmt = true;
            fmt_tag = le16(ck + 8);
            channels = le16(ck + 10);
            rate = (int)le32(ck + 12);
            block_align = le16(ck + 20);
            bits = le16(ck + 22);
            if (fmt_tag == 0xFFFE) {  // WAVE_FORMAT_EXTENSIBLE: the subformat GUID's first two bytes
                if (len < 40 || avail < 40) {
                    err = "truncated WAVE_FORMAT_EXTENSIBLE header in " + path;
                    return false;
                }
                fmt_tag = le16(ck + 32);
            }
        } else if (memcmp(ck, "data", 4) == 0) {
            data = ck + 8;
            data_len = (len <= avail) ? len : avail;
        }
        pos = body + (size_t)len + (len & 1);
    }
    if (!have_fmt || !data || channels <= 0) {
        err = "WAVE file without fmt/data chunk: " + path;
        return false;
    }
    const bool pcm_ok = fmt_tag == 1 && (bits == 8 || bits == 16 || bits == 24 || bits == 32);
    const bool flt_ok = fmt_tag == 3 && (bits == 32 || bits == 64);
    if (!pcm_ok && !flt_ok) {
        err = (fmt_tag == 1 || fmt_tag == 3) ? "unsupported bit depth " + std::to_string(bits) + " in " + path
                                              : "unsupported WAVE format tag in " + path;
        return false;
    }
    if ((long long)block_align < (long long)channels * (bits / 8)) {
        err = "WAVE block_align smaller than one frame of samples in " + path;
        return false;
    }
    if (info) {
        info->format_tag = fmt_tag;
        info->channels = channels;
        info->bits = bits;
        info->sample_rate = rate;
    }
    const size_t frames = data_len / (size_t)block_align;
    out.resize(frames);
    for (size_t i = 0; i < frames; i++) {
        const unsigned char *s = data + i * (size_t)block_align;  // first channel
        double v;
        if (fmt_tag == 1) {
            if (bits == 8) v = ((int)s[0] - 128) / 128.0;
            else if (bits == 16) v = (int16_t)le16(s) / 32768.0;
            else if (bits == 24) {
                int32_t x = (int32_t)((uint32_t)s[0] << 8 | (uint32_t)s[1] << 16 | (uint32_t)s[2] << 24) >> 8;
                v = x / 8388608.0;
            } else v = (int32_t)le32(s) / 2147483648.0;
        } else if (bits == 32) {
            float x;
            memcpy(&x, s, 4);
            v = x;
        } else {
            double x;
            memcpy(&x, s, 8);
            v = x;
        }
        out[i] = (float)v;
    }
    return true;
}

// Tokenised reading of the pc.config grammar (src/partconvMulti.cpp:36-134):
// positional tokens exactly as the reference's fscanf calls consume them; the
// optional last count of each line is consumed when it is an integer and then
// ignored (set to 0, recounted at setup).  Text after the numPCs records is
// ignored.  Repeats are muted copies.
bool parse_config(const std::string &path, Config &cfg, std::string &err)
{
    std::vector<unsigned char> raw;
    if (!read_file(path, raw)) {
        err = "Unable to open config file " + path;
        return false;
    }
    std::string text(raw.begin(), raw.end());
    std::istringstream ss(text);
    std::vector<std::string> tok;
    std::string t;
    while (ss >> t) tok.push_back(t);
    size_t p = 0;
    auto take_int = [&](int &v, const char *what) -> bool {
        if (p >= tok.size()) { err = std::string("Error reading config file! (") + what + ")"; return false; }
        char *end = nullptr;
        long x = strtol(tok[p].c_str(), &end, 10);
        if (end == tok[p].c_str() || *end != 0) {
            err = std::string("Error reading config file! (") + what + ": " + tok[p] + ")";
            return false;
        }
        v = (int)x;
        p++;
        return true;
    };
    int numRepeats = 1;
    if (!take_int(cfg.FS, "FS") || !take_int(cfg.buffer_size, "buffer_size") ||
        !take_int(cfg.nIn, "numInputChannels") || !take_int(cfg.nOut, "numOutputChannels") ||
        !take_int(cfg.numPCs, "numPCs") || !take_int(numRepeats, "numRepeats"))
        return false;
    if (cfg.numPCs < 1 || numRepeats < 1 || cfg.nIn < 1 || cfg.nOut < 1 || cfg.buffer_size < 1) {
        err = "Error reading config file! (header values must be positive)";
        return false;
    }
    std::string dir;
    size_t slash = path.find_last_of('/');
    if (slash != std::string::npos) dir = path.substr(0, slash + 1);
    std::vector<ConfigLine> base;
    for (int i = 0; i < cfg.numPCs; i++) {
        ConfigLine L;
        if (p >= tok.size()) { err = "Error reading config file! 2"; return false; }
        L.ir_file = tok[p++];
        if (!dir.empty() && L.ir_file[0] != '/') {
            std::ifstream probe(L.ir_file);
            if (!probe) L.ir_file = dir + L.ir_file;
        }
        if (!take_int(L.mute, "mute") || !take_int(L.in_ch, "inputChannel") ||
            !take_int(L.out_ch, "outputChannel"))
            return false;
        if (p >= tok.size()) { err = "Error reading config file! 2"; return false; }
        {
            char *end = nullptr;
            L.scale = strtof(tok[p].c_str(), &end);
            if (end == tok[p].c_str()) { err = "Error reading config file! 2 (scaleFactor)"; return false; }
            p++;
        }
        int M = 0;
        if (!take_int(M, "M")) return false;
        if (M < 1 || M > 16) { err = "Error reading config file! 3 (levels)"; return false; }
        for (int j = 0; j < M; j++) {
            int n = 0;
            if (!take_int(n, "N")) { err = "Error reading config file! 3"; return false; }
            L.N.push_back(n);
            if (j < M - 1) {
                int c = 0;
                if (!take_int(c, "num_parts")) { err = "Error reading config file! 3"; return false; }
                L.P.push_back(c);
            } else {
                if (p < tok.size()) {
                    char *end = nullptr;
                    strtol(tok[p].c_str(), &end, 10);
                    if (end != tok[p].c_str() && *end == 0) p++;
                }
                L.P.push_back(0);
            }
        }
        if (L.N[0] != cfg.buffer_size) {
            char msg[128];
            snprintf(msg, sizeof msg, "line %d: first block size must equal buffer_size", i + 1);
            err = msg;
            return false;
        }
        base.push_back(L);
    }
    cfg.lines = base;
    for (int r = 1; r < numRepeats; r++)
        for (const ConfigLine &b : base) {
            ConfigLine c = b;
            c.mute = 1;
            cfg.lines.push_back(c);
        }
    cfg.numPCs = (int)cfg.lines.size();
    return true;
}

}  // namespace pcio
