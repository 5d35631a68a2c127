"""The RIFF/WAVE impulse-response reader (pc_read_ir, the library's
readImpulseResponse, src/sndtools.cpp:3-47) on every branch: integer PCM of 8
(unsigned), 16, 24 and 32 bits, IEEE float of 32 and 64 bits, and
WAVE_FORMAT_EXTENSIBLE, each against an independent decoding (stdlib `wave`
or the raw bytes through numpy); first channel of multi-channel files; padded
frames; and malformed headers rejected before any sample is read.  No GPU."""
import struct
import wave

import numpy as np
import pytest

from partconv_amd import readImpulseResponse

RATE = 44100


def riff(chunks):
    body = b"WAVE" + b"".join(tag + struct.pack("<I", len(data)) + data + (b"\0" if len(data) & 1 else b"")
                              for tag, data in chunks)
    return b"RIFF" + struct.pack("<I", len(body)) + body


def fmt_chunk(tag, channels, bits, block_align=None, rate=RATE, ext_sub=None):
    ba = block_align if block_align is not None else channels * bits // 8
    d = struct.pack("<HHIIHH", tag, channels, rate, rate * ba, ba, bits)
    if ext_sub is not None:  # WAVE_FORMAT_EXTENSIBLE: cbSize, valid bits, channel mask, subformat GUID
        guid = struct.pack("<H", ext_sub) + bytes.fromhex("000000001000800000aa00389b71")
        d += struct.pack("<HHI", 22, bits, 0) + guid
    return (b"fmt ", d)


@pytest.mark.parametrize("bits,channels", [(8, 1), (16, 2), (24, 3), (32, 1), (16, 1)])
def test_integer_pcm_matches_stdlib_wave(tmp_path, bits, channels):
    rng = np.random.default_rng(bits * 10 + channels)
    n = 1001
    width = bits // 8
    if bits == 8:
        raw = rng.integers(0, 256, size=(n, channels), dtype=np.uint8).tobytes()
    else:
        lim = 1 << (bits - 1)
        v = rng.integers(-lim, lim, size=(n, channels), dtype=np.int64)
        v[0, 0], v[1, 0] = -lim, lim - 1          # full-scale extremes
        raw = b"".join(int(x).to_bytes(width, "little", signed=True) for x in v.ravel())
    p = tmp_path / f"pcm{bits}.wav"
    with wave.open(str(p), "wb") as w:
        w.setnchannels(channels)
        w.setsampwidth(width)
        w.setframerate(RATE)
        w.writeframes(raw)
    # independent decoding: stdlib wave back to bytes, numpy for the scaling
    with wave.open(str(p), "rb") as w:
        frames = w.readframes(w.getnframes())
    b = np.frombuffer(frames, dtype=np.uint8).reshape(n, channels, width)[:, 0, :]
    if bits == 8:
        want = (b[:, 0].astype(np.float64) - 128.0) / 128.0
    else:
        u = np.zeros(n, dtype=np.int64)
        for k in range(width):
            u |= b[:, k].astype(np.int64) << (8 * k)
        u = np.where(u >= (1 << (bits - 1)), u - (1 << bits), u)
        want = u / float(1 << (bits - 1))
    got, sr = readImpulseResponse(p)
    assert sr == RATE and got.dtype == np.float32 and got.shape == (n,)
    assert np.array_equal(got, want.astype(np.float32))


@pytest.mark.parametrize("bits,dtype", [(32, "<f4"), (64, "<f8")])
@pytest.mark.parametrize("extensible", [False, True])
def test_ieee_float_and_extensible(tmp_path, bits, dtype, extensible):
    rng = np.random.default_rng(bits + extensible)
    x = rng.standard_normal((777, 2)).astype(dtype)
    fmt = fmt_chunk(0xFFFE, 2, bits, ext_sub=3) if extensible else fmt_chunk(3, 2, bits)
    p = tmp_path / "f.wav"
    p.write_bytes(riff([fmt, (b"LIST", b"INFOjunk"), (b"data", x.tobytes())]))
    got, sr = readImpulseResponse(p)
    assert sr == RATE
    assert np.array_equal(got, x[:, 0].astype(np.float32))


def test_extensible_pcm24_and_padded_frames(tmp_path):
    """EXTENSIBLE with a PCM subformat, and frames padded beyond the samples
    (block_align 8 for one 24-bit channel): the reader steps by block_align."""
    rng = np.random.default_rng(3)
    v = rng.integers(-(1 << 23), 1 << 23, size=500)
    raw = b"".join(int(a).to_bytes(3, "little", signed=True) + b"\x55" * 5 for a in v)
    p = tmp_path / "e.wav"
    p.write_bytes(riff([fmt_chunk(0xFFFE, 1, 24, block_align=8, ext_sub=1), (b"data", raw)]))
    got, _ = readImpulseResponse(p)
    assert np.array_equal(got, (v / float(1 << 23)).astype(np.float32))


@pytest.mark.parametrize("case", ["short_block_align", "odd_bits", "bits_zero", "no_data", "no_fmt",
                                  "truncated_extensible", "unknown_tag", "not_riff", "missing"])
def test_malformed_headers_are_rejected(tmp_path, case):
    p = tmp_path / "bad.wav"
    data = (b"data", b"\x01\x02" * 9)
    if case == "short_block_align":   # 32-bit samples in 2-byte frames, data chunk last
        blob = riff([fmt_chunk(1, 1, 32, block_align=2), data])
    elif case == "odd_bits":
        blob = riff([fmt_chunk(1, 1, 12, block_align=2), data])
    elif case == "bits_zero":
        blob = riff([fmt_chunk(1, 1, 0, block_align=2), data])
    elif case == "no_data":
        blob = riff([fmt_chunk(1, 1, 16)])
    elif case == "no_fmt":
        blob = riff([data])
    elif case == "truncated_extensible":
        f = fmt_chunk(0xFFFE, 1, 16)
        blob = riff([f, data])
    elif case == "unknown_tag":
        blob = riff([fmt_chunk(2, 1, 16), data])
    elif case == "not_riff":
        blob = b"RIFX" + b"\0" * 40
    else:
        blob = None
    if blob is not None:
        p.write_bytes(blob)
    with pytest.raises(OSError):
        readImpulseResponse(p)


def test_config_with_bad_wav_reports_io_error(tmp_path):
    """pcm_setup reads the IR files named in the config (first channel): a
    malformed one fails with PC_ERR_IO before any GPU work."""
    from partconv_amd import PartConvMulti, last_error
    (tmp_path / "ir.wav").write_bytes(riff([fmt_chunk(1, 1, 32, block_align=2), (b"data", b"\0" * 8)]))
    cfg = tmp_path / "pc.config"
    cfg.write_text("48000 64 1 1 1 1\nir.wav 0 0 0 1.0 2 64 6 256\n")
    m = PartConvMulti(0)
    assert m.setup(str(cfg)) == 4
    assert "block_align" in last_error()
