"""Shared test helpers: WAV bytes, the reference's pinned inputs/outputs, and the oracle (CPU restatement of the
reference, test infrastructure only) through ctypes."""
import ctypes
import hashlib
import math
import os
import random
import struct
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "kbsrc")
GPU_CLI = os.path.join(ROOT, "kbsrc_amd", "bin", "kbsrc")

# Pins published by the reference's own test suite (tests/test_engines.py:26-38): SHA-256 of the input WAVs, of the
# CLI outputs, and the exact samples of one stereo frame 8 kHz -> 16 kHz --len 0.01 (planar, channel after channel).
PRIMARY_WAV = "59c2fbaef9de2e52113f766058c3cc8d9cdac07fab7839c3666cf7600a34d5fe"
SUBNORMAL_WAV = "83d27bd8fb3438a12503cd1c0a9e78a1c980917ab4c808d226f14aca1b5a8f12"
CLI_PINS = [  # input, rate, mode, sha256 of the output WAV
    ("primary", 96000, "exact", "355d557004d979d59358252264c63e689a136b51103c4390b2985a1c85f8398b"),
    ("primary", 96000, "fast", "5c71eba51cfc4958af6736b2c4657dd3cb6bf291f643b23e22f3abe799d5cb07"),
    ("primary", 44100, "exact", "752c3fe855e795ecbba79e83aaf00cfe47fa09c9d4905d2e3e61235197dc1ccc"),
    ("subnormal", 96000, "exact", "90ae0178d5ae1fea6505f0c03c130a6ddb3a880d2b74ad21dd25cf53dc5e7fea"),
    ("subnormal", 44100, "exact", "0a03683c6809085e3d1d2e926d21aa3faec9aecc3db66590f501c68440b7bef2"),
]
SUBNORMAL_96K_PLANAR = "b1a721cde705af8f2a055dc45dc4575a46b29999f48da18ae4da590a04d8c7c3"  # --len 0.1
FRAME1 = (-0.5955104827880859, 0.11259841918945312)
FRAME1_OUT = ["bfe003180f1a9fbe", "bfd77c39344e9ed3", "3fb8386a3126e979", "3fb1c322eee114b9"]

# Pins measured on the compiled reference for the benchmark workloads (float64 WAV in and out).
IN_48K_1S = "c4337677f970288e21daeb63dc74cc1605bbc3d5372d06d26ce50776f017fe04"
PINS_48K_1S = {("exact", True): "25cf220e8557f2b33649a7d7ab8824c8f89708b41d08597fb9e79d9fd5ba33b6",
           ("fast", True): "671316329d4984f9fad1286f53b56550bdffbd52de66f1f8f7b90436bf83a251",
           ("exact", False): "46b7159b71385b7c9b373fcb377bdc1d32cc5a03539697f814191a63f109be7b"}
IN_96K_180S = "f44fdd75bb4cd2fed49617582d5437b96a445d71770e6bd77558fbc59ee9901c"
PINS_96K_180S = {"exact_f64": "53dbcdc814aba411c04913eec37e65a59f508a7d7b45496e2289aa1d28a7fe2a",
           "fast_f64": "a7da412e086fbafc38026607bd44ab6848a9456e4e37f536c690999a6243fa46",
           "exact_s24": "56c469a411c7b2e45568ad2b38949c8269e0f031a7a5d4cdb04e515eab662c39",
           "exact_s16": "04f5ffe2af3971e0aee6dc4b97089c3aeca307a9ade93895118b000ca091a37f",
           "exact_s32": "06c4edb4e089bb988b91a98fe2aa04e6fdce59b4ce6762d7d6e2a97e13ca0d4f"}
IN_44K_180S = "4d52e1161c75429ba5ae91dd4eca67135a8cf749c957a1166e3a4ff49ed94166"
PIN_44K_TO_48K_180S = "6ed5128888063dca0ec7f8254dba14ac4312a008d69fbe9c6daa95940b5de58e"
# 10 MHz -> 48 kHz mono, --len 0.5 (5,000,000-tap kernel, 20,160,000-point block FFT: 3150 x 3200 four-step), on
# input_10m_short(): sha256 of the raw float64 output samples. Measured on the compiled reference; the oracle gives
# the same bits.
PINS_10M_SHORT = {"exact": "cb6ebf41f1876e33961823a63536dbe1576ba05355d50bc5690691b6af408820",
                  "fast": "045b10e3c725eec048444ac74117755bf2507286630753a972785fdf34986c77"}

# The same rates on 30 s (300,000,000 frames of input_10m_30s()): 31 blocks exact, and the fast plan's own blocks.
PINS_10M_30S = {"exact": "41ba6929f40b098befd88dc7036a0b42fd9b04912f91ff13cb043e7367612dc0",
                "fast": "f673ce48cae1a103858663581d33bbbce6ce4e566fda554955d46769b77c45cf"}


# 8-channel 384 kHz -> 44.1 kHz, --len 0.5 (ratio 147/1280): channel c is default_rng(20 + c).uniform(-0.5, 0.5)
FRAMES_384K_SHORT = 768_000        # 2 s, run at --block 1536 (the full-length plan's block: same transform shapes)
FRAMES_384K_FULL = 230_400_000     # 600 s


# sha256 of the planar float64 output, measured on the compiled reference (tests/golden/make_extreme_pins.py)
PINS_384K_SHORT = {"exact": "d0865900cdf580cbde077c69e9e0f9517ff551f62d01583309bc9830a3cdc9fa",
                   "fast": "c8e53844cbbe5ec98211be5bc68f9e126f4c53fea0b5836999338793cd5fedb9"}
PIN_384K_600S = "a5456c800f46db850a20cadab8344ae5e4d8e0f8686cf38e15b23c0c5afe55b6"   # exact, auto block (1536)


def input_384k_chunks(frames, chunk=1 << 22):
    """(n, 8) interleaved blocks of the 384 kHz input, in order (each channel's stream drawn chunk by chunk)."""
    rngs = [np.random.default_rng(20 + c) for c in range(8)]
    for i in range(0, frames, chunk):
        n = min(chunk, frames - i)
        yield np.stack([r.uniform(-0.5, 0.5, n) for r in rngs], 1)


def input_384k(frames):
    x = np.empty((frames, 8))
    i = 0
    for blk in input_384k_chunks(frames):
        x[i:i + len(blk)] = blk
        i += len(blk)
    return x


def input_10m_30s():
    return np.random.default_rng(5).uniform(-0.5, 0.5, 300_000_000)


def input_10m_short():
    return np.random.default_rng(5).uniform(-0.5, 0.5, 200000)


def primary97():
    rng = random.Random(0x4B42535243)
    return np.array([rng.randint(-1048576, 1048576) / 1048576 for _ in range(194)])


def subnormal97():
    rng = random.Random(12345)
    return np.array([math.ldexp(rng.randint(-1024, 1024) / 1024, -1024) for _ in range(194)])


def input_48k_1s():
    return np.random.default_rng(1).uniform(-0.5, 0.5, 48000)


def input_96k_180s(frames=17_280_000):
    """Stereo interleaved: channel c from default_rng(c + 1)."""
    x = np.empty((frames, 2))
    x[:, 0] = np.random.default_rng(1).uniform(-0.5, 0.5, frames)
    x[:, 1] = np.random.default_rng(2).uniform(-0.5, 0.5, frames)
    return x


def input_44k_180s(frames=44100 * 180):
    x = np.empty((frames, 2))
    x[:, 0] = np.random.default_rng(3).uniform(-0.5, 0.5, frames)
    x[:, 1] = np.random.default_rng(4).uniform(-0.5, 0.5, frames)
    return x


def sha(b: bytes) -> str:
    return hashlib.sha256(b).hexdigest()


def wav_header_f64(ch, rate, frames):
    data_len = frames * ch * 8
    fmt = struct.pack("<HHIIHHH", 3, ch, rate, rate * 8 * ch, 8 * ch, 64, 0)
    body_len = 4 + 8 + len(fmt) + 12 + 8 + data_len
    return (b"RIFF" + struct.pack("<I", body_len) + b"WAVE" + b"fmt " + struct.pack("<I", len(fmt)) + fmt
            + b"fact" + struct.pack("<II", 4, frames) + b"data" + struct.pack("<I", data_len))


def wav_f64(interleaved, ch, rate):
    """64-bit float WAV, 18-byte fmt chunk (cbSize 0) and a fact chunk: the reference's output layout."""
    a = np.ascontiguousarray(np.asarray(interleaved, dtype="<f8").reshape(-1))
    return wav_header_f64(ch, rate, len(a) // ch) + a.tobytes()


def wav_sha_planar(y, rate):
    """SHA-256 of the f64 WAV the reference CLI writes for planar output y (ch, n), streamed."""
    y = np.asarray(y)
    ch, n = y.shape
    h = hashlib.sha256(wav_header_f64(ch, rate, n))
    step = 1 << 20
    for i in range(0, n, step):
        h.update(np.ascontiguousarray(y[:, i:i + step].T).astype("<f8").tobytes())
    return h.hexdigest()


def read_wav_f64(path):
    with open(path, "rb") as f:
        blob = f.read()
    pos = 12
    ch = None
    while pos + 8 <= len(blob):
        tag, size = blob[pos:pos + 4], struct.unpack("<I", blob[pos + 4:pos + 8])[0]
        if tag == b"fmt ":
            ch = struct.unpack("<H", blob[pos + 10:pos + 12])[0]
        if tag == b"data":
            data = np.frombuffer(blob[pos + 8:pos + 8 + size], dtype="<f8")
            return data.reshape(-1, ch).T.copy()
        pos += 8 + size + (size & 1)
    return None


# ---- oracle ---------------------------------------------------------------------------------------------------
_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        lib = ctypes.CDLL(ORACLE_SO)
        P, LL, I, D = ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int, ctypes.c_double
        lib.oracle_resample.argtypes = [P, LL, I, I, I, I, D, D, I, I, P]
        lib.oracle_plan.argtypes = [LL, I, I, I, I, D, D, I, ctypes.POINTER(LL)]
        lib.oracle_last_error.restype = ctypes.c_char_p
        lib.oracle_cis.argtypes = [LL, LL, P]
        lib.oracle_cis_table.argtypes = [LL, LL, LL, P]
        lib.oracle_kernel_taps.argtypes = [I, LL, LL, D, P]
        lib.oracle_kernel_spectrum.argtypes = [P, I, I, D, P]
        lib.oracle_kernel_taps_at.argtypes = [I, LL, LL, D, P, I, P]
        lib.oracle_burg.argtypes = [P, LL, LL, P]
        lib.oracle_cfft.argtypes = [I, I, P, P]
        _oracle = lib
    return _oracle


def oracle_taps(L, a, b, beta, idx=None):
    """Oracle taps as an (n, 2) array of (hi, lo): all of h[0..L], or h[idx] only."""
    if idx is None:
        v = np.empty((L + 1, 2))
        rc = oracle().oracle_kernel_taps(L, a, b, beta, v.ctypes.data)
    else:
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        v = np.empty((len(idx), 2))
        rc = oracle().oracle_kernel_taps_at(L, a, b, beta, idx.ctypes.data, len(idx), v.ctypes.data)
    assert rc == 0, oracle().oracle_last_error()
    return v


def oracle_spectrum(taps, n, scale):
    t = np.ascontiguousarray(taps, dtype=np.float64).reshape(-1, 2)
    v = np.empty((n // 2 + 1, 2))
    rc = oracle().oracle_kernel_spectrum(t.ctypes.data, len(t) - 1, n, scale, v.ctypes.data)
    assert rc == 0, oracle().oracle_last_error()
    return v


ORACLE_PLAN_FIELDS = ["P", "Q", "nout", "nb", "marg", "Nb", "Mb", "hop", "nblk", "o", "tot", "base", "pre", "E", "L",
                      "Lk", "c2", "gen", "up", "equal"]


def oracle_plan(frames, ch, rin, rout, precision="exact", length=0.2, beta=40.0, block=0):
    v = (ctypes.c_longlong * 20)()
    rc = oracle().oracle_plan(frames, ch, rin, rout, 1 if precision == "exact" else 0, length, beta, block, v)
    if rc:
        raise ValueError(oracle().oracle_last_error().decode())
    return dict(zip(ORACLE_PLAN_FIELDS, list(v)))


def oracle_resample(x, ch, rin, rout, precision="exact", length=0.2, beta=40.0, block=0, lpc=True):
    """x: interleaved float64 (frames*ch). Returns planar (ch, nout)."""
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))
    frames = len(x) // ch
    g = math.gcd(rin, rout)
    nout = -(-frames * (rout // g) // (rin // g))
    out = np.zeros(ch * nout)
    rc = oracle().oracle_resample(x.ctypes.data if len(x) else None, frames, ch, rin, rout,
                                  1 if precision == "exact" else 0, length, beta, block, int(lpc), out.ctypes.data)
    if rc == 1:
        raise ValueError(oracle().oracle_last_error().decode())
    if rc:
        raise RuntimeError(oracle().oracle_last_error().decode())
    return out.reshape(ch, nout)


def ref_available():
    return os.path.exists(REF_BIN)


def ref_resample(tmp_path, x, ch, rin, rout, precision="exact", length=0.2, beta=40.0, block=0, lpc=True):
    """Run the compiled reference CLI (oracle/_ref) on a float64 WAV; returns planar output."""
    src, dst = os.path.join(tmp_path, "ref_in.wav"), os.path.join(tmp_path, "ref_out.wav")
    with open(src, "wb") as f:
        f.write(wav_f64(x, ch, rin))
    args = [REF_BIN, src, dst, "--rate", str(rout), "--mode", precision, "--len", repr(length), "--beta", repr(beta)]
    if block:
        args += ["--block", str(block)]
    if not lpc:
        args.append("--no-lpc")
    r = subprocess.run(args, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(r.stderr)
    return read_wav_f64(dst)
