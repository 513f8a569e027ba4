"""Pins for the 8-channel 384 kHz -> 44.1 kHz plans (exact, --len 0.5): sha256 of the planar float64 output of the
compiled reference (oracle/_ref/kbsrc, built from the reference's sources by oracle/ref.mk) on the inputs of
tests/helpers.input_384k(). Prints the digests that tests/helpers.py holds; run from the repository root:

    python tests/golden/make_extreme_pins.py [short|full]

The full input (230,400,000 frames x 8 channels, 14.7 GB as a float64 WAV) is written in chunks to a temporary
folder; the reference needs ~35 GB of memory for it."""
import hashlib
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import helpers as H  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "kbsrc")


def write_wav(path, frames, ch, rate, chunks):
    """float64 WAV (fmt 18 bytes, tag 3, cbSize 0, fact chunk) from an iterator of (n, ch) blocks. Above 4 GB the
    RIFF and data sizes are written as 0xFFFFFFFF, which the reference's reader takes as "to the end of the file"
    (src/io/wav.cpp:128-131)."""
    data_bytes = frames * ch * 8
    fmt = struct.pack("<HHIIHHH", 3, ch, rate, rate * ch * 8, ch * 8, 64, 0)
    riff = min(4 + 8 + len(fmt) + 12 + 8 + data_bytes, 0xFFFFFFFF)
    with open(path, "wb") as f:
        f.write(b"RIFF" + struct.pack("<I", riff) + b"WAVE")
        f.write(b"fmt " + struct.pack("<I", len(fmt)) + fmt)
        f.write(b"fact" + struct.pack("<II", 4, frames & 0xFFFFFFFF))
        f.write(b"data" + struct.pack("<I", min(data_bytes, 0xFFFFFFFF)))
        for blk in chunks:
            f.write(np.ascontiguousarray(blk, dtype="<f8").tobytes())


def planar_digest(path, ch):
    """sha256 of the output samples as the library returns them: channel after channel."""
    with open(path, "rb") as f:
        blob = f.read()
    pos = 12
    while pos + 8 <= len(blob):
        tag, size = blob[pos:pos + 4], struct.unpack("<I", blob[pos + 4:pos + 8])[0]
        if tag == b"data":
            y = np.frombuffer(blob, dtype="<f8", count=size // 8, offset=pos + 8).reshape(-1, ch)
            return hashlib.sha256(np.ascontiguousarray(y.T).tobytes()).hexdigest(), y.shape[0]
        pos += 8 + size + (size & 1)
    raise ValueError("no data chunk")


def run(frames, block, modes):
    with tempfile.TemporaryDirectory(prefix="kbsrc-pins-", dir=os.environ.get("PIN_TMP")) as td:
        src = os.path.join(td, "in.wav")
        write_wav(src, frames, 8, 384000, H.input_384k_chunks(frames))
        for mode in modes:
            dst = os.path.join(td, "out.wav")
            args = [REF, src, dst, "--rate", "44100", "--mode", mode, "--len", "0.5", "-t", "8", "-v"]
            if block:
                args += ["--block", str(block)]
            r = subprocess.run(args, capture_output=True, text=True)
            if r.returncode:
                raise RuntimeError(r.stderr)
            digest, nout = planar_digest(dst, 8)
            print(f"{mode} frames={frames} block={block}: {digest} nout={nout}  [{r.stderr.strip()}]", flush=True)
            os.remove(dst)


if __name__ == "__main__":
    which = sys.argv[1] if len(sys.argv) > 1 else "short"
    if which == "short":
        run(H.FRAMES_384K_SHORT, 1536, ("exact", "fast"))
    else:
        run(H.FRAMES_384K_FULL, 0, ("exact",))
