"""Python binding of the MI355X resampler (libkbsrc_gpu.so, C ABI in include/kbsrc_gpu.h).

Mirrors the reference's library call ``kbsrc::resample`` (include/kbsrc/kbsrc.hpp:41-43): interleaved float64
frames in, planar float64 channels out, with the same options (precision, length, beta, block, lpc_edges) and the
same errors (``ValueError`` where the reference throws ``std::invalid_argument``, with its message). The engine
seam (plan -> stage -> execute on device buffers) is exposed through :class:`Plan` for callers that keep data in
HBM. There is no CPU fallback: importing fails loudly when the built library is missing.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

try:   # one HIP runtime per process: when torch is present its runtime must be the one this library binds to
    import torch  # noqa: F401
except ImportError:
    pass

__all__ = ["resample", "output_frames", "Plan", "Stats", "version", "library_path", "FAST", "EXACT",
           "KbsrcError", "PLAN_FIELDS", "kernel_taps", "kernel_spectrum", "pack", "PACK_FORMATS", "Batch",
           "resample_batch", "tuned_shapes",
           "split_dft16", "plain_dft16", "plain_dft16_inwave", "Graph"]

FAST, EXACT, LONG_DOUBLE = 0, 1, 2
OK, EINVAL, ERUNTIME, ENOMEM = 0, 1, 2, 3
PLAN_FIELDS = ["P", "Q", "nout", "nb", "marg", "Nb", "Mb", "hop", "nblk", "o", "tot", "base", "pre", "E", "L", "Lk",
               "c2", "gen", "up", "equal", "tasks", "batch", "xstride"]

_HERE = os.path.dirname(os.path.abspath(__file__))


def library_path() -> str:
    """The engine library: KBSRC_LIB when set (an alternative build, e.g. for side-by-side timing), else the
    in-tree build."""
    return os.environ.get("KBSRC_LIB") or os.path.join(_HERE, "libkbsrc_gpu.so")


class KbsrcError(RuntimeError):
    """A device or internal error (std::runtime_error in the reference)."""


class _Options(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("length", ctypes.c_double), ("beta", ctypes.c_double),
                ("threads", ctypes.c_int), ("block", ctypes.c_int), ("lpc_edges", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("p", ctypes.c_longlong), ("q", ctypes.c_longlong), ("taps", ctypes.c_longlong),
                ("block", ctypes.c_int), ("threads", ctypes.c_int), ("general_fold", ctypes.c_int),
                ("design_seconds", ctypes.c_double), ("process_seconds", ctypes.c_double),
                ("passband", ctypes.c_double), ("stopband", ctypes.c_double), ("isa", ctypes.c_char_p)]


def _load():
    path = library_path()
    if not os.path.exists(path):
        raise ImportError(f"kbsrc_amd: {path} is not built (run __graft_entry__.build() or make -C kbsrc_amd)")
    lib = ctypes.CDLL(path)
    P, LL, I, D = ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int, ctypes.c_double
    sig = {
        "kbsrc_default_options": (None, [ctypes.POINTER(_Options)]),
        "kbsrc_version": (ctypes.c_char_p, []),
        "kbsrc_last_error": (ctypes.c_char_p, []),
        "kbsrc_output_frames": (I, [LL, I, I, ctypes.POINTER(LL)]),
        "kbsrc_resample": (I, [P, LL, I, I, I, ctypes.POINTER(_Options), P, ctypes.POINTER(Stats)]),
        "kbsrc_plan_create": (I, [LL, I, I, I, ctypes.POINTER(_Options), ctypes.POINTER(P)]),
        "kbsrc_plan_query": (I, [P, ctypes.POINTER(LL)]),
        "kbsrc_plan_stage": (I, [P, P, P, P]),
        "kbsrc_plan_execute": (I, [P, P, P, P]),
        "kbsrc_plan_execute_split": (I, [P, P, P, I, P]),
        "kbsrc_plan_execute_dd": (I, [P, P, P, I, P]),
        "kbsrc_plan_execute_noalloc": (I, [P, P, P, P]),
        "kbsrc_plan_execute_timed": (I, [P, P, P, P, ctypes.POINTER(D)]),
        "kbsrc_plan_destroy": (None, [P]),
        "kbsrc_plan_stage_device": (I, [P, P, P, P]),
        "kbsrc_plan_stats": (I, [P, ctypes.POINTER(Stats)]),
        "kbsrc_pack": (I, [P, I, LL, LL, LL, I, P, P, P]),
        "kbsrc_batch_create": (I, [I, P, ctypes.POINTER(_Options), ctypes.POINTER(P)]),
        "kbsrc_batch_query": (I, [P, I, ctypes.POINTER(LL)]),
        "kbsrc_batch_stage": (I, [P, I, P, I, P, P]),
        "kbsrc_batch_execute": (I, [P, P, P, P]),
        "kbsrc_batch_destroy": (None, [P]),
        "kbsrc_plan_tune": (I, [P, P, P, P]),
        "kbsrc_batch_tune": (I, [P, P, P, P]),
        "kbsrc_plan_capture": (I, [P, P, P, ctypes.POINTER(P)]),
        "kbsrc_batch_capture": (I, [P, P, P, ctypes.POINTER(P)]),
        "kbsrc_graph_launch": (I, [P, P]),
        "kbsrc_graph_destroy": (None, [P]),
        "kbsrc_tuned_shapes": (I, []),
        "kbsrc_split_dft16": (I, [P, LL, I, P, P]),
        "kbsrc_plain_dft16": (I, [P, LL, P, P]),
        "kbsrc_plain_dft16_inwave": (I, [P, LL, P, P]),
        "kbsrc_kernel_taps": (I, [I, LL, LL, D, P]),
        "kbsrc_kernel_spectrum": (I, [P, I, I, D, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = _load()


def _check(rc: int):
    if rc == OK:
        return
    msg = (_lib.kbsrc_last_error() or b"").decode()
    if rc == EINVAL:
        raise ValueError(msg)
    if rc == ENOMEM:
        raise MemoryError(msg)
    raise KbsrcError(msg)


def _options(precision, length, beta, block, lpc_edges, threads=0) -> _Options:
    o = _Options()
    _lib.kbsrc_default_options(ctypes.byref(o))
    o.precision = {"exact": EXACT, "fast": FAST, "ld": LONG_DOUBLE}.get(precision, precision) \
        if isinstance(precision, str) else int(precision)
    o.length, o.beta, o.block, o.lpc_edges, o.threads = float(length), float(beta), int(block), int(bool(lpc_edges)), \
        int(threads)
    return o


def split_dft16(d_x: int, nvec: int, slices: int, d_y: int, stream: int = 0):
    """Split-precision (int8 MFMA) forward DFTs of nvec 16-point complex dd vectors (device pointers)."""
    _check(_lib.kbsrc_split_dft16(ctypes.c_void_p(d_x), nvec, slices, ctypes.c_void_p(d_y), ctypes.c_void_p(stream)))


def plain_dft16(d_x: int, nvec: int, d_y: int, stream: int = 0):
    """The same DFTs by the engine's plain dd radix-4 x 4 pass."""
    _check(_lib.kbsrc_plain_dft16(ctypes.c_void_p(d_x), nvec, ctypes.c_void_p(d_y), ctypes.c_void_p(stream)))


def plain_dft16_inwave(d_x: int, nvec: int, d_y: int, stream: int = 0):
    """The same plain dd pass with its stage exchange inside the wave instead of LDS (same bits)."""
    _check(_lib.kbsrc_plain_dft16_inwave(ctypes.c_void_p(d_x), nvec, ctypes.c_void_p(d_y), ctypes.c_void_p(stream)))


def tuned_shapes() -> int:
    """Sub-FFT pass shapes measured so far in this process (Plan.tune, Batch.tune)."""
    return _lib.kbsrc_tuned_shapes()


def version() -> str:
    return _lib.kbsrc_version().decode()


def output_frames(frames: int, rate_in: int, rate_out: int) -> int:
    n = ctypes.c_longlong()
    _check(_lib.kbsrc_output_frames(frames, rate_in, rate_out, ctypes.byref(n)))
    return n.value


def resample(x, rate_in: int, rate_out: int, precision="exact", length=0.2, beta=40.0, block=0, lpc_edges=True,
             threads=0, return_stats=False):
    """Resample ``x`` (frames x channels, or 1-D mono) float64 from rate_in to rate_out on the GPU.

    Returns a (channels, nout) float64 array (planar, as the reference's vector of channels)."""
    a = np.asarray(x, dtype=np.float64)
    if a.ndim == 1:
        a = a[:, None]
    a = np.ascontiguousarray(a)
    frames, ch = a.shape
    if ch < 1:
        raise ValueError("bad input buffer")
    nout = output_frames(frames, rate_in, rate_out) if rate_in >= 1 and rate_out >= 1 else 0
    out = np.empty((ch, max(nout, 0)), dtype=np.float64)
    st = Stats()
    o = _options(precision, length, beta, block, lpc_edges, threads)
    _check(_lib.kbsrc_resample(a.ctypes.data if frames else None, frames, ch, rate_in, rate_out, ctypes.byref(o),
                               out.ctypes.data if out.size else None, ctypes.byref(st)))
    return (out, st) if return_stats else out


def kernel_taps(L: int, a: int, b: int, beta: float) -> np.ndarray:
    """The reference's kernel_taps (src/design.cpp:17-37) on the GPU: taps h[0..L] as an (L+1, 2) array of
    (hi, lo) double-double parts."""
    out = np.empty((L + 1, 2) if L >= 0 else (1, 2), dtype=np.float64)
    _check(_lib.kbsrc_kernel_taps(int(L), int(a), int(b), float(beta), out.ctypes.data))
    return out


def kernel_spectrum(taps, n: int, scale: float) -> np.ndarray:
    """The reference's kernel_spectrum (src/design.cpp:40-49) on the GPU: taps as an (Lk+1, 2) (hi, lo) array,
    placed circularly in length n; returns Re X[k] / scale for k = 0..n/2 as an (n/2+1, 2) array."""
    t = np.ascontiguousarray(np.asarray(taps, dtype=np.float64).reshape(-1, 2))
    out = np.empty((max(n, 0) // 2 + 1, 2), dtype=np.float64)
    _check(_lib.kbsrc_kernel_spectrum(t.ctypes.data, len(t) - 1, int(n), float(scale), out.ctypes.data))
    return out


PACK_FORMATS = {"f64": 0, "f32": 1, "s32": 2, "s24": 3, "s16": 4}


def pack(d_planar: int, channels: int, nout: int, frame0: int, frames: int, fmt: str, d_bytes: int, d_clipped: int,
         stream: int = 0):
    """Output packing on the GPU (WAV sample bytes, interleaved, no dither); device pointers as integers."""
    _check(_lib.kbsrc_pack(ctypes.c_void_p(d_planar), channels, nout, frame0, frames, PACK_FORMATS[fmt],
                           ctypes.c_void_p(d_bytes), ctypes.c_void_p(d_clipped), ctypes.c_void_p(stream)))


class _Job(ctypes.Structure):
    _fields_ = [("frames", ctypes.c_longlong), ("channels", ctypes.c_int), ("rate_in", ctypes.c_int),
                ("rate_out", ctypes.c_int)]


class Graph:
    """One plan or batch execute on fixed device buffers, captured as a HIP graph (kbsrc_*_capture); ``launch``
    replays it with one call. Keep the plan or batch alive while the graph exists."""

    def __init__(self, handle, owner):
        self._h = handle
        self._owner = owner

    def launch(self, stream: int = 0):
        _check(_lib.kbsrc_graph_launch(self._h, ctypes.c_void_p(stream)))

    def close(self):
        if self._h:
            _lib.kbsrc_graph_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many jobs in one pair of device buffers (kbsrc_batch_*): jobs = [(frames, channels, rate_in, rate_out)],
    one set of options. ``layout[j]`` = dict(in_offset, in_doubles, out_offset, out_doubles, nout, group) in
    doubles; ``in_doubles`` / ``out_doubles`` the buffer sizes."""

    def __init__(self, jobs, precision="exact", length=0.2, beta=40.0, block=0, lpc_edges=True):
        arr = (_Job * len(jobs))(*[_Job(int(f), int(c), int(a), int(b)) for f, c, a, b in jobs])
        self._h = ctypes.c_void_p()
        o = _options(precision, length, beta, block, lpc_edges)
        _check(_lib.kbsrc_batch_create(len(jobs), arr, ctypes.byref(o), ctypes.byref(self._h)))
        v = (ctypes.c_longlong * 6)()
        _check(_lib.kbsrc_batch_query(self._h, -1, v))
        self.in_doubles, self.out_doubles, self.groups = v[1], v[3], v[5]
        self.fused = bool(v[0])   # executes run one launch per step over every group
        self.layout = []
        keys = ["in_offset", "in_doubles", "out_offset", "out_doubles", "nout", "group"]
        for j in range(len(jobs)):
            _check(_lib.kbsrc_batch_query(self._h, j, v))
            self.layout.append(dict(zip(keys, list(v))))

    def stage(self, job: int, interleaved, d_in: int, stream: int = 0):
        """Host interleaved (frames, channels) float64 input of one job into the batch input buffer."""
        a = np.ascontiguousarray(np.asarray(interleaved, dtype=np.float64))
        _check(_lib.kbsrc_batch_stage(self._h, job, a.ctypes.data if a.size else None, 0, ctypes.c_void_p(d_in),
                                      ctypes.c_void_p(stream)))

    def stage_device(self, job: int, d_interleaved: int, d_in: int, stream: int = 0):
        _check(_lib.kbsrc_batch_stage(self._h, job, ctypes.c_void_p(d_interleaved), 1, ctypes.c_void_p(d_in),
                                      ctypes.c_void_p(stream)))

    def tune(self, d_in: int, d_out: int, stream: int = 0):
        """Execute once while measuring kernel shapes for the passes not measured before (same outputs)."""
        _check(_lib.kbsrc_batch_tune(self._h, ctypes.c_void_p(d_in), ctypes.c_void_p(d_out), ctypes.c_void_p(stream)))

    def execute(self, d_in: int, d_out: int, stream: int = 0):
        _check(_lib.kbsrc_batch_execute(self._h, ctypes.c_void_p(d_in), ctypes.c_void_p(d_out),
                                        ctypes.c_void_p(stream)))

    def capture(self, d_in: int, d_out: int) -> Graph:
        """The batch execute on these buffers as a replayable HIP graph."""
        g = ctypes.c_void_p()
        _check(_lib.kbsrc_batch_capture(self._h, ctypes.c_void_p(d_in), ctypes.c_void_p(d_out), ctypes.byref(g)))
        return Graph(g, self)

    def close(self):
        if self._h:
            _lib.kbsrc_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def resample_batch(signals, rates_in, rates_out, precision="exact", length=0.2, beta=40.0, block=0,
                   lpc_edges=True):
    """Resample several signals ((frames, channels) float64 each) in one batch on the GPU; returns a list of
    (channels, nout) arrays. Device buffers come from torch."""
    import torch
    xs = [np.asarray(x, dtype=np.float64) for x in signals]
    xs = [x[:, None] if x.ndim == 1 else x for x in xs]
    b = Batch([(x.shape[0], x.shape[1], ri, ro) for x, ri, ro in zip(xs, rates_in, rates_out)], precision, length,
              beta, block, lpc_edges)
    d_in = torch.zeros(max(1, b.in_doubles), dtype=torch.float64, device="cuda")
    d_out = torch.zeros(max(1, b.out_doubles), dtype=torch.float64, device="cuda")
    for j, x in enumerate(xs):
        b.stage(j, x, d_in.data_ptr())
    b.execute(d_in.data_ptr(), d_out.data_ptr())
    torch.cuda.synchronize()
    host = d_out.cpu().numpy()
    out = []
    for x, L in zip(xs, b.layout):
        out.append(host[L["out_offset"]:L["out_offset"] + L["out_doubles"]].reshape(x.shape[1], L["nout"]).copy())
    b.close()
    return out


class Plan:
    """The engine seam: plan/design once, then stage and execute on device buffers (integer device pointers,
    e.g. ``torch.Tensor.data_ptr()``), on a HIP stream handle (0 = default stream)."""

    def __init__(self, frames, channels, rate_in, rate_out, precision="exact", length=0.2, beta=40.0, block=0,
                 lpc_edges=True, threads=0):
        self._h = ctypes.c_void_p()
        o = _options(precision, length, beta, block, lpc_edges, threads)
        _check(_lib.kbsrc_plan_create(frames, channels, rate_in, rate_out, ctypes.byref(o), ctypes.byref(self._h)))
        v = (ctypes.c_longlong * len(PLAN_FIELDS))()
        _check(_lib.kbsrc_plan_query(self._h, v))
        self.info = dict(zip(PLAN_FIELDS, list(v)))
        self.channels = channels

    def stage(self, interleaved, d_padded: int, stream: int = 0):
        a = np.ascontiguousarray(np.asarray(interleaved, dtype=np.float64))
        _check(_lib.kbsrc_plan_stage(self._h, a.ctypes.data, ctypes.c_void_p(d_padded), ctypes.c_void_p(stream)))

    def stage_device(self, d_interleaved: int, d_padded: int, stream: int = 0):
        """Staging from interleaved input already in device memory (frames x channels float64)."""
        _check(_lib.kbsrc_plan_stage_device(self._h, ctypes.c_void_p(d_interleaved), ctypes.c_void_p(d_padded),
                                            ctypes.c_void_p(stream)))

    def stats(self) -> Stats:
        st = Stats()
        _check(_lib.kbsrc_plan_stats(self._h, ctypes.byref(st)))
        return st

    def execute(self, d_padded: int, d_out: int, stream: int = 0):
        _check(_lib.kbsrc_plan_execute(self._h, ctypes.c_void_p(d_padded), ctypes.c_void_p(d_out),
                                       ctypes.c_void_p(stream)))

    def execute_noalloc(self, d_padded: int, d_out: int, stream: int = 0):
        """The execute with its work buffers in the library's static device array (no device allocation); same
        bits. Calls on one GPU share the array and must not overlap."""
        _check(_lib.kbsrc_plan_execute_noalloc(self._h, ctypes.c_void_p(d_padded), ctypes.c_void_p(d_out),
                                               ctypes.c_void_p(stream)))

    def capture(self, d_padded: int, d_out: int) -> Graph:
        """The execute on these buffers as a replayable HIP graph."""
        g = ctypes.c_void_p()
        _check(_lib.kbsrc_plan_capture(self._h, ctypes.c_void_p(d_padded), ctypes.c_void_p(d_out), ctypes.byref(g)))
        return Graph(g, self)

    def execute_split(self, d_padded: int, d_out: int, slices: int, stream: int = 0):
        """The split-precision form (exact-mode plans): sub-FFTs whose first two stages are radix 4 run them as one
        int8-MFMA radix-16 stage with ``slices`` base-128 slices per operand. Not bit-exact; asynchronous."""
        _check(_lib.kbsrc_plan_execute_split(self._h, ctypes.c_void_p(d_padded), ctypes.c_void_p(d_out), int(slices),
                                             ctypes.c_void_p(stream)))

    def execute_dd(self, d_padded: int, d_out_pairs: int, slices: int = 0, stream: int = 0):
        """Every output sample as its dd value before the final rounding: d_out_pairs holds channels x nout (hi, lo)
        pairs (exact-mode plans; slices 0 the plain form, > 0 the split form). Asynchronous."""
        _check(_lib.kbsrc_plan_execute_dd(self._h, ctypes.c_void_p(d_padded), ctypes.c_void_p(d_out_pairs),
                                          int(slices), ctypes.c_void_p(stream)))

    def tune(self, d_padded: int, d_out: int, stream: int = 0):
        """Execute once while measuring kernel shapes for the passes not measured before (same outputs)."""
        _check(_lib.kbsrc_plan_tune(self._h, ctypes.c_void_p(d_padded), ctypes.c_void_p(d_out),
                                    ctypes.c_void_p(stream)))

    def execute_timed(self, d_padded: int, d_out: int, stream: int = 0):
        ms = (ctypes.c_double * 5)()
        _check(_lib.kbsrc_plan_execute_timed(self._h, ctypes.c_void_p(d_padded), ctypes.c_void_p(d_out),
                                             ctypes.c_void_p(stream), ms))
        return dict(zip(["fwd_col", "fwd_row", "mid", "inv_col", "inv_row"], list(ms)))

    def close(self):
        if self._h:
            _lib.kbsrc_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
