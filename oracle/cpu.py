"""ctypes front end of oracle/libmfcpu.so, the single-thread FP32 CPU
restatement (test infrastructure and the timed CPU baseline only)."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libmfcpu.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "libmfcpu.so"], check=True)
    return LIB


def build_native(outdir):
    """Build the same source with -march=native for the host it runs on (the
    timed CPU baseline) into `outdir`; returns the library path."""
    out = os.path.join(outdir, "libmfcpu_native.so")
    subprocess.run(["g++", "-O3", "-march=native", "-ffp-contract=off", "-fPIC", "-std=c++17", "-shared",
                    "-o", out, os.path.join(HERE, "mfcpu.cpp")], check=True)
    return out


def _declare(L):
    L.mfcpu_flat.restype = ctypes.c_long
    L.mfcpu_flat.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                             ctypes.c_void_p, ctypes.c_size_t, ctypes.c_float, ctypes.c_size_t,
                             ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.mfcpu_block_spectrum.restype = ctypes.c_int
    L.mfcpu_block_spectrum.argtypes = [ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_size_t, ctypes.c_void_p]
    return L


def load(path=None):
    """The checker library (default), or the library at `path`."""
    global _lib
    if path is not None:
        return _declare(ctypes.CDLL(path))
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = _declare(ctypes.CDLL(LIB))
    return _lib


def flat(D, T, binsize=None, threshold=0.0, window=None, lib=None):
    """FP32 CPU flat filter: (idx int64, val complex64, counts int32)."""
    L = load() if lib is None else lib
    D = np.ascontiguousarray(np.atleast_2d(D), dtype=np.complex64)
    T = np.ascontiguousarray(np.atleast_2d(T), dtype=np.complex64)
    nd, n = D.shape
    nt = T.shape[0]
    binsize = n if binsize is None else int(binsize)
    ws, we = (0, n) if window is None else (int(window[0]), int(window[1]))
    nb = -(-(min(we, n) - ws) // binsize)
    idx = np.empty((nd, nt, nb), np.int64)
    val = np.empty((nd, nt, nb), np.complex64)
    cnt = np.empty((nd, nt), np.int32)
    r = L.mfcpu_flat(n, nd, nt, D.ctypes.data, T.ctypes.data, binsize, float(threshold), ws, we,
                     idx.ctypes.data, val.ctypes.data, cnt.ctypes.data)
    if r < 0:
        raise ValueError("mfcpu_flat rejected its arguments")
    return idx, val, cnt


def block_spectrum(series, start, n):
    L = load()
    s = np.ascontiguousarray(series, dtype=np.complex64)
    out = np.empty(n, np.complex64)
    if L.mfcpu_block_spectrum(n, s.ctypes.data, s.size, int(start), out.ctypes.data) < 0:
        raise ValueError("bad length")
    return out
