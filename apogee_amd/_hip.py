"""ctypes binding of libmfhip.so and the device Context used by the filters.

The Context plays the role of the reference's backend Context protocol
(python/matchedfilter/_vkcompute.py: ``peaks`` :962-1063, ``hier_peaks``
:534-623, ``forward`` :855-906): the public classes call it at exactly the
places the reference calls ``self._gpu.peaks`` / ``.hier_peaks`` /
``.forward``.  Unlike the reference it keeps the spectra on the device from
the moment they are set, so nothing is re-uploaded per run.

There is no fallback: if the HIP library is missing or no HIP device is
visible, constructing a Context raises.
"""
import ctypes
import os
import threading

import numpy as np

from ._errors import UnsupportedSize

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmfhip.so")

_lib = None
_lock = threading.Lock()

_c_size = ctypes.c_size_t
_vp = ctypes.c_void_p


def _declare(lib):
    lib.mfh_last_error.restype = ctypes.c_char_p
    lib.mfh_device_count.restype = ctypes.c_int
    lib.mfh_device_info.argtypes = [ctypes.c_int, ctypes.c_char_p, _c_size, ctypes.c_char_p, _c_size]
    lib.mfh_supported.argtypes = [_c_size]
    lib.mfh_create.restype = _vp
    lib.mfh_create.argtypes = [ctypes.c_int, _c_size, ctypes.c_int, ctypes.c_int]
    lib.mfh_destroy.argtypes = [_vp]
    lib.mfh_destroy.restype = None
    lib.mfh_stream.argtypes = [_vp]
    lib.mfh_stream.restype = _vp
    lib.mfh_sync.argtypes = [_vp]
    for name in ("mfh_set_data", "mfh_set_templates", "mfh_set_data_device_async",
                 "mfh_set_templates_device_async"):
        getattr(lib, name).argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp]
    for name in ("mfh_bind_data", "mfh_bind_templates"):
        getattr(lib, name).argtypes = [_vp, _vp]
    lib.mfh_get_rows.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp]
    lib.mfh_hier_set_coarse_templates.argtypes = [_vp, _c_size, _vp, ctypes.c_int]
    lib.mfh_nbins.restype = _c_size
    lib.mfh_nbins.argtypes = [_c_size, _c_size, _c_size, _c_size]
    lib.mfh_flat_peaks.restype = ctypes.c_long
    lib.mfh_flat_peaks.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   _c_size, ctypes.c_float, _c_size, _c_size, _vp, _vp, _vp]
    lib.mfh_flat_peaks_rec.restype = ctypes.c_long
    lib.mfh_flat_peaks_rec.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       _c_size, ctypes.c_float, _c_size, _c_size, _vp, _vp]
    lib.mfh_flat_peaks_async.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         _c_size, ctypes.c_float, _c_size, _c_size, _vp, _vp]
    lib.mfh_hier_set_band.argtypes = [_vp, _c_size, ctypes.c_double]
    lib.mfh_hier_peaks.restype = ctypes.c_long
    lib.mfh_hier_peaks.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   _c_size, ctypes.c_float, _c_size, _c_size, ctypes.c_float,
                                   _vp, _vp, _vp, _vp]
    lib.mfh_hier_peaks_rec.restype = ctypes.c_long
    lib.mfh_hier_peaks_rec.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       _c_size, ctypes.c_float, _c_size, _c_size, ctypes.c_float,
                                       _vp, _vp, _vp]
    lib.mfh_hier_peaks_async.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         _c_size, ctypes.c_float, _c_size, _c_size, ctypes.c_float,
                                         _vp, _vp, _vp]
    lib.mfh_host_alloc.restype = _vp
    lib.mfh_host_alloc.argtypes = [_c_size]
    lib.mfh_host_free.argtypes = [_vp]
    for name in ("mfh_run_series", "mfh_hier_run_series", "mfh_run_series_rec", "mfh_hier_run_series_rec"):
        getattr(lib, name).restype = ctypes.c_long
    lib.mfh_run_series.argtypes = [_vp, _vp, _c_size, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, _c_size, ctypes.c_float, _vp, _vp, _vp]
    lib.mfh_hier_run_series.argtypes = [_vp, _vp, _c_size, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, _c_size, ctypes.c_float,
                                        ctypes.c_float, _vp, _vp, _vp, _vp]
    lib.mfh_run_series_rec.argtypes = [_vp, _vp, _c_size, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, _c_size, ctypes.c_float, _vp, _vp]
    lib.mfh_hier_run_series_rec.argtypes = [_vp, _vp, _c_size, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, _c_size, ctypes.c_float,
                                            ctypes.c_float, _vp, _vp, _vp]
    lib.mfh_series_forward.argtypes = [_vp, _vp, _c_size, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int]
    lib.mfh_hier_set_coarse_precision.argtypes = [_vp, ctypes.c_int, ctypes.c_float]
    lib.mfh_hier_coarse_max.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        _c_size, _c_size, ctypes.c_int, _vp]
    lib.mfh_selftest_mfma.argtypes = [ctypes.c_int, ctypes.c_int, _vp, _vp, _vp]
    lib.mfh_probe_valu.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    lib.mfh_probe_copy.argtypes = [ctypes.c_int, _c_size, ctypes.POINTER(ctypes.c_double)]
    lib.mfh_stream_wait.argtypes = [_vp, _vp]
    lib.mfh_set_tuning.argtypes = [_vp, ctypes.c_char_p, ctypes.c_long]
    lib.mfh_prof_enable.argtypes = [_vp, ctypes.c_int]
    lib.mfh_prof_read.argtypes = [_vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_long)]
    lib.mfh_peak_grad_async.argtypes = [_vp, _vp, ctypes.c_longlong, ctypes.c_int, _vp, ctypes.c_longlong,
                                         ctypes.c_int, _c_size, _vp, _vp, _vp, _vp]
    lib.mfh_peak_values_async.argtypes = [_vp, _vp, ctypes.c_longlong, ctypes.c_int, _vp, ctypes.c_longlong,
                                          ctypes.c_int, _c_size, _vp, _vp]
    lib.mfh_queue_create.restype = _vp
    lib.mfh_queue_create.argtypes = [ctypes.c_int]
    lib.mfh_queue_destroy.restype = None
    lib.mfh_queue_destroy.argtypes = [_vp]
    lib.mfh_queue_set_pool.argtypes = [_vp, _vp, _c_size, ctypes.c_int]
    lib.mfh_queue_plan.restype = ctypes.c_longlong
    lib.mfh_queue_plan.argtypes = [_vp, ctypes.c_int] + [_vp] * 9
    lib.mfh_queue_run_async.argtypes = [_vp, _vp, _vp]
    lib.mfh_queue_run.restype = ctypes.c_longlong
    lib.mfh_queue_run.argtypes = [_vp, _vp, _vp, _vp]
    lib.mfh_queue_set_tuning.argtypes = [_vp, ctypes.c_char_p, ctypes.c_long]
    lib.mfh_queue_sync.argtypes = [_vp]
    lib.mfh_queue_stream.restype = _vp
    lib.mfh_queue_stream.argtypes = [_vp]
    lib.mfh_queue_prof_enable.argtypes = [_vp, ctypes.c_int]
    lib.mfh_queue_prof_read.argtypes = [_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long)]
    return lib


def _share_torch_hip_runtime():
    """Make libmfhip.so and PyTorch use ONE HIP runtime.

    PyTorch-ROCm bundles its own libamdhip64 and names it in DT_NEEDED as
    "libamdhip64.so"; libmfhip.so names the system runtime by its soname
    "libamdhip64.so.7".  Loaded in the other order the process ends up with
    two runtimes and PyTorch's finds no device.  Preloading PyTorch's copy
    (same soname) first makes both resolve to it.  Nothing happens when
    PyTorch is not installed.
    """
    import importlib.util
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        spec = None
    if spec is None or not spec.submodule_search_locations:
        return
    path = os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    if os.path.exists(path):
        ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


def lib():
    """The loaded HIP library; raises RuntimeError if it was never built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    "the HIP library %s is missing: build it with "
                    "`python -m apogee_amd._build` (hipcc, gfx950)" % LIB_PATH)
            _share_torch_hip_runtime()
            _lib = _declare(ctypes.CDLL(LIB_PATH))
        return _lib


def last_error():
    return lib().mfh_last_error().decode(errors="replace")


class _HostBlock:
    """One page-locked buffer of mfh_host_alloc, back to the pool when the
    last array over it is gone."""

    __slots__ = ("ptr", "nbytes")

    def __init__(self, ptr, nbytes):
        self.ptr, self.nbytes = ptr, nbytes

    def __del__(self):
        try:
            _POOL.give(self.ptr, self.nbytes)
        except Exception:       # interpreter shutdown: the process frees it
            pass


class _HostPool:
    """Free page-locked result buffers by size.  A records array handed to a
    caller owns its buffer until the array is gone, so results never alias;
    the pool keeps at most `keep_bytes` of free buffers."""

    def __init__(self, keep_bytes=256 << 20):
        self.free = {}
        self.free_bytes = 0
        self.keep_bytes = keep_bytes
        # reentrant: a collection inside give() (it allocates) may free a
        # result and call give() again on the same thread
        self.lock = threading.RLock()

    def take(self, nbytes):
        with self.lock:
            got = self.free.get(nbytes)
            if got:
                self.free_bytes -= nbytes
                return got.pop()
        ptr = lib().mfh_host_alloc(nbytes)
        return int(ptr) if ptr else None

    def give(self, ptr, nbytes):
        with self.lock:
            if self.free_bytes + nbytes <= self.keep_bytes:
                self.free.setdefault(nbytes, []).append(ptr)
                self.free_bytes += nbytes
                return
        lib().mfh_host_free(ctypes.c_void_p(ptr))


_POOL = _HostPool()
# larger results go through the context's one staging buffer and a host copy,
# so that each live result does not pin memory of its own beyond this
_POOL_MAX_BYTES = 64 << 20


def _records_array(shape, dtype):
    """An empty records array in page-locked memory the device writes
    directly (a plain NumPy array when none can be had)."""
    count = int(np.prod(shape, dtype=np.int64))
    nbytes = max(1, count * dtype.itemsize)
    ptr = _POOL.take(nbytes) if nbytes <= _POOL_MAX_BYTES else None
    if ptr is None:
        return np.empty(shape, dtype=dtype)
    buf = (ctypes.c_char * nbytes).from_address(ptr)
    buf._block = _HostBlock(ptr, nbytes)
    return np.frombuffer(buf, dtype=dtype, count=count).reshape(shape)


def _raw_arrays(shape):
    """Empty (index int64, value complex64) arrays back to back in one
    page-locked buffer the device writes directly (plain NumPy arrays when
    none can be had)."""
    count = int(np.prod(shape, dtype=np.int64))
    nbytes = max(1, count * 16)
    ptr = _POOL.take(nbytes) if nbytes <= _POOL_MAX_BYTES else None
    if ptr is None:
        return np.empty(shape, dtype=np.int64), np.empty(shape, dtype=np.complex64)
    buf = (ctypes.c_char * nbytes).from_address(ptr)
    buf._block = _HostBlock(ptr, nbytes)
    whole = np.frombuffer(buf, dtype=np.uint8, count=count * 16)
    return (whole[:count * 8].view(np.int64).reshape(shape),
            whole[count * 8:].view(np.complex64).reshape(shape))


def _record_dtype(dtype):
    """The records layout the library writes: an int64 at byte 0 and a
    complex64 at byte 8, 16 bytes in all (little-endian)."""
    dt = np.dtype(dtype)
    names = dt.names or ()
    ok = (dt.itemsize == 16 and len(names) == 2 and dt.fields[names[0]][:2] == (np.dtype("<i8"), 0)
          and dt.fields[names[1]][:2] == (np.dtype("<c8"), 8))
    if not ok:
        raise ValueError("records need an (int64, complex64) structured dtype of 16 bytes")
    return dt


def check(status, what="mfh"):
    if status < 0:
        raise RuntimeError("%s failed: %s" % (what, last_error()))
    return status


def device_count():
    return int(lib().mfh_device_count())


def device_info(index):
    name = ctypes.create_string_buffer(256)
    arch = ctypes.create_string_buffer(64)
    check(lib().mfh_device_info(int(index), name, 256, arch, 64), "mfh_device_info")
    return name.value.decode(), arch.value.decode()


def supported(n):
    return bool(lib().mfh_supported(int(n)))


def selftest_mfma(device, variant, A, B):
    """Run one MFMA variant on FP32 matrices (see include/mfhip.h)."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    C = np.empty((16, 16), dtype=np.float32)
    check(lib().mfh_selftest_mfma(int(device), int(variant), _ptr(A), _ptr(B), _ptr(C)),
          "mfh_selftest_mfma")
    return C


def probe_valu(device, op, chains, blocks=2048, iters=1024):
    """Measured FP32 rate of independent scalar chains (op "fma" or "add"):
    (flop per second, milliseconds)."""
    ms = ctypes.c_double(0.0)
    flop = ctypes.c_double(0.0)
    check(lib().mfh_probe_valu(int(device), {"fma": 0, "add": 1}[op], int(chains), int(blocks), int(iters),
                               ctypes.byref(ms), ctypes.byref(flop)), "mfh_probe_valu")
    return flop.value / (ms.value * 1e-3), ms.value


def probe_copy(device, nbytes=256 << 20):
    """Measured device-to-device copy rate in bytes per second (read + write)."""
    ms = ctypes.c_double(0.0)
    check(lib().mfh_probe_copy(int(device), int(nbytes), ctypes.byref(ms)), "mfh_probe_copy")
    return 2.0 * nbytes / (ms.value * 1e-3)


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _device_pointer(a):
    """Device address of a ROCm tensor (torch), or None for host arrays."""
    if hasattr(a, "data_ptr") and getattr(a, "is_cuda", False):
        return int(a.data_ptr())
    return None


# ------------------------------------------------------------------ DLPack
# The exchange structs of the DLPack standard (dlpack.h, v0.8 unversioned
# capsule "dltensor"): what a producer hands over through __dlpack__().
DLPACK_ROCM = 10
_DL_COMPLEX = 5


class _DLDevice(ctypes.Structure):
    _fields_ = [("device_type", ctypes.c_int32), ("device_id", ctypes.c_int32)]


class _DLDataType(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint8), ("bits", ctypes.c_uint8), ("lanes", ctypes.c_uint16)]


class _DLTensor(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("device", _DLDevice), ("ndim", ctypes.c_int32),
                ("dtype", _DLDataType), ("shape", ctypes.POINTER(ctypes.c_int64)),
                ("strides", ctypes.POINTER(ctypes.c_int64)), ("byte_offset", ctypes.c_uint64)]


class _DLManagedTensor(ctypes.Structure):
    _fields_ = [("dl_tensor", _DLTensor), ("manager_ctx", ctypes.c_void_p), ("deleter", ctypes.c_void_p)]


_capsule_pointer = ctypes.pythonapi.PyCapsule_GetPointer
_capsule_pointer.restype = ctypes.c_void_p
_capsule_pointer.argtypes = [ctypes.py_object, ctypes.c_char_p]


class DeviceView:
    """A ROCm buffer borrowed from a DLPack producer.

    ``capsule`` holds the producer's memory alive: while the view exists the
    buffer stays valid (the producer's deleter runs when the capsule goes).
    """
    __slots__ = ("ptr", "shape", "complex64", "compact", "device", "capsule")

    def __init__(self, ptr, shape, complex64, compact, device, capsule):
        self.ptr, self.shape, self.complex64 = ptr, shape, complex64
        self.compact, self.device, self.capsule = compact, device, capsule


def _torch_stream(torch, device):
    """PyTorch's current stream of ``device`` as a raw hipStream_t value."""
    raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
    if raw is not None:
        return int(raw(device))
    return int(torch.cuda.current_stream(device).cuda_stream)


def device_view(producer, stream):
    """Borrow ``producer``'s ROCm buffer for work queued on ``stream``.

    ``producer.__dlpack__(stream=stream)`` is the DLPack hand-over: the
    producer orders ``stream`` after its own pending writes (PyTorch records
    an event on its current stream and makes ``stream`` wait on it), so no
    host synchronisation is needed.  A PyTorch tensor is detached and its
    lazy conjugation resolved first (DLPack carries neither).
    """
    try:
        import torch
        if isinstance(producer, torch.Tensor):
            producer = producer.detach().resolve_conj()
    except ImportError:
        pass
    capsule = producer.__dlpack__(stream=int(stream))
    addr = _capsule_pointer(capsule, b"dltensor")
    t = _DLManagedTensor.from_address(addr).dl_tensor
    if t.device.device_type != DLPACK_ROCM:
        raise TypeError("DLPack producer handed over device type %d, not ROCm" % t.device.device_type)
    shape = tuple(int(t.shape[i]) for i in range(t.ndim))
    compact = True
    if t.strides:
        expect = 1
        for axis in range(t.ndim - 1, -1, -1):
            if shape[axis] != 1 and int(t.strides[axis]) != expect:
                compact = False
            expect *= shape[axis]
    c64 = (t.dtype.code, t.dtype.bits, t.dtype.lanes) == (_DL_COMPLEX, 64, 1)
    return DeviceView(int(t.data or 0) + int(t.byte_offset), shape, c64, compact,
                      int(t.device.device_id), capsule)


class Context:
    """Device state of one filter: spectra resident in HBM plus a HIP stream.

    ``n``, ``ndata`` and ``ntemplates`` fix the device allocation; rows are
    written by :meth:`set_rows` and read by every run until overwritten.
    """

    def __init__(self, index, n, ndata, ntemplates):
        L = lib()
        if not supported(n):
            raise UnsupportedSize("transform length %d is not supported on this device" % n)
        self.index = int(index)
        self.n, self.ndata, self.ntemplates = int(n), int(ndata), int(ntemplates)
        h = L.mfh_create(self.index, self.n, self.ndata, self.ntemplates)
        if not h:
            raise RuntimeError("mfh_create failed: %s" % last_error())
        self._h = ctypes.c_void_p(h)
        self.last_refinements = 0
        self._bound = {}

    # ------------------------------------------------------------ lifetime
    def destroy(self):
        h, self._h = getattr(self, "_h", None), None
        if h and _lib is not None:
            _lib.mfh_destroy(h)

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    @property
    def handle(self):
        if not self._h:
            raise RuntimeError("context destroyed")
        return self._h

    def stream(self):
        return int(lib().mfh_stream(self.handle) or 0)

    def sync(self):
        check(lib().mfh_sync(self.handle), "mfh_sync")

    def wait_for(self, stream):
        """Order this context's stream after the work queued on ``stream``
        (a raw hipStream_t value) so far, with no host wait."""
        check(lib().mfh_stream_wait(self.handle, ctypes.c_void_p(int(stream))), "mfh_stream_wait")

    # -------------------------------------------------------------- ingest
    def set_rows(self, kind, row0, spectra):
        """Copy host spectra, a C-contiguous complex64 array (rows, n), into
        data (kind="data") or template rows from row0; returns when the copy
        is done."""
        L = lib()
        fn = L.mfh_set_data if kind == "data" else L.mfh_set_templates
        check(fn(self.handle, int(row0), int(spectra.shape[0]), _ptr(spectra)), "set %s" % kind)
        self._bound.pop(kind, None)

    def copy_rows(self, kind, row0, view):
        """Copy a borrowed ROCm buffer (:class:`DeviceView`, rows x n
        complex64) into rows from row0, device to device; returns when the
        copy is done, so the producer may reuse its buffer."""
        L = lib()
        fn = L.mfh_set_data_device_async if kind == "data" else L.mfh_set_templates_device_async
        rows = int(np.prod(view.shape, dtype=np.int64)) // self.n
        check(fn(self.handle, int(row0), rows, ctypes.c_void_p(view.ptr)), "set %s" % kind)
        self.sync()
        self._bound.pop(kind, None)

    def bind_rows(self, kind, view):
        """Read the whole data or template bank in place from a borrowed ROCm
        buffer (:class:`DeviceView` of shape (rows, n), complex64, compact).
        Nothing is copied; the view is kept until the bank is replaced, and
        the library never writes the buffer (a later single-row update first
        copies the bank into the context's own storage)."""
        L = lib()
        fn = L.mfh_bind_data if kind == "data" else L.mfh_bind_templates
        check(fn(self.handle, ctypes.c_void_p(view.ptr)), "bind %s" % kind)
        self._bound[kind] = view

    def borrow(self, producer):
        """:func:`device_view` of ``producer`` ordered on this context's stream.

        A PyTorch tensor takes the short path of the same hand-over: the
        context's stream waits on an event recorded on PyTorch's current
        stream of the tensor's device (``mfh_stream_wait``), and the view
        keeps the tensor alive -- what ``__dlpack__(stream=...)`` does,
        without building and parsing a capsule on every call.
        """
        try:
            import torch
        except ImportError:
            torch = None
        if torch is not None and isinstance(producer, torch.Tensor) and producer.is_cuda:
            t = producer.detach() if producer.requires_grad else producer
            if t.is_conj():
                t = t.resolve_conj()
            dev = t.get_device()
            self.wait_for(_torch_stream(torch, dev))
            return DeviceView(int(t.data_ptr()), tuple(t.shape), t.dtype == torch.complex64,
                              t.is_contiguous(), dev, t)
        return device_view(producer, self.stream())

    def bind_tensor(self, kind, tensor):
        """Bind a whole bank to a contiguous complex64 PyTorch tensor of
        shape (rows, n) in place.  The caller has finished the tensor's
        producer (a host wait on its stream); the tensor is kept until the
        bank is replaced."""
        L = lib()
        fn = L.mfh_bind_data if kind == "data" else L.mfh_bind_templates
        check(fn(self.handle, ctypes.c_void_p(int(tensor.data_ptr()))), "bind %s" % kind)
        self._bound[kind] = tensor

    # ---------------------------------------------------------------- runs
    def peaks(self, d0, nd, t0, nt, binsize, threshold, start, end, counts=False, records=None):
        """Flat filter over rows [d0, d0+nd) x [t0, t0+nt); host results.

        Returns ``(index int64, value complex64)`` of shape (nd, nt, nbins), and
        the per-pair crossing counts (int32, (nd, nt)) when ``counts``.  With
        ``records`` (a structured dtype of an int64 and a complex64 field,
        16 bytes) the first two are replaced by one array of that dtype,
        filled by the library, and None.
        """
        nb = int(lib().mfh_nbins(self.n, int(binsize), int(start), int(end)))
        cnt = np.empty((nd, nt), dtype=np.int32) if counts else None
        args = (self.handle, int(d0), int(nd), int(t0), int(nt), int(binsize), float(threshold), int(start),
                int(end))
        if records is not None:
            rec = _records_array((nd, nt, nb), _record_dtype(records))
            check(lib().mfh_flat_peaks_rec(*args, _ptr(rec), _ptr(cnt) if cnt is not None else None),
                  "mfh_flat_peaks_rec")
            return rec, None, cnt
        idx, val = _raw_arrays((nd, nt, nb))
        check(lib().mfh_flat_peaks(*args, _ptr(idx), _ptr(val), _ptr(cnt) if cnt is not None else None),
              "mfh_flat_peaks")
        return idx, val, cnt

    def peaks_async(self, d0, nd, t0, nt, binsize, threshold, start, end, idx_ptr, val_ptr):
        """Queue a flat run writing device buffers (int64 / complex64 pointers)."""
        check(lib().mfh_flat_peaks_async(self.handle, int(d0), int(nd), int(t0), int(nt), int(binsize),
                                         float(threshold), int(start), int(end),
                                         ctypes.c_void_p(int(idx_ptr)), ctypes.c_void_p(int(val_ptr))),
              "mfh_flat_peaks_async")

    # ---------------------------------------------------------- hierarchy
    def hier_set_band(self, band, ref_fraction=-1.0):
        """Coarse band and normalisation (bank-wide fraction, or < 0 per template)."""
        check(lib().mfh_hier_set_band(self.handle, int(band), float(ref_fraction)), "mfh_hier_set_band")

    def hier_peaks(self, d0, nd, t0, nt, binsize, threshold, start, end, coarse_threshold, counts=False,
                   records=None):
        """Hierarchical run; returns (index, value, counts or None, refined
        pairs), or with ``records`` (as in :meth:`peaks`) (records, None,
        counts or None, refined pairs)."""
        nb = int(lib().mfh_nbins(self.n, int(binsize), int(start), int(end)))
        cnt = np.empty((nd, nt), dtype=np.int32) if counts else None
        refined = ctypes.c_int64(0)
        args = (self.handle, int(d0), int(nd), int(t0), int(nt), int(binsize), float(threshold), int(start),
                int(end), float(coarse_threshold))
        if records is not None:
            rec = _records_array((nd, nt, nb), _record_dtype(records))
            check(lib().mfh_hier_peaks_rec(*args, _ptr(rec), _ptr(cnt) if cnt is not None else None,
                                           ctypes.byref(refined)), "mfh_hier_peaks_rec")
            first, second = rec, None
        else:
            idx, val = _raw_arrays((nd, nt, nb))
            check(lib().mfh_hier_peaks(*args, _ptr(idx), _ptr(val), _ptr(cnt) if cnt is not None else None,
                                       ctypes.byref(refined)), "mfh_hier_peaks")
            first, second = idx, val
        self.last_refinements = int(refined.value)
        return first, second, cnt, int(refined.value)

    def hier_set_coarse_precision(self, precision, delta):
        """0 = FP32 screen, 1 = fp16 MFMA, 2 = bf16 MFMA with gate widening delta."""
        check(lib().mfh_hier_set_coarse_precision(self.handle, int(precision), float(delta)),
              "mfh_hier_set_coarse_precision")

    def hier_coarse_max(self, d0, nd, t0, nt, start, end, precision):
        """Coarse statistic max |z_c| per pair (nd, nt) at a precision."""
        out = np.empty((nd, nt), dtype=np.float32)
        check(lib().mfh_hier_coarse_max(self.handle, int(d0), int(nd), int(t0), int(nt), int(start),
                                        int(end), int(precision), _ptr(out)), "mfh_hier_coarse_max")
        return out

    def hier_peaks_async(self, d0, nd, t0, nt, binsize, threshold, start, end, coarse_threshold,
                         idx_ptr, val_ptr):
        """Queue a hierarchical run writing device buffers."""
        check(lib().mfh_hier_peaks_async(self.handle, int(d0), int(nd), int(t0), int(nt), int(binsize),
                                         float(threshold), int(start), int(end), float(coarse_threshold),
                                         ctypes.c_void_p(int(idx_ptr)), ctypes.c_void_p(int(val_ptr)), None),
              "mfh_hier_peaks_async")

    # ------------------------------------------------------------ series
    def run_series(self, series, starts, win_start, win_end, t0, nt, binsize, threshold,
                   coarse_threshold=None, records=None):
        """Block spectra of a series on the device, then the flat (or, with a
        coarse threshold, hierarchical) filter; returns (index, value, refined),
        or (records, None, refined) when `records` names the record dtype."""
        view = self.borrow(series) if _device_pointer(series) is not None else None
        if view is not None:
            sp, ns, on_dev = ctypes.c_void_p(view.ptr), int(view.shape[0]), 1
        else:
            sp, ns, on_dev = _ptr(series), int(series.size), 0
        st = np.ascontiguousarray(starts, dtype=np.uintp)
        ws = np.ascontiguousarray(win_start, dtype=np.uintp)
        we = np.ascontiguousarray(win_end, dtype=np.uintp)
        nblk = int(st.size)
        nb = int(lib().mfh_nbins(self.n, int(binsize), int(ws[0]), int(we[0])))
        refined = ctypes.c_int64(0)
        if records is not None:
            rec = _records_array((nblk, nt, nb), _record_dtype(records))
            if coarse_threshold is None:
                check(lib().mfh_run_series_rec(self.handle, sp, ns, on_dev, _ptr(st), _ptr(ws), _ptr(we), nblk,
                                               int(t0), int(nt), int(binsize), float(threshold), _ptr(rec),
                                               None), "mfh_run_series_rec")
            else:
                check(lib().mfh_hier_run_series_rec(self.handle, sp, ns, on_dev, _ptr(st), _ptr(ws), _ptr(we),
                                                    nblk, int(t0), int(nt), int(binsize), float(threshold),
                                                    float(coarse_threshold), _ptr(rec), None,
                                                    ctypes.byref(refined)), "mfh_hier_run_series_rec")
            self.last_refinements = int(refined.value)
            return rec, None, int(refined.value)
        idx, val = _raw_arrays((nblk, nt, nb))
        if coarse_threshold is None:
            check(lib().mfh_run_series(self.handle, sp, ns, on_dev, _ptr(st), _ptr(ws), _ptr(we), nblk,
                                       int(t0), int(nt), int(binsize), float(threshold),
                                       _ptr(idx), _ptr(val), None), "mfh_run_series")
        else:
            check(lib().mfh_hier_run_series(self.handle, sp, ns, on_dev, _ptr(st), _ptr(ws), _ptr(we),
                                            nblk, int(t0), int(nt), int(binsize), float(threshold),
                                            float(coarse_threshold), _ptr(idx), _ptr(val), None,
                                            ctypes.byref(refined)), "mfh_hier_run_series")
        self.last_refinements = int(refined.value)
        return idx, val, int(refined.value)

    # ------------------------------------------------------- derivatives
    def peak_grad_async(self, data_ptr, nd, tmpl_ptr, nt, nb, idx_ptr, gval_ptr, gdata_ptr, gtmpl_ptr):
        """Queue dL/dD and/or dL/dT (device pointers; a null output is skipped)."""
        vp = lambda p: ctypes.c_void_p(int(p)) if p else None
        check(lib().mfh_peak_grad_async(self.handle, vp(data_ptr), self.n, int(nd), vp(tmpl_ptr), self.n,
                                        int(nt), int(nb), vp(idx_ptr), vp(gval_ptr), vp(gdata_ptr),
                                        vp(gtmpl_ptr)), "mfh_peak_grad_async")

    def peak_values_async(self, data_ptr, nd, tmpl_ptr, nt, nb, idx_ptr, out_ptr):
        """Queue z at stored lags (device pointers)."""
        vp = lambda p: ctypes.c_void_p(int(p))
        check(lib().mfh_peak_values_async(self.handle, vp(data_ptr), self.n, int(nd), vp(tmpl_ptr), self.n,
                                          int(nt), int(nb), vp(idx_ptr), vp(out_ptr)),
              "mfh_peak_values_async")

    def series_forward(self, series, starts, row0=0):
        """Block spectra of a series into data rows [row0, row0 + len(starts))."""
        view = self.borrow(series) if _device_pointer(series) is not None else None
        if view is not None:
            sp, ns, on_dev = ctypes.c_void_p(view.ptr), int(view.shape[0]), 1
        else:
            sp, ns, on_dev = _ptr(series), int(series.size), 0
        st = np.ascontiguousarray(starts, dtype=np.uintp)
        check(lib().mfh_series_forward(self.handle, sp, ns, on_dev, _ptr(st), int(st.size), int(row0)),
              "mfh_series_forward")
        self._bound.pop("data", None)      # the library wrote data rows of its own copy

    # ------------------------------------------------------------- tuning
    def set_tuning(self, key, value):
        check(lib().mfh_set_tuning(self.handle, key.encode(), int(value)), "mfh_set_tuning")

    def prof_enable(self, on=True):
        check(lib().mfh_prof_enable(self.handle, 1 if on else 0), "mfh_prof_enable")

    def prof_read(self, name):
        ms = ctypes.c_double(0.0)
        cnt = ctypes.c_long(0)
        check(lib().mfh_prof_read(self.handle, name.encode(), ctypes.byref(ms), ctypes.byref(cnt)),
              "mfh_prof_read")
        return ms.value, cnt.value


class Queue:
    """Device side of :class:`apogee_amd.refine.RefineQueue`: a spectrum pool,
    a planned work queue and the persistent refine launch that drains it."""

    def __init__(self, index):
        L = lib()
        self.index = int(index)
        h = L.mfh_queue_create(self.index)
        if not h:
            raise RuntimeError("mfh_queue_create failed: %s" % last_error())
        self._h = ctypes.c_void_p(h)
        self._pool_ref = None

    def destroy(self):
        h, self._h = getattr(self, "_h", None), None
        if h and _lib is not None:
            _lib.mfh_queue_destroy(h)

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    @property
    def handle(self):
        if not self._h:
            raise RuntimeError("queue destroyed")
        return self._h

    def set_pool(self, pool):
        """pool: 1-D complex64 host array (copied) or ROCm tensor (referenced)."""
        if _device_pointer(pool) is not None:
            view = device_view(pool, self.stream())
            check(lib().mfh_queue_set_pool(self.handle, ctypes.c_void_p(view.ptr), int(view.shape[0]), 1),
                  "mfh_queue_set_pool")
            self._pool_ref = view     # the queue reads it in place
        else:
            check(lib().mfh_queue_set_pool(self.handle, _ptr(pool), int(pool.size), 0),
                  "mfh_queue_set_pool")
            self._pool_ref = None

    def plan(self, n, data_off, tmpl_off, ncand, binsize, threshold, ws, we):
        """Returns (per-request output offsets int64, total output elements)."""
        arrs = [np.ascontiguousarray(a, dtype=np.int64) for a in (n, data_off, tmpl_off)]
        nc = np.ascontiguousarray(ncand, dtype=np.int32)
        bs = np.ascontiguousarray(binsize, dtype=np.int64)
        th = np.ascontiguousarray(threshold, dtype=np.float32)
        w0 = np.ascontiguousarray(ws, dtype=np.int64)
        w1 = np.ascontiguousarray(we, dtype=np.int64)
        nreq = int(nc.size)
        off = np.empty(nreq, dtype=np.int64)
        total = lib().mfh_queue_plan(self.handle, nreq, _ptr(arrs[0]), _ptr(arrs[1]), _ptr(arrs[2]),
                                     _ptr(nc), _ptr(bs), _ptr(th), _ptr(w0), _ptr(w1), _ptr(off))
        check(total, "mfh_queue_plan")
        return off, int(total)

    def run(self, nout, npairs):
        """Host results: (index int64 [nout], value complex64 [nout], counts int32 [npairs])."""
        idx = np.empty(max(nout, 1), dtype=np.int64)
        val = np.empty(max(nout, 1), dtype=np.complex64)
        cnt = np.empty(max(npairs, 1), dtype=np.int32)
        check(lib().mfh_queue_run(self.handle, _ptr(idx), _ptr(val), _ptr(cnt)), "mfh_queue_run")
        return idx[:nout], val[:nout], cnt[:npairs]

    def run_async(self, idx_ptr, val_ptr):
        check(lib().mfh_queue_run_async(self.handle, ctypes.c_void_p(int(idx_ptr)),
                                        ctypes.c_void_p(int(val_ptr))), "mfh_queue_run_async")

    def sync(self):
        check(lib().mfh_queue_sync(self.handle), "mfh_queue_sync")

    def stream(self):
        return int(lib().mfh_queue_stream(self.handle) or 0)

    def set_tuning(self, key, value):
        check(lib().mfh_queue_set_tuning(self.handle, key.encode(), int(value)), "mfh_queue_set_tuning")

    def prof_enable(self, on=True):
        check(lib().mfh_queue_prof_enable(self.handle, 1 if on else 0), "mfh_queue_prof_enable")

    def prof_read(self):
        ms = ctypes.c_double(0.0)
        cnt = ctypes.c_long(0)
        check(lib().mfh_queue_prof_read(self.handle, ctypes.byref(ms), ctypes.byref(cnt)),
              "mfh_queue_prof_read")
        return ms.value, cnt.value
