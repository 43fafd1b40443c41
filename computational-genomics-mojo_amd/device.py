"""GenomicsDevice: device context, buffers and uploads over libcgm_align.so.

Mirrors genomics/gpu/device.mojo:21-138 of the reference (a DeviceContext
wrapper with typed allocation and SequenceBatch uploads), adding the int32
download the reference lacks.  Buffers are DeviceArray objects that own a
hipMalloc'ed region; streams are HIP streams created by the library.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


class DeviceArray:
    """A typed device buffer (the reference's DeviceBuffer[dtype])."""

    def __init__(self, count: int, dtype):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        p = ctypes.c_void_p()
        _lib.call("cgm_malloc", ctypes.byref(p), self.nbytes)
        self.dptr = p.value
        self._freed = False

    def free(self) -> None:
        if not self._freed and self.dptr:
            _lib.load().cgm_free(self.dptr)
        self._freed = True

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __len__(self):
        return self.count


class GenomicsDevice:
    """device.mojo:21-29 -- one device, one stream."""

    def __init__(self, device: int = 0):
        n = ctypes.c_int32()
        _lib.call("cgm_device_count", ctypes.byref(n))
        if n.value <= device:
            raise _lib.CgmError(_lib.CGM_ERR_NODEV, f"no GPU {device} (found {n.value})")
        _lib.call("cgm_set_device", device)
        self.device = device
        s = ctypes.c_void_p()
        _lib.call("cgm_stream_create", ctypes.byref(s))
        self.stream = s.value

    def close(self) -> None:
        if self.stream:
            _lib.load().cgm_stream_synchronize(self.stream)
            _lib.load().cgm_stream_destroy(self.stream)
            self.stream = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        name = ctypes.create_string_buffer(256)
        cu, clk, lds = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        hbm = ctypes.c_int64()
        _lib.call("cgm_device_info", self.device, name, 256, ctypes.byref(cu), ctypes.byref(clk),
                  ctypes.byref(hbm), ctypes.byref(lds))
        return {"name": name.value.decode(), "compute_units": cu.value,
                "max_clock_mhz": clk.value / 1000.0, "hbm_bytes": hbm.value,
                "lds_bytes_per_cu": lds.value}

    def synchronize(self) -> None:
        _lib.call("cgm_stream_synchronize", self.stream)

    # typed allocation (device.mojo:38-54)
    def alloc(self, count: int, dtype) -> DeviceArray:
        return DeviceArray(count, dtype)

    def alloc_uint64(self, count: int) -> DeviceArray:
        return DeviceArray(count, np.uint64)

    def alloc_int32(self, count: int) -> DeviceArray:
        return DeviceArray(count, np.int32)

    # copies
    def upload(self, host: np.ndarray, dtype=None) -> DeviceArray:
        a = np.ascontiguousarray(host, dtype=dtype or host.dtype)
        d = DeviceArray(max(a.size, 1), a.dtype)
        if a.size:
            _lib.call("cgm_memcpy_htod", d.dptr, _lib.ptr(a), a.nbytes, self.stream)
            self.synchronize()
        return d

    def download(self, d: DeviceArray, count: int | None = None) -> np.ndarray:
        n = d.count if count is None else int(count)
        out = np.empty(n, dtype=d.dtype)
        if n:
            _lib.call("cgm_memcpy_dtoh", _lib.ptr(out), d.dptr, n * d.dtype.itemsize, self.stream)
            self.synchronize()
        return out

    def download_int32(self, d: DeviceArray, count: int | None = None) -> np.ndarray:
        return self.download(d, count)

    # SequenceBatch uploads (device.mojo:60-108)
    def upload_batch_packed(self, batch) -> DeviceArray:
        return self.upload(batch.packed, np.uint64)

    def upload_lengths(self, batch) -> DeviceArray:
        return self.upload(batch.lengths, np.int32)

    def upload_offsets(self, batch) -> DeviceArray:
        return self.upload(batch.offsets, np.int32)


class Timer:
    """HIP-event timer on a stream: elapsed milliseconds between start/stop."""

    def __init__(self, stream):
        self.stream = stream
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        _lib.call("cgm_event_create", ctypes.byref(a))
        _lib.call("cgm_event_create", ctypes.byref(b))
        self._a, self._b = a.value, b.value

    def start(self):
        _lib.call("cgm_event_record", self._a, self.stream)

    def stop(self):
        _lib.call("cgm_event_record", self._b, self.stream)

    def elapsed_ms(self) -> float:
        _lib.call("cgm_event_synchronize", self._b)
        ms = ctypes.c_float()
        _lib.call("cgm_event_elapsed_ms", self._a, self._b, ctypes.byref(ms))
        return float(ms.value)

    def __del__(self):
        try:
            _lib.load().cgm_event_destroy(self._a)
            _lib.load().cgm_event_destroy(self._b)
        except Exception:
            pass
