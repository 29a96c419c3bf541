"""Padded device arrays and small helpers (replaces GPUSimulators/Common.py:215-235,
436-487, 498-616, 741-779 for the hot path).

``HipArray2D`` keeps the layout and methods of ``CudaArray2D``: a zero
initialised ``(ny+2*y_halo, nx+2*x_halo)`` float32 buffer (pitched), interior
upload at construction from a ``(ny, nx)`` or padded array, and ``download`` /
``upload`` of an ``[x, y, w, h]`` extent.  ``ArakawaA2D`` groups the fields of
one state.  Unlike the reference, ``download`` always synchronises the stream
before returning the host array.
"""

import ctypes
import logging
import time

import numpy as np

from . import _lib


class Timer:
    """Context manager that logs the elapsed milliseconds (Common.py:215-235)."""

    def __init__(self, tag, log_level=logging.DEBUG):
        self.tag = tag
        self.log_level = log_level
        self.logger = logging.getLogger(__name__)

    def __enter__(self):
        self.start = time.time()
        return self

    def __exit__(self, *args):
        self.end = time.time()
        self.secs = self.end - self.start
        self.msecs = self.secs * 1000
        self.logger.log(self.log_level, "%s: %f ms", self.tag, self.msecs)

    def elapsed(self):
        return time.time() - self.start


class ProgressPrinter:
    """Returns a progress string at most every ``print_every`` seconds (Common.py:436-487)."""

    def __init__(self, total_steps, print_every=5):
        self.total_steps = total_steps
        self.print_every = print_every
        self.next_print_time = print_every
        self.start = time.time()

    def getPrintString(self, step):
        elapsed = time.time() - self.start
        if elapsed < self.next_print_time:
            return None
        self.next_print_time = elapsed + self.print_every
        frac = step / self.total_steps if self.total_steps else 1.0
        eta = elapsed / frac - elapsed if frac > 0 else float("nan")
        return "%5.1f%% elapsed %.1fs, remaining %.1fs" % (100.0 * frac, elapsed, eta)


def _stream_handle(stream):
    return stream.handle if stream is not None else None


class HipArray2D:
    """Pitched, zero-initialised 2-D float32 device buffer with ghost cells."""

    def __init__(self, stream, nx, ny, x_halo, y_halo, cpu_data=None, dtype=np.float32):
        self.logger = logging.getLogger(__name__)
        if np.dtype(dtype) != np.float32:
            raise TypeError("only float32 fields are supported")
        self.nx = int(nx)
        self.ny = int(ny)
        self.x_halo = int(x_halo)
        self.y_halo = int(y_halo)
        nx_halo = self.nx + 2 * self.x_halo
        ny_halo = self.ny + 2 * self.y_halo
        self.shape = (ny_halo, nx_halo)
        ptr = ctypes.c_void_p()
        pitch = ctypes.c_size_t()
        _lib.call("fvmi_malloc_pitch", ctypes.byref(ptr), ctypes.byref(pitch), nx_halo * 4, ny_halo)
        self.ptr = ptr
        self.pitch = int(pitch.value)
        _lib.call("fvmi_fill_f32", self.ptr, self.pitch, nx_halo, ny_halo, 0.0, _stream_handle(stream))
        if cpu_data is None:
            return
        cpu_data = np.asarray(cpu_data)
        if cpu_data.shape not in ((ny_halo, nx_halo), (self.ny, self.nx)):
            raise ValueError("Wrong shape of data %s vs %s / %s"
                             % (str(cpu_data.shape), str((self.ny, self.nx)), str((ny_halo, nx_halo))))
        if cpu_data.itemsize != 4:
            raise TypeError("Wrong size of data type")
        if np.isfortran(cpu_data):
            raise TypeError("Wrong datatype (Fortran, expected C)")
        x = (nx_halo - cpu_data.shape[1]) // 2
        y = (ny_halo - cpu_data.shape[0]) // 2
        self.upload(stream, cpu_data, extent=[x, y, cpu_data.shape[1], cpu_data.shape[0]])
        if stream is not None:
            stream.synchronize()

    @property
    def data_ptr(self):
        return self.ptr.value

    def row_ptr(self, row, col=0):
        return self.ptr.value + row * self.pitch + col * 4

    def download(self, stream, cpu_data=None, asynch=False, extent=None):
        if extent is None:
            x, y, nx, ny = self.x_halo, self.y_halo, self.nx, self.ny
        else:
            x, y, nx, ny = (int(v) for v in extent)
        if cpu_data is None:
            cpu_data = np.empty((ny, nx), dtype=np.float32)
        if cpu_data.shape != (ny, nx) or cpu_data.dtype != np.float32 or not cpu_data.flags.c_contiguous:
            raise ValueError("download target must be a C-contiguous float32 (%d, %d) array" % (ny, nx))
        if x < 0 or y < 0 or x + nx > self.shape[1] or y + ny > self.shape[0]:
            raise ValueError("extent outside the buffer")
        _lib.call("fvmi_memcpy2d_async", cpu_data.ctypes.data, nx * 4, self.row_ptr(y, x), self.pitch, nx * 4, ny,
                  _lib.MEMCPY_DTOH, _stream_handle(stream))
        # The reference can return before the copy has finished (Common.py:767-769);
        # here the copy is always complete on return, whatever ``asynch`` says.
        if stream is not None:
            stream.synchronize()
        else:
            _lib.call("fvmi_stream_synchronize", None)
        return cpu_data

    def upload(self, stream, cpu_data, extent=None):
        if extent is None:
            x, y, nx, ny = self.x_halo, self.y_halo, self.nx, self.ny
        else:
            x, y, nx, ny = (int(v) for v in extent)
        src = np.ascontiguousarray(cpu_data, dtype=np.float32)
        if src.shape != (ny, nx):
            raise ValueError("upload source shape %s does not match extent %s" % (src.shape, (ny, nx)))
        if x < 0 or y < 0 or x + nx > self.shape[1] or y + ny > self.shape[0]:
            raise ValueError("extent outside the buffer")
        self._keepalive = src
        _lib.call("fvmi_memcpy2d_async", self.row_ptr(y, x), self.pitch, src.ctypes.data, nx * 4, nx * 4, ny,
                  _lib.MEMCPY_HTOD, _stream_handle(stream))
        if stream is not None:
            stream.synchronize()

    def free(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value and _lib._lib is not None:
            _lib.lib().fvmi_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        self.free()


class ArakawaA2D:
    """The fields of one unstaggered grid state (Common.py:741-779)."""

    def __init__(self, stream, nx, ny, halo_x, halo_y, cpu_variables):
        self.logger = logging.getLogger(__name__)
        self.gpu_variables = [HipArray2D(stream, nx, ny, halo_x, halo_y, v) for v in cpu_variables]

    def __getitem__(self, key):
        if not isinstance(key, int):
            raise TypeError("Indexing is int based")
        if key >= len(self.gpu_variables) or key < 0:
            raise IndexError("Out of bounds")
        return self.gpu_variables[key]

    def __len__(self):
        return len(self.gpu_variables)

    def download(self, stream, variables=None):
        if variables is None:
            variables = range(len(self.gpu_variables))
        out = []
        for i in variables:
            if i >= len(self.gpu_variables):
                raise IndexError("Variable {:d} is out of range".format(i))
            out.append(self.gpu_variables[i].download(stream, asynch=True))
        if stream is not None:
            stream.synchronize()
        return out

    def check(self, stream=None):
        """Asserts that no field holds a non-finite interior value (Common.py:772-779,
        which asserts on NaN in the device sum)."""
        scratch = ctypes.c_void_p()
        _lib.call("fvmi_malloc", ctypes.byref(scratch), 16)
        try:
            for i, v in enumerate(self.gpu_variables):
                _lib.call("fvmi_fill_f32", scratch, 16, 4, 1, 0.0, _stream_handle(stream))
                _lib.call("fvmi_field_stats", v.nx, v.ny, v.row_ptr(0), v.pitch, scratch.value,
                          scratch.value + 8, _stream_handle(stream))
                host = np.zeros(2, dtype=np.float64)
                _lib.call("fvmi_memcpy_async", host.ctypes.data, scratch, 16, _lib.MEMCPY_DTOH, _stream_handle(stream))
                if stream is not None:
                    stream.synchronize()
                else:
                    _lib.call("fvmi_stream_synchronize", None)
                var_sum = float(host[0])
                bad = int(host[1:2].view(np.uint32)[0])
                self.logger.debug("Data %d with size [%d x %d] has average %f", i, v.nx, v.ny,
                                  var_sum / (v.nx * v.ny))
                assert bad == 0, "Data contains NaN values!"
        finally:
            _lib.lib().fvmi_free(scratch)


# Drop-in name for code written against the reference.
CudaArray2D = HipArray2D
