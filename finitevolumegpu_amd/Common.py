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
import json
import logging
import os
import subprocess
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
                _lib.call("fvmi_field_stats", v.nx, v.ny, v.x_halo, v.y_halo, v.row_ptr(0), v.pitch,
                          scratch.value, scratch.value + 8, _stream_handle(stream))
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


# ---------------------------------------------------------------------------
# Simulation output (Common.py:48-208, 361-431)
# ---------------------------------------------------------------------------


def safeCall(cmd):
    """Output of ``cmd`` run in this package's folder (Common.py:48-66).  Unlike
    the reference, a failing command (no git, no repository) returns "" and
    logs a warning instead of raising, so output files can be written from an
    installed tree."""
    logger = logging.getLogger(__name__)
    try:
        return subprocess.check_output(cmd, stderr=subprocess.STDOUT, cwd=os.path.dirname(os.path.realpath(__file__)),
                                       universal_newlines=True, shell=False)
    except (OSError, subprocess.CalledProcessError) as e:
        logger.warning("%s failed: %s", " ".join(cmd), e)
        return ""


def getGitHash():
    return safeCall(["git", "rev-parse", "HEAD"])


def getGitStatus():
    return safeCall(["git", "status", "--porcelain", "-uno"])


def toJson(in_dict, compressed=True):
    """JSON text of a dict: arrays as lists, anything else JSON cannot encode as
    its str() (Common.py:74-90)."""
    logger = logging.getLogger(__name__)
    out = dict(in_dict)
    for key, value in out.items():
        if isinstance(value, np.ndarray):
            out[key] = value.tolist()
        else:
            try:
                json.dumps(value)
            except (TypeError, ValueError):
                out[key] = str(value)
                logger.warning("JSON: Converting %s to string (%s)", key, out[key])
    return json.dumps(out)


def quantize(data, least_significant_digit):
    """The rounding netCDF4-python applies to a variable created with
    ``least_significant_digit=d`` (its ``_quantize``; the reference's
    environment files do not pin a version): keep
    ``bits = ceil(log2(10**-e))`` binary digits after the point, with e =
    floor(log10(10**-d)) (ceil for d < 0), i.e. ``around(2**bits * x) / 2**bits``
    -- for d = 3, multiples of 1/1024.  float32 input stays float32 (the scaling
    by a power of two is exact)."""
    data = np.asarray(data)
    exp = np.log10(10.0 ** -least_significant_digit)
    exp = int(np.floor(exp)) if exp < 0 else int(np.ceil(exp))
    scale = 2.0 ** np.ceil(np.log2(10.0 ** -exp))
    if data.dtype == np.float32:
        scale = np.float32(scale)
    return np.around(scale * data) / scale


class _Netcdf3Variable:
    """One variable of a :class:`Netcdf3Dataset`; writes are rounded to the
    variable's ``least_significant_digit`` as netCDF4-python rounds them."""

    def __init__(self, var, least_significant_digit=None):
        self._var = var
        self._lsd = least_significant_digit

    def __setitem__(self, index, value):
        value = np.asarray(value)
        if self._lsd is not None:
            value = quantize(value.astype(self._var.data.dtype.newbyteorder("="), copy=False), self._lsd)
        self._var[index] = value

    def __getitem__(self, index):
        return self._var[index]

    @property
    def shape(self):
        return self._var.shape

    @property
    def dimensions(self):
        return self._var.dimensions

    def __getattr__(self, name):
        return getattr(self._var, name)


class Netcdf3Dataset:
    """The part of ``netCDF4.Dataset`` that the output path uses
    (createDimension, createVariable with zlib / least_significant_digit,
    variables, dimensions, global attributes, close) over scipy's netCDF-3
    writer, for environments without netCDF4: a 64-bit-offset netCDF-3 file
    that netCDF4 and scipy both read.  netCDF-3 has no compression, so ``zlib``
    is accepted and ignored; ``least_significant_digit`` rounds the data and is
    recorded as the variable attribute netCDF4 writes."""

    def __init__(self, filename, mode="r", clobber=True, format=None, **kwargs):
        from scipy.io import netcdf_file
        mode = (mode or "r")[0]
        if mode == "w" and not clobber and os.path.exists(filename):
            raise FileExistsError(filename)
        object.__setattr__(self, "_f", netcdf_file(filename, mode, mmap=False, version=2))
        object.__setattr__(self, "_vars", {})
        object.__setattr__(self, "filepath", lambda: filename)

    def createDimension(self, name, size):
        self._f.createDimension(name, size)

    def createVariable(self, name, datatype, dimensions=(), zlib=False, least_significant_digit=None, **kwargs):
        dims = (dimensions,) if isinstance(dimensions, str) else tuple(dimensions)
        var = self._f.createVariable(name, np.dtype(datatype).char, dims)
        if least_significant_digit is not None:
            var.least_significant_digit = np.int32(least_significant_digit)
        wrapped = _Netcdf3Variable(var, least_significant_digit)
        self._vars[name] = wrapped
        return wrapped

    @property
    def variables(self):
        out = {name: _Netcdf3Variable(v) for name, v in self._f.variables.items()}
        out.update(self._vars)
        return out

    @property
    def dimensions(self):
        return self._f.dimensions

    def ncattrs(self):
        return list(self._f._attributes)

    def getncattr(self, name):
        return self._f._attributes[name]

    def __setattr__(self, name, value):
        setattr(self._f, name, value)  # a global attribute

    def __getattr__(self, name):
        attrs = self.__dict__["_f"]._attributes
        if name in attrs:
            return attrs[name]
        raise AttributeError(name)

    def close(self):
        self._f.close()


def open_dataset(filename, *args, **kwargs):
    """``netCDF4.Dataset(filename, ...)`` where netCDF4 is installed, else a
    :class:`Netcdf3Dataset`."""
    try:
        import netCDF4
    except ImportError:
        return Netcdf3Dataset(filename, *args, **kwargs)
    return netCDF4.Dataset(filename, *args, **kwargs)


def unique_path(path):
    """``path`` itself if no file has that name, else the first free name
    ``<stem>_NNNN<ext>`` (NNNN = 0000, 0001, ...): output files are never
    overwritten."""
    if not os.path.isfile(path):
        return path
    stem, ext = os.path.splitext(path)
    n = 0
    while os.path.isfile("%s_%04d%s" % (stem, n, ext)):
        n += 1
    return "%s_%04d%s" % (stem, n, ext)


class DataDumper:
    """Context manager around one netCDF dataset (the role of the reference's
    DataDumper, Common.py:361-416).  Creates the parent folder when needed; a
    dataset opened for writing gets a name no existing file has (unique_path).

        with DataDumper("out.nc", mode="w") as dump:
            dump.ncfile.createDimension("x", 10)

    ``filename`` is the path actually used; ``ncfile`` the open dataset."""

    def __init__(self, filename, *args, **kwargs):
        self.logger = logging.getLogger(__name__)
        path = os.path.abspath(filename)
        folder = os.path.dirname(path)
        if folder and not os.path.isdir(folder):
            os.makedirs(folder, exist_ok=True)
            self.logger.info("made output folder %s", folder)
        mode = kwargs.get("mode", args[0] if args else "r") or "r"
        self.writing = any(c in mode for c in "wa+")
        self.filename = unique_path(path) if self.writing else path
        self.args, self.kwargs = args, kwargs
        self.ncfile = None

    def __enter__(self):
        self.ncfile = open_dataset(self.filename, *self.args, **self.kwargs)
        self.logger.info("%s opened (%s)", self.filename, "write" if self.writing else "read")
        return self

    def __exit__(self, *exc):
        if self.ncfile is not None:
            self.ncfile.close()
            self.ncfile = None
            self.logger.info("%s closed", self.filename)
        return False


# Arguments holding initial fields: too large for the sim_args attribute.
_FIELD_ARGS = ("rho", "rho_u", "rho_v", "E", "h0", "hu0", "hv0")


class _PhaseClock:
    """Wall-clock sums per phase in the layout runSimulation returns:
    {"start": {phase: sum of start times}, "end": {phase: sum of end times}},
    so end - start is the total time spent in a phase."""

    PHASES = ("t_sim_init", "t_nc_write", "t_full_step")

    def __init__(self):
        self.data = {"start": dict.fromkeys(self.PHASES, 0), "end": dict.fromkeys(self.PHASES, 0)}

    def start(self, phase):
        self.data["start"][phase] += time.time()

    def stop(self, phase):
        self.data["end"][phase] += time.time()


class _SnapshotFile:
    """The netCDF layout of a simulation's output: dimensions (time, y, x),
    float32 coordinate variables time, x, y, one float32 (time, y, x) variable
    per saved field (zlib, least_significant_digit 3), and the global
    attributes created, git_hash, git_status, simulator and sim_args (JSON).
    This layout is the reference's (Common.py:117-163), so that its readers
    open these files."""

    def __init__(self, ncfile, simulator, simulator_args, sim, save_times, names):
        ncfile.created = time.ctime(time.time())
        ncfile.git_hash = getGitHash()
        ncfile.git_status = getGitStatus()
        ncfile.simulator = str(simulator)
        ncfile.sim_args = toJson({k: v for k, v in simulator_args.items() if k not in _FIELD_ARGS})
        nx, ny = int(simulator_args["nx"]), int(simulator_args["ny"])
        for name, size in (("time", len(save_times)), ("x", nx), ("y", ny)):
            ncfile.createDimension(name, size)
        f32 = np.dtype(np.float32).char
        x0, x1, y0, y1 = sim.getExtent()
        for name, values in (("time", save_times), ("x", np.linspace(x0, x1, nx)), ("y", np.linspace(y0, y1, ny))):
            ncfile.createVariable(name, f32, name)[:] = values
        self.fields = [ncfile.createVariable(name, f32, ("time", "y", "x"), zlib=True, least_significant_digit=3)
                       for name in names]

    def write(self, index, values):
        for var, value in zip(self.fields, values):
            var[index, :] = value


def runSimulation(simulator, simulator_args, outfile, save_times, save_var_names=[], dt=None):
    """Constructs ``simulator(**simulator_args)``, advances it to each of
    ``save_times`` in turn and stores the variables named in
    ``save_var_names`` (an entry None skips the variable of that index) at
    each of them, in a new netCDF file named after ``outfile`` (the layout of
    _SnapshotFile).  ``dt``: a fixed step for simulate(), None for its CFL
    policy.  The simulator's check() runs before every interval; a failing
    check ends the run early, logged, with the frames written so far.

    Returns ``(filename, phase_times, profiling_data_mpi)``: the file written,
    the wall-clock sums of _PhaseClock, and the simulator's
    ``profiling_data_mpi`` (None for a simulator without one).  Signature and
    return tuple of the reference's runSimulation (Common.py:92-208)."""
    logger = logging.getLogger(__name__)
    save_times = np.asarray(save_times)
    if save_times.size == 0:  # the reference asserts here
        raise AssertionError("runSimulation: save_times is empty")
    clock = _PhaseClock()
    clock.start("t_sim_init")
    t0 = time.time()
    sim = simulator(**simulator_args)
    logger.info("%s constructed in %.3f s", simulator.__name__ if hasattr(simulator, "__name__") else simulator,
                time.time() - t0)
    keep = [(i, name) for i, name in enumerate(save_var_names) if name is not None]
    # interval k advances from save_times[k-1] (0 for the first) to save_times[k]
    intervals = np.diff(save_times, prepend=0)

    with DataDumper(outfile, mode="w", clobber=False) as dump:
        snaps = _SnapshotFile(dump.ncfile, simulator, simulator_args, sim, save_times, [n for _, n in keep])
        clock.stop("t_sim_init")
        printer = ProgressPrinter(save_times[-1], print_every=10)
        for k, interval in enumerate(intervals):
            try:
                sim.check()
            except AssertionError as e:
                logger.error("stopped before interval %d: check failed at step %d, t=%g (%s)", k, sim.simSteps(),
                             sim.simTime(), e)
                break
            clock.start("t_full_step")
            if interval > 0.0:
                sim.simulate(interval, dt)
            clock.stop("t_full_step")
            clock.start("t_nc_write")
            snaps.write(k, sim.download([i for i, _ in keep]))
            clock.stop("t_nc_write")
            progress = printer.getPrintString(save_times[k])
            if progress:
                logger.debug(progress)
        if sim.simSteps() > 0:
            logger.debug("%d steps to t=%g, mean dt %g", sim.simSteps(), sim.simTime(),
                         sim.simTime() / sim.simSteps())
    return dump.filename, clock.data, getattr(sim, "profiling_data_mpi", None)
