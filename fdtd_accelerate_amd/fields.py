"""Device field storage with the KEMP ``Fields`` surface (``K7/cpu/fields.py:7-145``,
``K7/gpu/fields.py``).

Six float32 tensors ``(nx, ny, nz_pitch)`` on the GPU, z contiguous, with nz padded to
a multiple of ``segment_nbytes`` (64 B by default, as ``K7/gpu/fields.py:45-50``) and
pad lanes held at 0; optional CE/CH coefficient tensors (default 0.5).  Update
objects (Core, Cpml, IncidentDirect, ...) register themselves with
``append_instance`` and run in priority order (core < pml < incident) from
``update_e()`` / ``update_h()``.  All work goes to one HIP stream (torch's current
stream when the Fields is created); ``enqueue_barrier()`` synchronises it, replacing
the reference's QueueTask worker (``K7/cpu/queue_task.py:13-45``).
"""
from __future__ import annotations

import numpy as np
import torch

from . import util


def _device():
    if not torch.cuda.is_available():
        raise RuntimeError("fdtd_accelerate_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


# Arrays at least this large are carved from one allocation with staggered starts.
ARENA_MIN_BYTES = 1 << 30
_ARENA_ROUND = 2 << 20


def staggered_arrays(count, shape, dtype, device, zero=True, offset=0):
    """``count`` arrays of ``shape``.  When one array holds 1 GiB or more they are views
    of one allocation, array k starting k * (size + gap) bytes in, gap = size / 8 rounded
    up to 2 MiB, which fixes their relative placement: twelve separately allocated 4 GiB
    arrays ran the 1024^3 fused step in 11.6-14.0 ms depending on where they landed
    (13.7 ms as the first allocation of a process), carved this way 12.81-12.87 ms in six
    allocations on two GPUs; the in-place step 16.47-16.54 ms against 16.08-16.97 ms.
    Smaller arrays are allocated one by one.  ``offset`` (elements) starts every array
    that far past an alignment of 2 MiB (arena) or 512 B (own allocation), so that its
    first element sits offset * element-size bytes past a 1 KiB boundary."""
    n = 1
    for v in shape:
        n *= v
    esize = torch.empty((), dtype=dtype).element_size()
    nbytes = n * esize
    make = torch.zeros if zero else torch.empty
    if nbytes < ARENA_MIN_BYTES:
        if not offset:
            return [make(shape, dtype=dtype, device=device) for _ in range(count)]
        return [make(n + offset, dtype=dtype, device=device)[offset:].view(shape) for _ in range(count)]
    gap = -(-(nbytes // 8) // _ARENA_ROUND) * _ARENA_ROUND
    step = (nbytes + gap) // esize
    arena = make((count - 1) * step + n + offset, dtype=dtype, device=device)
    return [arena[k * step + offset:k * step + offset + n].view(shape) for k in range(count)]


class Fields:
    def __init__(self, nx, ny, nz, coeff_use="", precision_float="single",
                 segment_nbytes=64, device=None, stream=None):
        util.check_type("nx", nx, int)
        util.check_type("ny", ny, int)
        util.check_type("nz", nz, int)
        util.check_value("coeff_use", coeff_use, ("", "e", "h", "eh"))
        util.check_value("precision_float", precision_float, ("single", "double"))
        util.check_type("segment_nbytes", segment_nbytes, int)
        if segment_nbytes % 16 or segment_nbytes <= 0:
            raise ValueError(f"argument 'segment_nbytes' must be a positive multiple of 16 : {segment_nbytes}")
        if nx < 2 or ny < 2 or nz < 2:
            raise ValueError(f"grid ({nx}, {ny}, {nz}) is too small")

        dtype = torch.float32 if precision_float == "single" else torch.float64
        esize = 4 if precision_float == "single" else 8
        align = segment_nbytes // esize
        pad = int(np.ceil(nz / align) * align) - nz
        nz_pitch = nz + pad
        self.device = torch.device(device) if device is not None else _device()
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)

        self.nx, self.ny, self.nz = nx, ny, nz
        self.ns = [nx, ny, nz]
        self.ns_pitch = [nx, ny, nz_pitch]
        self.ns_pad = [nx, ny, pad]
        self.nz_pitch = nz_pitch
        self.align_size = align
        self.pad = pad
        self.slice_z = slice(None) if pad == 0 else slice(None, -pad)
        self.precision_float = precision_float
        self.dtype = np.float32 if precision_float == "single" else np.float64
        self.torch_dtype = dtype
        self.coeff_use = coeff_use
        self.ce_on = "e" in coeff_use
        self.ch_on = "h" in coeff_use
        self.dx = 1.0
        self.dt = 0.5

        self.ehs = staggered_arrays(6, (nx, ny, nz_pitch), dtype, self.device)
        self.ces = [self._half() for _ in range(3)] if self.ce_on else None
        self.chs = [self._half() for _ in range(3)] if self.ch_on else None
        self.ce_mat = None  # per-component uint8 indices into ce_lut (compress_ce)
        self.ce_lut = None
        self.instance_list = []

    def _half(self):
        return torch.full(tuple(self.ns_pitch), 0.5, dtype=self.torch_dtype, device=self.device)

    # component accessors
    ex = property(lambda s: s.ehs[0])
    ey = property(lambda s: s.ehs[1])
    ez = property(lambda s: s.ehs[2])
    hx = property(lambda s: s.ehs[3])
    hy = property(lambda s: s.ehs[4])
    hz = property(lambda s: s.ehs[5])

    _NAMES = ("ex", "ey", "ez", "hx", "hy", "hz")

    def get(self, str_f):
        names = list(self._NAMES)
        if self.ce_on:
            names += ["cex", "cey", "cez"]
        if self.ch_on:
            names += ["chx", "chy", "chz"]
        util.check_value("str_f", str_f, names)
        if str_f in self._NAMES:
            return self.ehs[self._NAMES.index(str_f)]
        if str_f.startswith("ce"):
            return self.ces["xyz".index(str_f[2])]
        return self.chs["xyz".index(str_f[2])]

    def require_single(self, what):
        """The update objects other than Core run single precision only."""
        if self.precision_float != "single":
            raise ValueError(f"{what} supports precision_float 'single' only : {self.precision_float!r} is given "
                             "(the 'double' programs cover the core update_e/update_h)")

    def append_instance(self, instance):
        util.append_instance(self.instance_list, instance)

    def _fill(self, dst_list, src_list, what):
        for d, s in zip(dst_list, src_list):
            if isinstance(s, np.ndarray):
                s = torch.from_numpy(np.ascontiguousarray(s, dtype=self.dtype))
            elif not isinstance(s, torch.Tensor):
                raise TypeError(f"argument '{what}' type must be a numpy.ndarray or torch.Tensor : {type(s)} is given")
            if tuple(s.shape) != (self.nx, self.ny, self.nz):
                raise ValueError(f"argument '{what}' shape must be {(self.nx, self.ny, self.nz)} : {tuple(s.shape)}")
            with torch.cuda.stream(self.stream):
                d[:, :, self.slice_z].copy_(s.to(self.device, self.torch_dtype), non_blocking=False)

    def set_ehs(self, ex, ey, ez, hx, hy, hz):
        self._fill(self.ehs, [ex, ey, ez, hx, hy, hz], "eh")

    def set_ces(self, cex, cey, cez):
        if not self.ce_on:
            raise AttributeError("The Fields instance has no ce arrays. You should add 'e' in the option 'coeff_use'")
        self._fill(self.ces, [cex, cey, cez], "ce")
        self.ce_mat = self.ce_lut = None

    def compress_ce(self):
        """Let the core E update read the CE arrays as uint8 material indices into a
        3 x 256 table of their exact values (39 instead of 48 bytes per cell; results
        unchanged).  Possible when each component has at most 256 distinct values;
        returns whether it was done.  The float arrays stay (the CPML kernels and host
        access use them); call again after changing CE."""
        if not self.ce_on or self.precision_float != "single":
            return False
        lut = torch.zeros((3, 256), dtype=torch.float32, device=self.device)
        mats = []
        with torch.cuda.stream(self.stream):
            for c, ce in enumerate(self.ces):
                vals, inv = torch.unique(ce, return_inverse=True)
                if vals.numel() > 256:
                    self.ce_mat = self.ce_lut = None
                    return False
                lut[c, :vals.numel()] = vals
                mats.append(inv.to(torch.uint8).contiguous())
                del inv
        self.ce_mat, self.ce_lut = mats, lut
        return True

    def set_chs(self, chx, chy, chz):
        if not self.ch_on:
            raise AttributeError("The Fields instance has no ch arrays. You should add 'h' in the option 'coeff_use'")
        self._fill(self.chs, [chx, chy, chz], "ch")

    def update_e(self):
        for inst in self.instance_list:
            inst.update_e()

    def update_h(self):
        for inst in self.instance_list:
            inst.update_h()

    def enqueue_barrier(self):
        self.stream.synchronize()

    # raw pointers for the C ABI
    def ptrs(self, tensors):
        return [None if t is None else t.data_ptr() for t in tensors]
