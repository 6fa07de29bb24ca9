"""Field snapshots to HDF5 files and the geometry-h5 loader.

``SaveFields`` is KEMP's ``K7/node/save_fields.py:10-48``: every call of
``save_fields(tstep)`` writes one file ``<dir_path><tstep, zero-padded>_tag<tag>.h5`` that
holds one gzip-compressed dataset per component of the box [pt0, pt1] (shape of
``K7/util/common.py:155-179``, unit axes dropped). The reference copies the box to the
host and writes it before the call returns. Here the call only enqueues work:
1. a device copy of the box into a staging slot, on the fields' stream, ordered after
   the step that produced the values;
2. a device-to-host copy of the slot into pinned host memory, on a copy stream of its
   own, so the next steps run while it moves;
3. the HDF5 write on a writer thread (``libmifdtd_h5.so``), which deflates the chunks
   on several host threads.
``nslots`` staging slots are in flight at most; a call blocks only while the oldest slot
is still being written. ``wait()`` returns when every file is complete, and raises
the first write error.

``read_geometry_h5`` reads the attributes ``K7/main.py:70-75`` reads (``coeff_use``,
``nx``, ``ny``, ``nz``); ``write_geometry_h5`` writes such a file as
``K7/test/create_sample_h5.py:3-8`` does. ``read_dataset`` / ``dataset_info`` read a
snapshot back.
"""
from __future__ import annotations

import concurrent.futures as cf
import ctypes
import os
import pickle
import threading

import numpy as np

from . import util
from .build import H5_LIB

_I, _LL, _P, _S = ctypes.c_int, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_char_p
_SIGS = {
    "mifdtd_h5_last_error": (_S, []),
    "mifdtd_h5_write_fields": (_I, [_S, _I, ctypes.POINTER(_S), ctypes.POINTER(_P), _I, _I,
                                    ctypes.POINTER(_LL), _I, _I]),
    "mifdtd_h5_write_geometry": (_I, [_S, _LL, _LL, _LL, _S]),
    "mifdtd_h5_read_geometry": (_I, [_S, ctypes.POINTER(_LL), ctypes.POINTER(_LL), ctypes.POINTER(_LL),
                                     ctypes.c_char_p, _I]),
    "mifdtd_h5_dataset_info": (_I, [_S, _S, ctypes.POINTER(_I), ctypes.POINTER(_LL), ctypes.POINTER(_I),
                                    ctypes.POINTER(_I)]),
    "mifdtd_h5_read_dataset": (_I, [_S, _S, _P, _LL, _I]),
}
EXPORTS = tuple(_SIGS)
MAX_RANK = 8
_FIELD_NAMES = ("ex", "ey", "ez", "hx", "hy", "hz")

_lock = threading.Lock()
_lib = None


def load():
    """Load (once) and return ``libmifdtd_h5.so``; raise if it is not built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(H5_LIB):
                raise RuntimeError(f"{H5_LIB} is not built (it needs the HDF5 C library): run "
                                   "`python -m fdtd_accelerate_amd.build`")
            lib = ctypes.CDLL(H5_LIB)
            for name, (res, args) in _SIGS.items():
                fn = getattr(lib, name)
                fn.restype, fn.argtypes = res, args
            _lib = lib
    return _lib


def _check(rc, what):
    if rc != 0:
        msg = load().mifdtd_h5_last_error()
        raise RuntimeError(f"{what} failed (rc={rc}): {msg.decode() if msg else ''}")


def _b(s):
    return os.fsencode(s) if not isinstance(s, bytes) else s


# ------------------------------------------------------------------ files
def write_fields(path, arrays, gzip_level=4, nthreads=8):
    """Write {name: array} (one shape, float32 or float64, rank 1-3) into a new file."""
    names = list(arrays)
    if not names:
        raise ValueError("no datasets to write")
    arrs = [np.ascontiguousarray(arrays[n]) for n in names]
    a0 = arrs[0]
    if a0.dtype not in (np.float32, np.float64):
        raise TypeError(f"dataset dtype must be float32 or float64 : {a0.dtype}")
    for a in arrs:
        if a.shape != a0.shape or a.dtype != a0.dtype:
            raise ValueError("every dataset of a snapshot has the same shape and dtype")
    dims = (_LL * a0.ndim)(*a0.shape)
    cnames = (_S * len(names))(*[n.encode() for n in names])
    ptrs = (_P * len(arrs))(*[a.ctypes.data for a in arrs])
    _check(load().mifdtd_h5_write_fields(_b(path), len(names), cnames, ptrs, int(a0.dtype == np.float64),
                                         a0.ndim, dims, int(gzip_level), int(nthreads)), f"writing {path}")


def dataset_info(path, name):
    """(shape, dtype, deflate level or -1) of dataset ``name``."""
    rank, f64, lvl = _I(), _I(), _I()
    dims = (_LL * MAX_RANK)()
    _check(load().mifdtd_h5_dataset_info(_b(path), name.encode(), ctypes.byref(rank), dims, ctypes.byref(f64),
                                         ctypes.byref(lvl)), f"reading {path}:{name}")
    return tuple(dims[:rank.value]), (np.float64 if f64.value else np.float32), lvl.value


def read_dataset(path, name, dtype=None):
    """Dataset ``name`` as a numpy array (its own float type unless ``dtype`` is given)."""
    shape, dt, _ = dataset_info(path, name)
    dt = np.dtype(dtype or dt)
    if dt not in (np.float32, np.float64):
        raise TypeError(f"dtype must be float32 or float64 : {dt}")
    out = np.empty(shape, dt)
    _check(load().mifdtd_h5_read_dataset(_b(path), name.encode(), out.ctypes.data, out.nbytes,
                                         int(dt == np.float64)), f"reading {path}:{name}")
    return out


def write_geometry_h5(path, nx, ny, nz, coeff_use=""):
    """The geometry file of K7/test/create_sample_h5.py:3-8."""
    util.check_value("coeff_use", coeff_use, ("", "e", "h", "eh"))
    _check(load().mifdtd_h5_write_geometry(_b(path), int(nx), int(ny), int(nz), coeff_use.encode()),
           f"writing {path}")


def read_geometry_h5(path):
    """(coeff_use, nx, ny, nz) of a geometry file, as K7/main.py:70-75 reads them."""
    nx, ny, nz = _LL(), _LL(), _LL()
    buf = ctypes.create_string_buffer(64)
    _check(load().mifdtd_h5_read_geometry(_b(path), ctypes.byref(nx), ctypes.byref(ny), ctypes.byref(nz), buf, 64),
           f"reading {path}")
    coeff_use = buf.value.decode("ascii")
    util.check_value("coeff_use", coeff_use, ("", "e", "h", "eh"))
    return coeff_use, int(nx.value), int(ny.value), int(nz.value)


def fpath_form(dir_path, tag, max_tstep):
    """The file-name pattern of K7/node/save_fields.py:30-31: the step number padded to
    one digit more than the decimal exponent of max_tstep."""
    ndigit = int(("%e" % max_tstep).split("+")[1]) + 1
    return dir_path + "%%.%dd_tag%d.h5" % (ndigit, tag)


# ------------------------------------------------------------------ SaveFields
class SaveFields:
    """Asynchronous snapshots of components of the box [pt0, pt1] (K7/node/save_fields.py:10-48).

    ``fields`` is a ``Fields`` (the engines keep it pointing at the current state).
    ``gzip_level``, ``nslots`` and ``nthreads`` have no reference counterpart: deflate
    level (h5py's gzip default 4), staging slots in flight, compression threads.
    """

    def __init__(self, fields, str_f, pt0, pt1, dir_path, tag, max_tstep, gzip_level=4, nslots=2, nthreads=8):
        import torch
        from .fields import Fields
        from .incident import _points
        util.check_type("fields", fields, Fields)
        util.check_type("str_f", str_f, (str, list, tuple), str)
        util.check_type("dir_path", dir_path, str)
        util.check_type("tag", tag, int)
        util.check_type("max_tstep", max_tstep, int)
        util.check_type("nslots", nslots, int)
        str_fs = (str_f,) if isinstance(str_f, str) else tuple(str_f)
        for s in str_fs:
            util.check_value("str_f", s, _FIELD_NAMES)
        if nslots < 1:
            raise ValueError(f"argument 'nslots' must be at least 1 : {nslots}")
        load()
        p0, p1 = _points(fields, pt0, pt1)
        self.mainf = fields
        self.str_fs = str_fs
        self.pt0, self.pt1 = tuple(p0), tuple(p1)
        self.slices = util.slices_two_points(p0, p1)
        self.shape = util.shape_two_points(p0, p1)
        self.fpath_form = fpath_form(dir_path, tag, max_tstep)
        self.gzip_level, self.nthreads = int(gzip_level), int(nthreads)
        tdt = torch.float32 if fields.precision_float == "single" else torch.float64
        n = len(str_fs)
        self._stage = [torch.empty((n,) + self.shape, dtype=tdt, device=fields.device) for _ in range(nslots)]
        self._host = [torch.empty((n,) + self.shape, dtype=tdt, pin_memory=True) for _ in range(nslots)]
        self._copy_stream = torch.cuda.Stream(fields.device)
        self._pending = [None] * nslots
        self._slot = 0
        self._writer = cf.ThreadPoolExecutor(max_workers=1, thread_name_prefix="savefields")
        self.saved = []

    def save_fields(self, tstep):
        """Snapshot the box as the fields stand after the work enqueued so far, into the
        file of step ``tstep``; returns at once (see ``wait``)."""
        import torch
        s = self._slot
        self._slot = (s + 1) % len(self._stage)
        if self._pending[s] is not None:   # the slot's previous file is written, its copies done
            self._pending[s].result()
        f = self.mainf
        stage, host = self._stage[s], self._host[s]
        with torch.cuda.stream(f.stream):
            for c, name in enumerate(self.str_fs):
                src = f.get(name)[self.slices]
                stage[c].copy_(src.reshape(1) if src.dim() == 0 else src)
            staged = torch.cuda.Event()
            staged.record(f.stream)
        self._copy_stream.wait_event(staged)
        with torch.cuda.stream(self._copy_stream):
            host.copy_(stage, non_blocking=True)
            copied = torch.cuda.Event()
            copied.record(self._copy_stream)
        path = self.fpath_form % tstep
        arrays = {name: host[c].numpy() for c, name in enumerate(self.str_fs)}

        def write(stage=stage, host=host):  # both stay referenced until their copy has completed
            copied.synchronize()
            write_fields(path, arrays, self.gzip_level, self.nthreads)
            return path

        self._pending[s] = self._writer.submit(write)
        self.saved.append(path)

    def wait(self):
        """Block until every snapshot file is complete; raise the first write error."""
        for k, p in enumerate(self._pending):
            if p is not None:
                self._pending[k] = None
                p.result()

    def close(self):
        self.wait()
        self._writer.shutdown(wait=True)


def write_divide_info(dir_path, shape, pt0, pt1, nx, str_fs, tmax, tgap):
    """``divide_info.pkl`` of a single-device run (K7/main.py:275-289): the saved box,
    the device x ranges (accumulated nx-1, K7/node/fields.py:29-31), components and the
    time range; pickle protocol 2, as the Python 2 reference reads it."""
    info = {"shape": tuple(shape), "pt0": tuple(pt0), "pt1": tuple(pt1), "anx_list": [0, nx - 1],
            "str_fs": tuple(str_fs), "tmax": tmax, "tgap": tgap}
    with open(dir_path + "divide_info.pkl", "wb") as fh:
        pickle.dump(info, fh, protocol=2)
    return info
