"""KEMP ``IncidentDirect`` (``K7/cpu/incident_direct.py:9-85``) and ``GetFields`` /
``SetFields`` (``K7/cpu/get_set_fields.py:7-127``) on device fields.

IncidentDirect adds (or writes) ``float32(tfunc(tstep)) * spatial_value`` into the box
[pt0, pt1] of one component once per E (or H) half step, after the core and CPML
updates (priority 'incident'); ``tstep`` starts at 1 and advances after each use.
"""
from __future__ import annotations

import numbers

import numpy as np
import torch

from . import _lib, util
from .fields import Fields

_FIELD_NAMES = ("ex", "ey", "ez", "hx", "hy", "hz")


def _points(fields, pt0, pt1):
    util.check_type("pt0", pt0, (list, tuple), (int, float, np.integer, np.floating))
    util.check_type("pt1", pt1, (list, tuple), (int, float, np.integer, np.floating))
    p0 = list(util.convert_indices(fields.ns, pt0))
    p1 = list(util.convert_indices(fields.ns, pt1))
    for axis, n, a, b in zip("xyz", fields.ns, p0, p1):
        util.check_value(f"pt0 {axis}", a, range(n))
        util.check_value(f"pt1 {axis}", b, range(n))
    return p0, p1


class IncidentDirect:
    priority_type = "incident"

    def __init__(self, fields, str_f, pt0, pt1, tfunc, spatial_value=1.0, is_overwrite=False):
        util.check_type("fields", fields, Fields)
        fields.require_single("IncidentDirect")
        util.check_value("str_f", str_f, _FIELD_NAMES)
        if not callable(tfunc):
            raise TypeError(f"argument 'tfunc' type must be a function : {type(tfunc)} is given")
        util.check_type("spatial_value", spatial_value, (np.ndarray, torch.Tensor, numbers.Real))
        util.check_type("is_overwrite", is_overwrite, bool)
        p0, p1 = _points(fields, pt0, pt1)
        if any(a > b for a, b in zip(p0, p1)):
            raise ValueError(f"pt0 {p0} must not exceed pt1 {p1}")
        self.mainf = fields
        self.str_f = str_f
        self.pt0, self.pt1 = p0, p1
        self.tfunc = tfunc
        self.is_overwrite = is_overwrite
        self.e_or_h = str_f[0]
        self.tstep = 1
        if isinstance(spatial_value, (np.ndarray, torch.Tensor)):
            shape = util.shape_two_points(p0, p1)
            if tuple(spatial_value.shape) != shape:
                raise AssertionError(f"shape mismatch : {shape}, {tuple(spatial_value.shape)}")
            sv = torch.as_tensor(spatial_value)
            if sv.dtype != torch.float32:
                raise AssertionError(f"dtype mismatch : float32, {sv.dtype}")
            self.sv = sv.to(fields.device).contiguous()
            self.svs = 0.0
        else:
            self.sv = None
            self.svs = float(np.float32(spatial_value))
        _lib.load()
        fields.append_instance(self)

    def tvalue(self, tstep):
        """The host value of the time function, cast to float32 as the reference does."""
        return float(np.float32(self.tfunc(tstep)))

    def update(self):
        f = self.mainf
        p0, p1 = self.pt0, self.pt1
        _lib.call("mifdtd_source", f.get(self.str_f).data_ptr(), f.nx, f.ny, f.nz_pitch,
                  p0[0], p1[0], p0[1], p1[1], p0[2], p1[2], self.tvalue(self.tstep),
                  None if self.sv is None else self.sv.data_ptr(), self.svs,
                  int(self.is_overwrite), _lib.stream_handle(f.stream))
        self.tstep += 1

    def update_e(self):
        if self.e_or_h == "e":
            self.update()

    def update_h(self):
        if self.e_or_h == "h":
            self.update()


class GetFields:
    """Copy the box [pt0, pt1] of one or more components to a host array."""

    def __init__(self, fields, str_f, pt0, pt1):
        util.check_type("fields", fields, Fields)
        fields.require_single("GetFields")
        util.check_type("str_f", str_f, (str, list, tuple), str)
        str_fs = (str_f,) if isinstance(str_f, str) else tuple(str_f)
        for s in str_fs:
            util.check_value("str_f", s, _FIELD_NAMES)
        p0, p1 = _points(fields, pt0, pt1)
        self.mainf = fields
        self.str_fs = str_fs
        self.slice_xyz = util.slices_two_points(p0, p1)
        shape = util.shape_two_points(p0, p1, len(str_fs))
        self.host_array = np.zeros(shape, dtype=np.float32)
        self.split_host_array_dict = dict(zip(str_fs, np.split(self.host_array, len(str_fs))))

    def copy_to_host_array(self):
        for s in self.str_fs:
            with torch.cuda.stream(self.mainf.stream):
                v = self.mainf.get(s)[self.slice_xyz].cpu().numpy()
            self.split_host_array_dict[s][:] = v.reshape(self.split_host_array_dict[s].shape)

    def get_event(self):
        self.copy_to_host_array()
        return _Done()

    def get_fields(self, str_f=""):
        return self.host_array if str_f == "" else self.split_host_array_dict[str_f]


class SetFields:
    """Write (or add) a scalar or an array into the box [pt0, pt1] of components."""

    def __init__(self, fields, str_f, pt0, pt1, is_array=False, is_overwrite=True):
        util.check_type("fields", fields, Fields)
        fields.require_single("SetFields")
        util.check_type("str_f", str_f, (str, list, tuple), str)
        util.check_type("is_array", is_array, bool)
        util.check_type("is_overwrite", is_overwrite, bool)
        str_fs = (str_f,) if isinstance(str_f, str) else tuple(str_f)
        for s in str_fs:
            util.check_value("str_f", s, _FIELD_NAMES)
        p0, p1 = _points(fields, pt0, pt1)
        self.mainf = fields
        self.str_fs = str_fs
        self.slice_xyz = util.slices_two_points(p0, p1)
        self.shape = util.shape_two_points(p0, p1, len(str_fs))
        self.is_array = is_array
        self.is_overwrite = is_overwrite

    def set_fields(self, value, wait_list=()):
        f = self.mainf
        with torch.cuda.stream(f.stream):
            if self.is_array:
                util.check_value("value.dtype", np.asarray(value).dtype, np.float32)
                util.check_value("value.shape", tuple(np.shape(value)), [self.shape])
                parts = np.split(np.asarray(value), len(self.str_fs))
                for s, part in zip(self.str_fs, parts):
                    dst = f.get(s)[self.slice_xyz]
                    t = torch.from_numpy(np.ascontiguousarray(part)).to(f.device).reshape(dst.shape)
                    if self.is_overwrite:
                        dst.copy_(t)
                    else:
                        dst.add_(t)
            else:
                v = float(np.float32(value))
                for s in self.str_fs:
                    dst = f.get(s)[self.slice_xyz]
                    if self.is_overwrite:
                        dst.fill_(v)
                    else:
                        dst.add_(v)
        return _Done()


class _Done:
    def wait(self):
        return None
