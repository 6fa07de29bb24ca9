"""Argument checks and index helpers with the behaviour of the reference's
``K7/util/common.py`` (same names, same exceptions), so code written against the
KEMP facade gets the same errors from this package."""
from __future__ import annotations

import numpy as np

_INTS = (int, np.integer)


def check_type(arg_name, arg, arg_type, element_type=None):
    """Raise TypeError unless ``arg`` is an ``arg_type``; for lists/tuples each element
    must be an ``element_type`` (``K7/util/common.py:6-36``)."""
    if not isinstance(arg, arg_type):
        raise TypeError(f"argument '{arg_name}' type must be a {arg_type!r} : {type(arg)} is given")
    if isinstance(arg, (list, tuple)):
        if element_type is None:
            raise TypeError("When the 'arg' is a list or a tuple, argument 'element_type' must be specified.")
        for e in arg:
            if not isinstance(e, element_type):
                raise TypeError(f"argument '{arg_name}'s element type must be a {element_type!r} : {type(e)} is given")


def check_value(arg_name, arg, value):
    """Raise ValueError unless ``arg`` is one of ``value`` (``K7/util/common.py:40-52``)."""
    vals = value if isinstance(value, (list, tuple, range, set, frozenset)) else (value,)
    if arg not in vals:
        r = repr(value)
        if isinstance(value, (list, tuple)) and len(r) > 40:
            r = str(list(value[:2]) + ["..."] + list(value[-2:])).replace("'", "")
        raise ValueError(f"argument '{arg_name}' value must be one of {r} : {arg!r} is given")


def convert_index(n, num):
    """Negative ints count from the end, floats are fractions of ``n``
    (``K7/util/common.py:219-241``)."""
    check_type("n", n, _INTS)
    check_type("num", num, (int, float, np.integer, np.floating))
    if isinstance(num, (float, np.floating)):
        return int(n * num)
    num = int(num)
    return num if num >= 0 else n + num


def convert_indices(ns, pt):
    check_type("pt", pt, (list, tuple), (int, float, np.integer, np.floating))
    return tuple(convert_index(n, p) for n, p in zip(ns, pt))


def slices_two_points(pt0, pt1):
    """Slices (or ints for equal coordinates) spanning two inclusive points
    (``K7/util/common.py:114-131``)."""
    return tuple(slice(p0, p1 + 1) if p0 != p1 else p0 for p0, p1 in zip(pt0, pt1))


def shape_two_points(pt0, pt1, mul_x=1, is_dummy=False):
    """Shape of the box between two inclusive points with unit axes dropped
    (unless ``is_dummy``), first axis multiplied by ``mul_x`` (``K7/util/common.py:155-197``)."""
    shape = []
    for p0, p1 in zip(pt0, pt1):
        v = abs(p1 - p0) + 1
        if v != 1 or is_dummy:
            shape.append(v)
    if not shape:
        return (mul_x,)
    shape[0] *= mul_x
    return tuple(shape)


# Update order within a half step (``K7/util/common.py:383-397``).
PRIORITY = {"core": 0, "current": 1, "pml": 2, "incident": 3, "pbc": 4, "mpi": 5, "exchange": 6}


def append_instance(instance_list, instance):
    """Insert ``instance`` after every instance of equal or lower priority."""
    new = PRIORITY[instance.priority_type]
    index = len(instance_list)
    for i, inst in enumerate(instance_list):
        if new < PRIORITY[inst.priority_type]:
            index = i
            break
    instance_list.insert(index, instance)
