"""Serialization savepoints on the Python side: the analogue of the
reference's serialbox calls (advection_CEtimestep_twostep_GPUBC.f90:122-142,
355-375 under SERIALIZE).

``Serializer(dir)`` opens a directory for ``mpdata_serialize_field_*``;
``write(savepoint, name, tensor)`` writes one field (a torch tensor on the
GPU or a numpy array, in the library's memory layout: the Fortran array
(n1, n2, n3) is a C-ordered tensor of shape (n3, n2, n1)); ``read(dir)``
returns {savepoint: {field: numpy array of the Fortran shape}} from the
directory's index, with numpy's loader (no pickled objects)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._lib import LIB, check


class Serializer:
    def __init__(self, directory):
        self.directory = os.fspath(directory)
        check(LIB.mpdata_serializer_open(self.directory.encode()))

    def close(self):
        LIB.mpdata_serializer_close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def write(self, savepoint, name, field, stream=None):
        """One field at one savepoint.  ``field``: contiguous float64/float32
        torch tensor (device or host) or numpy array, C-ordered (n3, n2, n1)
        or (n2, n1) / (n1,), written as the Fortran array (n1, n2, n3)."""
        shape = tuple(int(s) for s in field.shape)
        if len(shape) > 3 or len(shape) == 0:
            raise ValueError(f"{name}: 1 to 3 dimensions, got {shape}")
        n3, n2, n1 = (1,) * (3 - len(shape)) + shape
        if hasattr(field, "data_ptr"):  # torch
            import torch
            if not field.is_contiguous():
                raise ValueError(f"{name}: not contiguous")
            dt = {torch.float64: "_f64", torch.float32: "_f32"}.get(field.dtype)
            ptr = field.data_ptr()
            if stream is None and field.is_cuda:
                stream = torch.cuda.current_stream(field.device).cuda_stream
        else:
            field = np.ascontiguousarray(field)
            dt = {np.dtype(np.float64): "_f64", np.dtype(np.float32): "_f32"}.get(field.dtype)
            ptr = field.ctypes.data
        if dt is None:
            raise TypeError(f"{name}: float64 or float32 expected")
        check(getattr(LIB, "mpdata_serialize_field" + dt)(
            savepoint.encode(), name.encode(), ctypes.c_void_p(ptr), n1, n2, n3,
            ctypes.c_void_p(stream or 0)))


def read(directory):
    """{savepoint: {field: array}} of a serializer directory; a field written
    twice at one savepoint reads as its last version.  Arrays have the
    Fortran shape (n1, n2, n3) in Fortran order."""
    out = {}
    with open(os.path.join(directory, "savepoints.txt")) as f:
        for line in f:
            parts = line.split()
            if len(parts) != 7:
                continue
            sp, name, fname = parts[:3]
            arr = np.load(os.path.join(directory, fname), allow_pickle=False)
            if arr.shape != tuple(int(v) for v in parts[4:7]):
                raise ValueError(f"{fname}: shape {arr.shape} differs from the index")
            out.setdefault(sp, {})[name] = arr
    return out
