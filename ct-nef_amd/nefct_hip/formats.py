"""On-disk formats of the reference's object model: HDF5 in deepdish's
layout and JSON.

Replaces nefct/io/local_io_mixin.py:14-61 (``save``/``load`` of an object's
``asdict(recurse=True)`` through ``deepdish.io.save/load``, no compression),
nefct/io/data_io_mixin.py:15-51 (``dump_data``/``load_data``: the ``data``
field alone to its own file, the field then holding the path) and
nefct/io/json_io_mixin.py:17-72 (``json_dumps``/``json_dump``/``json_load``/
``json_loads``, refusing objects whose ``data`` is not yet dumped).

deepdish and PyTables are not needed: the HDF5 files are written and read
through the HDF5 C library (``libhdf5``, found through ``$NEFCT_HDF5_LIB``,
the loader's search path or the conda prefix) with ctypes, in the layout
deepdish gives a nested dict:

* a dict is a group (PyTables title ``dict:<n>``), each string key a member;
* a numpy array (or tensor) is a dataset of its dtype and shape;
* a list / tuple of fewer than 256 entries is a group titled ``list`` /
  ``tuple`` whose members are named ``i0``, ``i1``, ...;
* None is an empty group titled ``nonetype``;
* a scalar (int, float, bool, str) is an attribute of the enclosing group;
* a value that is not a dict is saved under the name ``data`` with the root
  attribute ``DEEPDISH_IO_UNPACK`` set, and loaded back unwrapped.

Nodes carry the PyTables system attributes (CLASS, TITLE, VERSION) that
PyTables reads; the reader skips them, so files written by deepdish load too.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import dataclasses
import hashlib
import json
import os
import time
from typing import Optional

import numpy as np

# --------------------------------------------------------------------------
# libhdf5 through ctypes
# --------------------------------------------------------------------------
_hid = ctypes.c_int64
_herr = ctypes.c_int
_hsize = ctypes.c_ulonglong
H5P_DEFAULT = 0
H5S_ALL = 0
H5S_SCALAR = 0
H5F_ACC_RDONLY = 0
H5F_ACC_TRUNC = 2
H5_INDEX_NAME = 0
H5_ITER_INC = 0
H5I_GROUP = 2
H5I_DATASET = 5
H5T_INTEGER, H5T_FLOAT, H5T_STRING, H5T_BITFIELD, H5T_ENUM = 0, 1, 3, 4, 8
H5T_CSET_UTF8 = 1
H5T_STR_NULLTERM = 0
H5T_SGN_NONE = 0

_LINK_CB = ctypes.CFUNCTYPE(_herr, _hid, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p)
_ATTR_CB = ctypes.CFUNCTYPE(_herr, _hid, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p)

_h5 = None
_T = {}


class FormatError(RuntimeError):
    pass


def _find_hdf5() -> str:
    env = os.environ.get("NEFCT_HDF5_LIB")
    if env:
        return env
    found = ctypes.util.find_library("hdf5")
    if found:
        return found
    for prefix in (os.environ.get("CONDA_PREFIX"), "/opt/conda", "/usr"):
        if prefix:
            for sub in ("lib", "lib64", "lib/x86_64-linux-gnu"):
                p = os.path.join(prefix, sub, "libhdf5.so")
                if os.path.exists(p):
                    return p
    raise FormatError("libhdf5 not found (set NEFCT_HDF5_LIB to the HDF5 C library)")


def h5lib():
    """The HDF5 C library with the signatures used here (loaded once)."""
    global _h5
    if _h5 is not None:
        return _h5
    L = ctypes.CDLL(_find_hdf5())
    sig = {
        "H5open": ([], _herr), "H5Fcreate": ([ctypes.c_char_p, ctypes.c_uint, _hid, _hid], _hid),
        "H5Fopen": ([ctypes.c_char_p, ctypes.c_uint, _hid], _hid), "H5Fclose": ([_hid], _herr),
        "H5Gcreate2": ([_hid, ctypes.c_char_p, _hid, _hid, _hid], _hid),
        "H5Gopen2": ([_hid, ctypes.c_char_p, _hid], _hid), "H5Gclose": ([_hid], _herr),
        "H5Oopen": ([_hid, ctypes.c_char_p, _hid], _hid), "H5Oclose": ([_hid], _herr),
        "H5Iget_type": ([_hid], ctypes.c_int),
        "H5Screate": ([ctypes.c_int], _hid),
        "H5Screate_simple": ([ctypes.c_int, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize)], _hid),
        "H5Sclose": ([_hid], _herr), "H5Sget_simple_extent_ndims": ([_hid], ctypes.c_int),
        "H5Sget_simple_extent_dims": ([_hid, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize)], ctypes.c_int),
        "H5Dcreate2": ([_hid, ctypes.c_char_p, _hid, _hid, _hid, _hid, _hid], _hid),
        "H5Dopen2": ([_hid, ctypes.c_char_p, _hid], _hid), "H5Dclose": ([_hid], _herr),
        "H5Dwrite": ([_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p], _herr),
        "H5Dread": ([_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p], _herr),
        "H5Dget_type": ([_hid], _hid), "H5Dget_space": ([_hid], _hid),
        "H5Acreate2": ([_hid, ctypes.c_char_p, _hid, _hid, _hid, _hid], _hid),
        "H5Aopen": ([_hid, ctypes.c_char_p, _hid], _hid), "H5Aclose": ([_hid], _herr),
        "H5Awrite": ([_hid, _hid, ctypes.c_void_p], _herr), "H5Aread": ([_hid, _hid, ctypes.c_void_p], _herr),
        "H5Aget_type": ([_hid], _hid), "H5Aget_space": ([_hid], _hid),
        "H5Aexists": ([_hid, ctypes.c_char_p], ctypes.c_int),
        "H5Aiterate2": ([_hid, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_hsize), _ATTR_CB, ctypes.c_void_p],
                        _herr),
        "H5Literate": ([_hid, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_hsize), _LINK_CB, ctypes.c_void_p],
                       _herr),
        "H5Tcopy": ([_hid], _hid), "H5Tclose": ([_hid], _herr), "H5Tset_size": ([_hid, ctypes.c_size_t], _herr),
        "H5Tset_cset": ([_hid, ctypes.c_int], _herr), "H5Tset_strpad": ([_hid, ctypes.c_int], _herr),
        "H5Tget_class": ([_hid], ctypes.c_int), "H5Tget_size": ([_hid], ctypes.c_size_t),
        "H5Tget_sign": ([_hid], ctypes.c_int), "H5Tis_variable_str": ([_hid], ctypes.c_int),
        "H5Tget_super": ([_hid], _hid),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    if L.H5open() < 0:
        raise FormatError("H5open failed")
    for key, sym in (("f32", "H5T_NATIVE_FLOAT_g"), ("f64", "H5T_NATIVE_DOUBLE_g"),
                     ("i8", "H5T_NATIVE_INT8_g"), ("i16", "H5T_NATIVE_INT16_g"), ("i32", "H5T_NATIVE_INT32_g"),
                     ("i64", "H5T_NATIVE_INT64_g"), ("u8", "H5T_NATIVE_UINT8_g"), ("u16", "H5T_NATIVE_UINT16_g"),
                     ("u32", "H5T_NATIVE_UINT32_g"), ("u64", "H5T_NATIVE_UINT64_g"), ("b8", "H5T_NATIVE_B8_g"),
                     ("s1", "H5T_C_S1_g")):
        _T[key] = _hid.in_dll(L, sym).value
    _h5 = L
    return L


def _ok(v, what):
    if v < 0:
        raise FormatError(f"HDF5 call failed: {what}")
    return v


_NP2H5 = {np.dtype(np.float32): "f32", np.dtype(np.float64): "f64", np.dtype(np.int8): "i8",
          np.dtype(np.int16): "i16", np.dtype(np.int32): "i32", np.dtype(np.int64): "i64",
          np.dtype(np.uint8): "u8", np.dtype(np.uint16): "u16", np.dtype(np.uint32): "u32",
          np.dtype(np.uint64): "u64", np.dtype(np.bool_): "b8"}


def _str_type(nbytes: int) -> int:
    L = h5lib()
    t = _ok(L.H5Tcopy(_T["s1"]), "H5Tcopy")
    _ok(L.H5Tset_size(t, max(1, nbytes)), "H5Tset_size")
    _ok(L.H5Tset_cset(t, H5T_CSET_UTF8), "H5Tset_cset")
    _ok(L.H5Tset_strpad(t, H5T_STR_NULLTERM), "H5Tset_strpad")
    return t


def _space(shape) -> int:
    L = h5lib()
    if len(shape) == 0:
        return _ok(L.H5Screate(H5S_SCALAR), "H5Screate")
    dims = (_hsize * len(shape))(*shape)
    return _ok(L.H5Screate_simple(len(shape), dims, None), "H5Screate_simple")


def _write_attr(obj: int, name: str, value):
    """Scalar or string attribute (PyTables stores plain values this way)."""
    L = h5lib()
    if isinstance(value, (str, np.str_)):
        raw = str(value).encode("utf-8") + b"\0"
        t = _str_type(len(raw))
        sp = _space(())
        a = _ok(L.H5Acreate2(obj, name.encode(), t, sp, H5P_DEFAULT, H5P_DEFAULT), f"H5Acreate2 {name}")
        buf = ctypes.create_string_buffer(raw, len(raw))
        _ok(L.H5Awrite(a, t, buf), "H5Awrite")
        L.H5Aclose(a)
        L.H5Sclose(sp)
        L.H5Tclose(t)
        return
    arr = np.array(value)          # keeps a scalar 0-d (ascontiguousarray would make it 1-d)
    if arr.dtype == np.bool_:
        arr = arr.astype(np.bool_)
    elif arr.dtype.kind == "i":
        arr = arr.astype(np.int64)
    elif arr.dtype.kind == "u":
        arr = arr.astype(np.uint64)
    elif arr.dtype.kind == "f":
        arr = arr.astype(np.float64) if arr.dtype != np.float32 else arr
    else:
        raise FormatError(f"cannot store attribute {name!r} of type {type(value).__name__}")
    t = _T[_NP2H5[arr.dtype]]
    sp = _space(arr.shape)
    a = _ok(L.H5Acreate2(obj, name.encode(), t, sp, H5P_DEFAULT, H5P_DEFAULT), f"H5Acreate2 {name}")
    _ok(L.H5Awrite(a, t, arr.ctypes.data_as(ctypes.c_void_p)), "H5Awrite")
    L.H5Aclose(a)
    L.H5Sclose(sp)


def _mark(obj: int, cls: str, title: str, version: str):
    _write_attr(obj, "CLASS", cls)
    _write_attr(obj, "TITLE", title)
    _write_attr(obj, "VERSION", version)


def _group(parent: int, name: str, title: str) -> int:
    g = _ok(h5lib().H5Gcreate2(parent, name.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), f"H5Gcreate2 {name}")
    _mark(g, "GROUP", title, "1.0")
    return g


def _write_array(parent: int, name: str, arr: np.ndarray):
    L = h5lib()
    arr = np.array(arr, order="C", copy=True)
    if arr.dtype not in _NP2H5:
        raise FormatError(f"cannot store array {name!r} of dtype {arr.dtype}")
    t = _T[_NP2H5[arr.dtype]]
    sp = _space(arr.shape)
    d = _ok(L.H5Dcreate2(parent, name.encode(), t, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), f"H5Dcreate2 {name}")
    if arr.size:
        _ok(L.H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, arr.ctypes.data_as(ctypes.c_void_p)), "H5Dwrite")
    _mark(d, "ARRAY", "", "2.4")
    _write_attr(d, "FLAVOR", "numpy")
    L.H5Dclose(d)
    L.H5Sclose(sp)


def _as_array(value):
    if hasattr(value, "detach") and hasattr(value, "cpu"):        # torch tensor
        return value.detach().cpu().numpy()
    return value


def _save_level(parent: int, name: str, value):
    L = h5lib()
    value = _as_array(value)
    if isinstance(value, dict):
        g = _group(parent, name, f"dict:{len(value)}")
        for k, v in value.items():
            if not isinstance(k, str):
                raise FormatError("only string keys are stored")
            _save_level(g, k, v)
        L.H5Gclose(g)
    elif isinstance(value, (list, tuple)) and len(value) < 256:
        g = _group(parent, name, "list" if isinstance(value, list) else "tuple")
        for i, v in enumerate(value):
            _save_level(g, f"i{i}", v)
        L.H5Gclose(g)
    elif isinstance(value, np.ndarray):
        _write_array(parent, name, value)
    elif value is None:
        L.H5Gclose(_group(parent, name, "nonetype"))
    elif isinstance(value, (list, tuple)):
        _write_array(parent, name, np.asarray(value))
    else:
        _write_attr(parent, name, value)


def h5_save(path: str, data) -> str:
    """Write ``data`` (a nested dict, or any value) to ``path`` in deepdish's
    layout, uncompressed (deepdish.io.save(path, data, compression=None))."""
    L = h5lib()
    f = _ok(L.H5Fcreate(os.fsencode(path), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), f"H5Fcreate {path}")
    try:
        root = _ok(L.H5Gopen2(f, b"/", H5P_DEFAULT), "H5Gopen2 /")
        _mark(root, "GROUP", "", "1.0")
        _write_attr(root, "PYTABLES_FORMAT_VERSION", "2.1")
        _write_attr(root, "DEEPDISH_IO_VERSION", 12)
        if isinstance(data, dict):
            for k, v in data.items():
                if not isinstance(k, str):
                    raise FormatError("only string keys are stored")
                _save_level(root, k, v)
        else:
            _save_level(root, "data", data)
            _write_attr(root, "DEEPDISH_IO_UNPACK", True)
        L.H5Gclose(root)
    finally:
        L.H5Fclose(f)
    return path


_SYSTEM_ATTRS = {"CLASS", "TITLE", "VERSION", "FLAVOR", "PYTABLES_FORMAT_VERSION", "FILTERS", "EXTDIM",
                 "AUTO_INDEX", "NROWS"}


def _read_typed(read, space: int, t: int, what: str):
    """Read an attribute or dataset whose type is t and dataspace space."""
    L = h5lib()
    nd = L.H5Sget_simple_extent_ndims(space)
    dims = (_hsize * max(nd, 1))()
    if nd > 0:
        L.H5Sget_simple_extent_dims(space, dims, None)
    shape = tuple(int(dims[i]) for i in range(nd))
    cls = L.H5Tget_class(t)
    size = L.H5Tget_size(t)
    if cls == H5T_STRING:
        if L.H5Tis_variable_str(t) > 0:
            raise FormatError(f"{what}: variable-length strings are not read")
        n = int(np.prod(shape)) if shape else 1
        mt = _str_type(size)
        buf = ctypes.create_string_buffer(size * n)
        _ok(read(mt, buf), f"read {what}")
        L.H5Tclose(mt)
        vals = [buf.raw[i * size:(i + 1) * size].split(b"\0", 1)[0].decode("utf-8") for i in range(n)]
        return vals[0] if not shape else np.array(vals).reshape(shape)
    if cls == H5T_INTEGER:
        signed = L.H5Tget_sign(t) != H5T_SGN_NONE
        dt = np.dtype(("i" if signed else "u") + str(size))
    elif cls == H5T_FLOAT:
        dt = np.dtype("f" + str(size))
    elif cls == H5T_BITFIELD and size == 1:
        dt = np.dtype(np.bool_)
    elif cls == H5T_ENUM and size == 1:        # PyTables' boolean enum
        dt = np.dtype(np.bool_)
    else:
        raise FormatError(f"{what}: HDF5 type class {cls} is not read")
    arr = np.empty(shape, dt)
    if arr.size:
        # a one-byte bitfield or enum (booleans) is read in its own type
        mt = _T[_NP2H5[dt]] if dt != np.bool_ else t
        view = arr if dt != np.bool_ else arr.view(np.uint8)
        _ok(read(mt, view.ctypes.data_as(ctypes.c_void_p)), f"read {what}")
    if not shape:
        return arr[()].item() if dt.kind in "biuf" else arr[()]
    return arr


def _attr_names(obj: int):
    L = h5lib()
    names = []

    def cb(_loc, name, _info, _data):
        names.append(name.decode("utf-8"))
        return 0

    fn = _ATTR_CB(cb)
    idx = _hsize(0)
    _ok(L.H5Aiterate2(obj, H5_INDEX_NAME, H5_ITER_INC, ctypes.byref(idx), fn, None), "H5Aiterate2")
    return names


def _read_attr(obj: int, name: str):
    L = h5lib()
    a = _ok(L.H5Aopen(obj, name.encode(), H5P_DEFAULT), f"H5Aopen {name}")
    t = L.H5Aget_type(a)
    sp = L.H5Aget_space(a)
    try:
        return _read_typed(lambda mt, buf: L.H5Aread(a, mt, buf), sp, t, f"attribute {name}")
    finally:
        L.H5Sclose(sp)
        L.H5Tclose(t)
        L.H5Aclose(a)


def _title(obj: int) -> str:
    if h5lib().H5Aexists(obj, b"TITLE") > 0:
        v = _read_attr(obj, "TITLE")
        return v if isinstance(v, str) else ""
    return ""


def _members(g: int):
    L = h5lib()
    names = []

    def cb(_grp, name, _info, _data):
        names.append(name.decode("utf-8"))
        return 0

    fn = _LINK_CB(cb)
    idx = _hsize(0)
    _ok(L.H5Literate(g, H5_INDEX_NAME, H5_ITER_INC, ctypes.byref(idx), fn, None), "H5Literate")
    return names


def _load_node(parent: int, name: str):
    L = h5lib()
    o = _ok(L.H5Oopen(parent, name.encode(), H5P_DEFAULT), f"H5Oopen {name}")
    try:
        kind = L.H5Iget_type(o)
        if kind == H5I_DATASET:
            t = L.H5Dget_type(o)
            sp = L.H5Dget_space(o)
            try:
                return _read_typed(lambda mt, buf: L.H5Dread(o, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf), sp, t,
                                   f"dataset {name}")
            finally:
                L.H5Sclose(sp)
                L.H5Tclose(t)
        if kind == H5I_GROUP:
            return _load_group(o)
        raise FormatError(f"{name}: unsupported HDF5 object")
    finally:
        L.H5Oclose(o)


def _load_group(g: int):
    title = _title(g)
    if title.startswith("nonetype"):
        return None
    out = {n: _load_node(g, n) for n in _members(g)}
    for a in _attr_names(g):
        if a in _SYSTEM_ATTRS or a.startswith("DEEPDISH_IO"):
            continue
        out[a] = _read_attr(g, a)
    if title.startswith("list") or title.startswith("tuple"):
        # entries i0, i1, ...: members (arrays, groups) or attributes (scalars)
        items = [out[f"i{i}"] for i in range(len(out))]
        return items if title.startswith("list") else tuple(items)
    return out


def h5_load(path: str, partial: Optional[str] = None):
    """Read a file written by ``h5_save`` or deepdish.io.save; ``partial`` a
    group path ('/a/b') reads only that node (deepdish.io.load(path, group))."""
    L = h5lib()
    f = _ok(L.H5Fopen(os.fsencode(path), H5F_ACC_RDONLY, H5P_DEFAULT), f"H5Fopen {path}")
    try:
        if partial not in (None, "", "/"):
            p = partial.rstrip("/")
            parent, _, leaf = p.rpartition("/")
            pg = _ok(L.H5Gopen2(f, (parent or "/").encode(), H5P_DEFAULT), f"H5Gopen2 {parent or '/'}")
            try:
                if leaf in _members(pg):
                    return _load_node(pg, leaf)
                if leaf in _attr_names(pg):
                    return _read_attr(pg, leaf)
                raise FormatError(f"{partial!r} not found in {path}")
            finally:
                L.H5Gclose(pg)
        root = _ok(L.H5Gopen2(f, b"/", H5P_DEFAULT), "H5Gopen2 /")
        try:
            unpack = L.H5Aexists(root, b"DEEPDISH_IO_UNPACK") > 0 and bool(_read_attr(root, "DEEPDISH_IO_UNPACK"))
            data = _load_group(root)
        finally:
            L.H5Gclose(root)
        return data["data"] if unpack else data
    finally:
        L.H5Fclose(f)


# --------------------------------------------------------------------------
# object model <-> dict (attr.asdict(recurse=True) / NefBaseClass.from_dict)
# --------------------------------------------------------------------------
def asdict(obj) -> dict:
    """Fields of a dataclass object by name, nested objects as dicts (the
    reference's ``asdict(recurse=True)``); tensors stay as they are."""
    out = {}
    for f in dataclasses.fields(obj):
        v = getattr(obj, f.name)
        out[f.name] = asdict(v) if dataclasses.is_dataclass(v) and not isinstance(v, type) else v
    return out


def from_dict(cls, dct: Optional[dict]):
    """Object of dataclass ``cls`` from a dict (NefBaseClass.from_dict,
    nefct/base/base.py): missing keys take the field default, nested
    dataclass fields are built from their dicts; arrays stored for tensor
    fields come back as CPU tensors."""
    import typing
    if dct is None:
        return cls()
    hints = typing.get_type_hints(cls)
    kwargs = {}
    for f in dataclasses.fields(cls):
        if f.name not in dct:
            continue
        v = dct[f.name]
        tp = hints.get(f.name)
        if isinstance(v, dict) and dataclasses.is_dataclass(tp):
            v = from_dict(tp, v)
        elif isinstance(v, np.ndarray) and _is_tensor_type(tp):
            import torch
            v = torch.from_numpy(np.ascontiguousarray(v))
        elif isinstance(v, np.bool_):
            v = bool(v)
        kwargs[f.name] = v
    return cls(**kwargs)


def _is_tensor_type(tp) -> bool:
    try:
        import torch
    except ImportError:
        return False
    return tp is torch.Tensor


def resource_dir() -> str:
    """Where files without an explicit path go (the reference's RESOURCE_DIR,
    nefct/config.py:23): ``$NEFCT_RESOURCE_DIR`` or ~/.cache/nefct_hip/resources."""
    d = os.environ.get("NEFCT_RESOURCE_DIR") or os.path.join(os.path.expanduser("~"), ".cache", "nefct_hip",
                                                             "resources")
    os.makedirs(d, exist_ok=True)
    return d


def _hash_of_timestamp() -> str:
    """SHA-256 of the current time's decimal string (nefct/utils.py:73-78)."""
    return hashlib.sha256(str(time.time()).encode("utf-8")).hexdigest()


def _new_path(ext: str) -> str:
    return os.path.join(resource_dir(), _hash_of_timestamp() + ext)


# local_io_mixin.py:14-61
def save(obj, path: Optional[str] = None) -> str:
    """Save the object's fields (recursively) to HDF5; returns the path."""
    return h5_save(path or _new_path(".hdf5"), asdict(obj))


def load(cls, path: str, partial: Optional[str] = None):
    if not path:
        raise ValueError("loading path is necessary")
    return from_dict(cls, h5_load(path, partial))


# data_io_mixin.py:15-51
def dump_data(obj):
    """The object with its ``data`` written to its own HDF5 file and replaced
    by that file's path."""
    path = _new_path(".hdf5")
    h5_save(path, _as_array(obj.data))
    return dataclasses.replace(obj, data=path)


def dump_all_data(obj):
    out = {}
    for f in dataclasses.fields(obj):
        v = getattr(obj, f.name)
        if f.name == "data":
            if isinstance(v, str):
                continue
            path = _new_path(".hdf5")
            h5_save(path, _as_array(v))
            out["data"] = path
        elif dataclasses.is_dataclass(v) and not isinstance(v, type):
            out[f.name] = dump_all_data(v)
    return dataclasses.replace(obj, **out)


def _loaded(obj, v):
    import typing
    tp = typing.get_type_hints(type(obj)).get("data")
    if isinstance(v, np.ndarray) and _is_tensor_type(tp):
        import torch
        return torch.from_numpy(np.ascontiguousarray(v))
    return v


def load_data(obj):
    return dataclasses.replace(obj, data=_loaded(obj, h5_load(obj.data)))


def load_all_data(obj):
    out = {}
    for f in dataclasses.fields(obj):
        v = getattr(obj, f.name)
        if f.name == "data":
            if isinstance(v, str):
                out["data"] = _loaded(obj, h5_load(v))
        elif dataclasses.is_dataclass(v) and not isinstance(v, type):
            out[f.name] = load_all_data(v)
    return dataclasses.replace(obj, **out)


# json_io_mixin.py:17-72
def _jsonable(v):
    if isinstance(v, dict):
        return {k: _jsonable(x) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
        return [_jsonable(x) for x in v]
    v = _as_array(v)
    if isinstance(v, np.ndarray):
        return v.tolist()
    if isinstance(v, np.generic):
        return v.item()
    return v


def _json_dict(obj) -> dict:
    dct = asdict(obj)
    if "data" in dct and not isinstance(dct["data"], str) and dct["data"] is not None:
        raise ValueError("please dump data first")
    return _jsonable(dct)


def json_dumps(obj) -> str:
    return json.dumps(_json_dict(obj))


def json_dump(obj, path: Optional[str] = None) -> str:
    path = path or _new_path(".json")
    with open(path, "w") as fout:
        json.dump(_json_dict(obj), fout, indent=4)
    return path


def json_load(cls, path: str):
    with open(path) as fin:
        return from_dict(cls, json.load(fin))


def json_loads(cls, json_str: str):
    return from_dict(cls, json.loads(json_str))


class IOMixin:
    """save/load, dump_data/load_data and the JSON methods as methods, as the
    reference's SaveMixin, LoadMixin, DumpDataMixin, LoadDataMixin,
    JsonDumpMixin and JsonLoadMixin give its classes."""

    def save(self, path: Optional[str] = None) -> str:
        return save(self, path)

    @classmethod
    def load(cls, path: str, partial: Optional[str] = None):
        return load(cls, path, partial)

    def dump_data(self):
        return dump_data(self)

    def dump_all_data(self):
        return dump_all_data(self)

    def load_data(self):
        return load_data(self)

    def load_all_data(self):
        return load_all_data(self)

    def json_dumps(self) -> str:
        return json_dumps(self)

    def json_dump(self, path: Optional[str] = None) -> str:
        return json_dump(self, path)

    @classmethod
    def json_loads(cls, json_str: str):
        return json_loads(cls, json_str)

    @classmethod
    def json_load(cls, path: str):
        return json_load(cls, path)

