"""The on-disk form of a run: grouped npz files.

The reference writes a run to NetCDF4 (steam/output.py:30-435) and appends
each nest as a group of the same file (refinements/r0, r1, ...,
steam/simulate.py:3878-3890); netCDF4 is not part of this image, so the same
layout is kept in one npz file: the entries of group ``g`` are stored under
``g|name`` (the root group ``/`` as ``root``), arrays as arrays and everything else
(scalars, strings, tuples) as one JSON entry per group, so the file loads
with ``numpy.load(allow_pickle=False)``.  Lists of records (the class
increments) are stored element by element.  Writes are atomic: the whole
file is rewritten through a temporary file and renamed over the old one.
"""

import json
import os
import tempfile

import numpy as np
import torch

SEP = '|'


def _gname(group):
    g = str(group).strip('/')
    return g or 'root'


def _key(group, name):
    return f"{_gname(group)}{SEP}{name}"


def _to_host(v):
    if isinstance(v, torch.Tensor):
        return v.detach().cpu().numpy()
    return v


def _jsonable(v):
    if isinstance(v, (np.floating,)):
        return float(v)
    if isinstance(v, (np.integer,)):
        return int(v)
    if isinstance(v, np.bool_):
        return bool(v)
    if isinstance(v, (tuple, list)):
        return [_jsonable(x) for x in v]
    return v


def pack(record, group):
    """Entries of one group for ``record`` (a dict of arrays, device tensors,
    scalars, strings, tuples and lists of such dicts)."""
    arrays, meta, lists = {}, {}, {}
    for name, value in record.items():
        value = _to_host(value)
        if isinstance(value, np.ndarray):
            arrays[_key(group, name)] = value
        elif isinstance(value, list) and value and all(isinstance(x, dict) for x in value):
            lists[name] = len(value)
            for i, item in enumerate(value):
                arrays.update(pack(item, f"{_gname(group)}/{name}.{i}"))
        else:
            meta[name] = _jsonable(value)
    arrays[_key(group, '__meta__')] = np.array(json.dumps({'meta': meta, 'lists': lists}))
    return arrays


def unpack(data, group):
    """The record of one group from loaded npz ``data``."""
    g = _gname(group)
    head = json.loads(str(data[_key(group, '__meta__')]))
    out = dict(head['meta'])
    prefix = g + SEP
    for k in data.files:
        if k.startswith(prefix):
            name = k[len(prefix):]
            if name != '__meta__':
                out[name] = np.asarray(data[k])
    for name, n in head['lists'].items():
        out[name] = [unpack(data, f"{g}/{name}.{i}") for i in range(n)]
    return out


def groups(path):
    """Names of the groups a file holds ('/' for the root), list-item
    subgroups left out."""
    with np.load(os.fspath(path), allow_pickle=False) as data:
        names = {k.split(SEP, 1)[0] for k in data.files if k.endswith(SEP + '__meta__')}
    names = {n for n in names if '.' not in n.rsplit('/', 1)[-1]}
    return sorted('/' if n == 'root' else n for n in names)


def write_group(path, group, record, replace_file=False):
    """Write ``record`` as group ``group`` of the npz file at ``path``,
    keeping the file's other groups (all of them dropped with
    ``replace_file``); atomic."""
    path = os.fspath(path)
    entries = {}
    if not replace_file and os.path.exists(path):
        g = _gname(group)
        with np.load(path, allow_pickle=False) as data:
            for k in data.files:
                top = k.split(SEP, 1)[0]
                if top == g or top.startswith(g + '/'):
                    continue
                entries[k] = np.asarray(data[k])
    entries.update(pack(record, group))
    directory = os.path.dirname(os.path.abspath(path)) or '.'
    fd, tmp = tempfile.mkstemp(prefix='.steam-', suffix='.npz', dir=directory)
    try:
        with os.fdopen(fd, 'wb') as handle:
            np.savez(handle, **entries)
            handle.flush()
            os.fsync(handle.fileno())
        os.replace(tmp, path)
    except BaseException:
        if os.path.exists(tmp):
            os.unlink(tmp)
        raise
    return path


def read_group(path, group='/'):
    """The record stored as ``group`` of the file at ``path``."""
    with np.load(os.fspath(path), allow_pickle=False) as data:
        if _key(group, '__meta__') not in data.files:
            raise KeyError(f"{os.fspath(path)} has no group {group!r}")
        return unpack(data, group)
