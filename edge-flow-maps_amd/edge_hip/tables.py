"""Hermite tables of backward maps, cached by storage and version.

A table packs, per component c and cell, the 8 coefficients the tricubic
Hermite needs into two 16-byte words, (psi_c, T_c) in a low plane and
(dT_c, d2T_c) in a high plane, so one map evaluation reads two 16-byte words
per corner instead of eight planes, and one that needs only psi and T (a face
centre) reads only the low plane.
Three kinds correspond to the reference's three Hermite variants:
  "garm"  (sim/hermite.py:298-456)  dT stored, d2T derived from dT
  "full"  (sim/hermite.py:459-607)  dT and d2T stored
  "fd"    (sim/hermite.py:64-295)   only T stored, dT and d2T derived
"""
import torch

from . import _lib
from .field import as_ptr, stream_handle, version_of

_KIND = {"garm": 0, "full": 1}
_cache = {}
_spare = {}   # per cache key: a second table buffer that a march writes the next map's table into
_MAX_ENTRIES = 6


def _key_and_version(kind, fields, dx):
    ptrs = tuple(as_ptr(f) for f in fields)
    return (kind, ptrs, float(dx)), tuple(version_of(p) for p in ptrs)


def _alloc(psi):
    n = psi.data[0].numel()
    return torch.empty(3 * n * 8, dtype=torch.float32, device=psi.data.device)


def table_version(kind, psi, T, dT=None, d2T=None, dx=None):
    """(key, version) identifying the map's current contents for derived caches."""
    fields = [psi, *T] + (list(dT) if dT is not None else []) + (list(d2T) if d2T is not None else [])
    return _key_and_version(kind, fields, dx)


def hermite_table(kind, psi, T, dT=None, d2T=None, dx=None):
    """Return the (possibly cached) table of the map psi with rows T=(T_x,T_y,T_z)."""
    fields = [psi, *T] + (list(dT) if dT is not None else []) + (list(d2T) if d2T is not None else [])
    key, ver = _key_and_version(kind, fields, dx)
    ent = _cache.get(key)
    if ent is not None and ent[0] == ver:
        return ent[1]
    tab = ent[1] if ent is not None else _alloc(psi)
    nx, ny, nz = psi.shape
    s = stream_handle()
    if kind == "fd":
        # derived dT (calc_fd_3_from_grad) and d2T, then a full-Hermite table
        scratch = _fd_scratch(psi)
        dTs, d2Ts = scratch[:3], scratch[3:]
        _lib.call("edge_calc_fd3", 1, *[as_ptr(t) for t in T], *[t.data_ptr() for t in dTs],
                  *[t.data_ptr() for t in d2Ts], nx, ny, nz, float(dx), s)
        _lib.call("edge_pack_hermite_table", 1, as_ptr(psi), *[as_ptr(t) for t in T], *[t.data_ptr() for t in dTs],
                  *[t.data_ptr() for t in d2Ts], tab.data_ptr(), nx, ny, nz, float(dx), s)
    else:
        k = _KIND[kind]
        d2 = [as_ptr(t) for t in d2T] if d2T is not None else [None, None, None]
        _lib.call("edge_pack_hermite_table", k, as_ptr(psi), *[as_ptr(t) for t in T], *[as_ptr(t) for t in dT], *d2,
                  tab.data_ptr(), nx, ny, nz, float(dx), s)
    if len(_cache) >= _MAX_ENTRIES and key not in _cache:
        evicted = next(iter(_cache))
        _cache.pop(evicted)
        _spare.pop(evicted, None)
    _cache[key] = (ver, tab)
    return tab


def spare_table(kind, psi, T, dT=None, d2T=None, dx=None):
    """A table buffer for the next contents of the map psi (never the one
    hermite_table returns for its current contents), to be filled by a march
    and then registered with install_table."""
    fields = [psi, *T] + (list(dT) if dT is not None else []) + (list(d2T) if d2T is not None else [])
    key, _ = _key_and_version(kind, fields, dx)
    tab = _spare.get(key)
    cur = _cache.get(key)
    if tab is None or (cur is not None and cur[1] is tab):
        tab = _spare[key] = _alloc(psi)
    return tab


def install_table(kind, tab, psi, T, dT=None, d2T=None, dx=None):
    """Register `tab` as the table of the map's current contents (after the
    march that wrote both has touched the fields); the table it replaces
    becomes the spare buffer of the next march."""
    fields = [psi, *T] + (list(dT) if dT is not None else []) + (list(d2T) if d2T is not None else [])
    key, ver = _key_and_version(kind, fields, dx)
    old = _cache.get(key)
    if len(_cache) >= _MAX_ENTRIES and key not in _cache:
        evicted = next(iter(_cache))
        _cache.pop(evicted)
        _spare.pop(evicted, None)
    _cache[key] = (ver, tab)
    if old is not None and old[1] is not tab:
        _spare[key] = old[1]
    else:
        _spare.pop(key, None)


_fd_bufs = {}


def _fd_scratch(psi):
    key = (psi.shape, psi.data.device)
    b = _fd_bufs.get(key)
    if b is None:
        nx, ny, nz = psi.shape
        b = [torch.empty((3, nx, ny, nz), dtype=torch.float32, device=psi.data.device) for _ in range(3)] + \
            [torch.empty((nx, ny, nz), dtype=torch.float32, device=psi.data.device) for _ in range(3)]
        _fd_bufs[key] = b
    return b


def clear():
    _cache.clear()
    _spare.clear()
    _fd_bufs.clear()
