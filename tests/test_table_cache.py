"""Host logic of the Hermite-table cache (edge_hip/tables.py): the two buffers a
marched map alternates between.  CPU tensors stand in for device memory; no
kernel runs."""
import torch

from edge_hip import tables
from edge_hip.field import touch


def _map(n=4):
    psi = torch.zeros(3, n, n, n)
    T = tuple(torch.zeros(3, n, n, n) for _ in range(3))
    dT = tuple(torch.zeros(3, n, n, n) for _ in range(3))
    return psi, T, dT


def _current(psi, T, dT, dx):
    key, ver = tables._key_and_version("garm", [psi, *T, *dT], dx)
    ent = tables._cache.get(key)
    return None if ent is None or ent[0] != ver else ent[1]


def test_spare_and_current_tables_alternate():
    tables.clear()
    psi, T, dT = _map()
    dx = 0.25
    key, ver = tables._key_and_version("garm", [psi, *T, *dT], dx)
    first = tables._alloc(psi)
    tables._cache[key] = (ver, first)          # as hermite_table leaves it after a pack
    seen = [first]
    for _ in range(4):
        cur = _current(psi, T, dT, dx)
        spare = tables.spare_table("garm", psi, T, dT, None, dx)
        assert spare is not cur                 # a march never writes the table it reads
        assert tables.spare_table("garm", psi, T, dT, None, dx) is spare
        touch(psi, *T, *dT)                     # the march rewrote the map
        assert _current(psi, T, dT, dx) is None
        tables.install_table("garm", spare, psi, T, dT, None, dx)
        assert _current(psi, T, dT, dx) is spare
        seen.append(spare)
    # two buffers in all, used in turn
    assert len({id(t) for t in seen}) == 2
    assert seen[0] is seen[2] is seen[4] and seen[1] is seen[3]
    tables.clear()


def test_install_without_a_cached_table_keeps_no_spare():
    tables.clear()
    psi, T, dT = _map()
    dx = 0.5
    tab = tables._alloc(psi)
    tables.install_table("garm", tab, psi, T, dT, None, dx)
    assert _current(psi, T, dT, dx) is tab
    spare = tables.spare_table("garm", psi, T, dT, None, dx)
    assert spare is not tab and spare.shape == tab.shape
    tables.clear()
    assert not tables._cache and not tables._spare


def test_eviction_drops_the_spare_with_its_entry():
    tables.clear()
    dx = 1.0
    maps = [_map() for _ in range(tables._MAX_ENTRIES + 1)]
    for psi, T, dT in maps[:-1]:
        tables.install_table("garm", tables._alloc(psi), psi, T, dT, None, dx)
        tables.spare_table("garm", psi, T, dT, None, dx)
    psi0, T0, dT0 = maps[0]
    key0, _ = tables._key_and_version("garm", [psi0, *T0, *dT0], dx)
    assert key0 in tables._spare
    psi, T, dT = maps[-1]
    tables.install_table("garm", tables._alloc(psi), psi, T, dT, None, dx)
    assert key0 not in tables._cache and key0 not in tables._spare
    assert len(tables._cache) == tables._MAX_ENTRIES
    tables.clear()
