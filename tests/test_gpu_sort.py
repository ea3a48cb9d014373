"""The cell sort's radix sort (sde_sort_pairs_device, the sort every
gridding, weighting and degridding pass runs on its cell keys) against
numpy's stable argsort of the same keys: keys and values equal, bit for
bit, over tile-sized edges (one key, one tile of 4096 keys, one more),
keys of one value, sorted and reversed runs, few distinct keys, every
width of the key from 1 to 32 bits, flag bits above the key width that
ride along, keys not aligned to 16 bytes, and an input of 2e7 keys."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

sde = pytest.importorskip("sde_amd")
from sde_amd import DeviceArray  # noqa: E402


def _sort(keys, vals, nbits, offset=0):
    """offset: the keys start that many elements into their device array
    (offset 1: not 16-byte aligned)."""
    L = sde.lib()
    n = len(keys)
    dk = DeviceArray.from_numpy(np.concatenate([np.zeros(offset, np.uint32), keys])) if n else \
        DeviceArray((1,), np.uint32)
    dv = DeviceArray.from_numpy(vals) if n else DeviceArray((1,), np.uint64)
    ko = DeviceArray((max(n, 1),), np.uint32)
    vo = DeviceArray((max(n, 1),), np.uint64)
    rc = L.sde_sort_pairs_device(C.c_void_p(dk.ptr + 4 * offset), C.c_void_p(ko.ptr), C.c_void_p(dv.ptr),
                                 C.c_void_p(vo.ptr), C.c_int64(n), C.c_int(nbits))
    if rc != 0:
        msg = sde.last_error()
        sde.clear_error()
        raise sde.SdeError(msg)
    return ko.numpy()[:n], vo.numpy()[:n]


def _check(keys, nbits, seed=0, label="", offset=0):
    keys = np.ascontiguousarray(keys, np.uint32)
    n = len(keys)
    rng = np.random.default_rng(seed)
    vals = (np.arange(n, dtype=np.uint64) | (rng.integers(0, 1 << 31, n, dtype=np.uint64) << np.uint64(32)))
    ko, vo = _sort(keys, vals, nbits, offset)
    mask = np.uint32(0xFFFFFFFF) if nbits == 32 else np.uint32((1 << nbits) - 1)
    order = np.argsort(keys & mask, kind="stable")
    np.testing.assert_array_equal(ko, keys[order], err_msg=label + " keys")
    np.testing.assert_array_equal(vo, vals[order], err_msg=label + " values")


@pytest.mark.parametrize("n", [0, 1, 2, 63, 64, 65, 4095, 4096, 4097, 8192, 12289, 100003])
def test_sizes(n):
    rng = np.random.default_rng(n)
    _check(rng.integers(0, 1 << 24, n, dtype=np.uint32), 24, seed=n, label="n=%d" % n)


@pytest.mark.parametrize("nbits", [1, 7, 8, 9, 13, 16, 17, 24, 25, 31, 32])
def test_key_widths(nbits):
    rng = np.random.default_rng(100 + nbits)
    n = 300_000
    hi = 1 << nbits
    keys = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
    _check(keys, nbits, seed=nbits, label="nbits=%d" % nbits)


def test_flag_bits_ride_along():
    """Bits at and above nbits do not order the keys and come out as they went
    in (the prep pass keeps a flag in bit 31 of the 24-bit cell key)."""
    rng = np.random.default_rng(5)
    n = 250_000
    keys = rng.integers(0, 1 << 24, n, dtype=np.uint32) | (rng.integers(0, 2, n, dtype=np.uint32) << np.uint32(31))
    keys |= rng.integers(0, 2, n, dtype=np.uint32) << np.uint32(27)
    _check(keys, 24, label="flags")


@pytest.mark.parametrize("kind", ["equal", "sorted", "reversed", "few", "sawtooth"])
def test_key_patterns(kind):
    n = 500_000
    rng = np.random.default_rng(9)
    if kind == "equal":
        keys = np.full(n, 0x00ABCDEF, np.uint32)
    elif kind == "sorted":
        keys = np.sort(rng.integers(0, 1 << 24, n, dtype=np.uint32))
    elif kind == "reversed":
        keys = np.sort(rng.integers(0, 1 << 24, n, dtype=np.uint32))[::-1].copy()
    elif kind == "few":
        keys = rng.choice(np.array([3, 0x10000, 0xFFFFFF, 0x800080], np.uint32), n)
    else:
        keys = (np.arange(n, dtype=np.uint32) * np.uint32(7919)) % np.uint32(4099)
    _check(keys, 24, label=kind)


def test_cell_sort_size():
    """2e7 keys over the 24-bit cell keys of a 4096^2 field (4883 tiles, not a
    multiple of the XCD count), skewed to the inner cells as a uv coverage is."""
    rng = np.random.default_rng(77)
    n = 20_000_000
    r = np.abs(rng.standard_normal(n)) * 0.25
    keys = np.minimum((r * (1 << 24)).astype(np.int64), (1 << 24) - 1).astype(np.uint32)
    _check(keys, 24, label="cell sort size")


@pytest.mark.parametrize("n", [5000, 300_001])
def test_unaligned_keys(n):
    """Keys that do not start on 16 bytes are counted one load a key."""
    rng = np.random.default_rng(n + 3)
    _check(rng.integers(0, 1 << 24, n, dtype=np.uint32), 24, label="unaligned", offset=1)


def test_bad_arguments():
    with pytest.raises(sde.SdeError):
        _sort(np.zeros(4, np.uint32), np.zeros(4, np.uint64), 33)
    with pytest.raises(sde.SdeError):
        _sort(np.zeros(4, np.uint32), np.zeros(4, np.uint64), 0)
