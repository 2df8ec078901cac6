"""Hand-built device-memory images for the gather pass (processDrawables.comp).

An Image is a byte array standing for one device allocation at address `base`.
The builder lays out handle tables (1, 2 or 3 levels of 2048 uint64 entries,
HandleTable.h:33-35), Geometry allocations with PrimitiveSet records
{count, first} (PrimitiveSet.h), MatrixLists {numMatrices, capacity, pad[7]} +
mat4[] (MatrixList.h:54-59) and a DrawableGpuData list (Drawable.h:32-43), and
records the ground truth of every drawable directly from what it placed —
the closed-form expectation the oracle and the kernel are checked against.
"""
from __future__ import annotations

import numpy as np

from cadr_amd._native import DRAWABLE_DTYPE

TABLE_ENTRIES = 2048
TABLE_BYTES = TABLE_ENTRIES * 8


class Image:
    def __init__(self, base: int, size: int):
        self.base = base
        self.buf = np.zeros(size, dtype=np.uint8)
        self.top = 0

    def alloc(self, nbytes: int, align: int = 64) -> int:
        self.top = (self.top + align - 1) // align * align
        addr = self.base + self.top
        self.top += nbytes
        assert self.top <= self.buf.size, "image too small"
        return addr

    def write(self, addr: int, data) -> None:
        b = np.ascontiguousarray(data).view(np.uint8).reshape(-1)
        off = addr - self.base
        self.buf[off: off + b.size] = b

    def read(self, addr: int, dtype, count: int) -> np.ndarray:
        dtype = np.dtype(dtype)
        off = addr - self.base
        return self.buf[off: off + dtype.itemsize * count].view(dtype).copy()


class HandleTables:
    """Tables of a given level for a sparse set of handles (handle -> address)."""

    def __init__(self, img: Image, level: int):
        self.img, self.level = img, level
        self.root = img.alloc(TABLE_BYTES, 64)
        self.children = {}   # (depth, index path) -> table address

    def _table(self, key) -> int:
        if key not in self.children:
            self.children[key] = self.img.alloc(TABLE_BYTES, 64)
        return self.children[key]

    def set(self, handle: int, addr: int) -> None:
        lo = handle & 0x7FF
        if self.level == 1:
            assert handle < TABLE_ENTRIES
            self.img.write(self.root + 8 * lo, np.array([addr], np.uint64))
        elif self.level == 2:
            hi = handle >> 11
            assert hi < TABLE_ENTRIES
            t = self._table((1, hi))
            self.img.write(self.root + 8 * hi, np.array([t], np.uint64))
            self.img.write(t + 8 * lo, np.array([addr], np.uint64))
        else:
            top, mid = handle >> 22, (handle >> 11) & 0x7FF
            assert top < TABLE_ENTRIES
            t2 = self._table((1, top))
            t3 = self._table((2, top, mid))
            self.img.write(self.root + 8 * top, np.array([t2], np.uint64))
            self.img.write(t2 + 8 * mid, np.array([t3], np.uint64))
            self.img.write(t3 + 8 * lo, np.array([addr], np.uint64))


def handle_sequence(level: int, count: int, rng) -> list:
    """Distinct handles for a level, covering the table boundaries the
    reference's HandleTable crosses (HandleTable.cpp:142-338)."""
    if level == 1:
        special = [1, 2, 2046, 2047]
        pool = list(range(1, 2048))
    elif level == 2:
        special = [2047, 2048, 2049, 4095, 4096, 4194303, 4194302, 2048 * 1000 + 5]
        pool = None
    else:
        special = [4194303, 4194304, 4194305, (5 << 22) + (17 << 11) + 3, (2047 << 22) + (2047 << 11) + 2047,
                   (1 << 22) + 2047, (1 << 22) + 2048]
        pool = None
    out = []
    seen = set()
    for h in special:
        if h not in seen and len(out) < count:
            out.append(h); seen.add(h)
    # the rest: runs of consecutive handles from random starting points, as the
    # reference's sequential numbering produces (few tables per thousand handles)
    while len(out) < count:
        if pool is not None:
            h = int(rng.choice(pool))
            if h not in seen:
                out.append(h); seen.add(h)
            continue
        lo, hi = (1, 4194304) if level == 2 else (4194304, 2048 << 22)
        start = int(rng.integers(lo, hi - 5000))
        for h in range(start, start + int(rng.integers(1, 4096))):
            if len(out) >= count:
                break
            if h not in seen:
                out.append(h); seen.add(h)
    return out


def build_gather_scene(base: int, level: int, n: int, rng, num_geometries: int = 8, ps_per_geometry: int = 4,
                       max_matrices: int = 3, with_drawable_data: bool = True, image_size: int | None = None):
    """Returns (image, push (table, list, indirect, pointers), expected indirect
    uint32[n,4], expected pointers uint64[n,4])."""
    num_ml = n if n > 0 else 1
    needed = 2 + num_geometries * 3 + num_ml + (n if with_drawable_data else 0)
    if level == 1:
        assert needed < 2047, "level 1 holds at most 2047 handles"
    handles = handle_sequence(level, needed, rng)
    if level == 1:
        est_tables = 1
    elif level == 2:
        est_tables = 1 + len({h >> 11 for h in handles})
    else:
        est_tables = 1 + len({h >> 22 for h in handles}) + len({h >> 11 for h in handles})
    size = image_size or (est_tables * TABLE_BYTES + num_geometries * 2048 + num_ml * (64 + 64 * max_matrices + 64)
                          + n * (48 + 16 + 32 + 48) + (1 << 16))
    img = Image(base, size)
    tables = HandleTables(img, level)
    hi = iter(handles)

    geoms = []
    for g in range(num_geometries):
        vh, ih, ph = next(hi), next(hi), next(hi)
        va = img.alloc(int(rng.integers(16, 512)), 64)
        ia = img.alloc(int(rng.integers(16, 512)), 64)
        ps = np.zeros((ps_per_geometry, 2), np.uint32)
        ps[:, 0] = rng.integers(1, 100000, ps_per_geometry)
        ps[:, 1] = rng.integers(0, 1 << 31, ps_per_geometry)
        pa = img.alloc(ps.nbytes, 16)
        img.write(pa, ps)
        for h, a in ((vh, va), (ih, ia), (ph, pa)):
            tables.set(h, a)
        geoms.append((vh, ih, ph, va, ia, pa, ps))

    mls = []
    for m in range(num_ml):
        h = next(hi)
        k = int(rng.integers(0, max_matrices + 1))
        hdr = np.zeros(16, np.uint32)
        hdr[0] = k
        hdr[1] = k + int(rng.integers(0, 3))
        mats = rng.standard_normal(16 * max(k, 0)).astype(np.float32)
        a = img.alloc(64 + 64 * k, 64)
        img.write(a, hdr)
        if k:
            img.write(a + 64, mats)
        tables.set(h, a)
        mls.append((h, a, k))

    drawables = np.zeros(n, dtype=DRAWABLE_DTYPE)
    exp_ind = np.zeros((n, 4), np.uint32)
    exp_ptr = np.zeros((n, 4), np.uint64)
    for i in range(n):
        g = geoms[int(rng.integers(0, num_geometries))]
        ml = mls[i]
        p = int(rng.integers(0, ps_per_geometry))
        ddh, dda = 0, 0
        if with_drawable_data and rng.random() < 0.7:
            ddh = next(hi)
            dda = img.alloc(32, 16)
            tables.set(ddh, dda)
        drawables[i] = (g[0], g[1], ml[0], ddh, g[2], 8 * p, 0)
        exp_ind[i] = (g[6][p, 0], ml[2], g[6][p, 1], 0)
        exp_ptr[i] = (g[3], g[4], ml[1], dda)

    dl = img.alloc(max(n, 1) * 48, 64)
    img.write(dl, drawables)
    ind = img.alloc(max(n, 1) * 16, 64)
    ptrs = img.alloc(max(n, 1) * 32, 64)
    return img, (tables.root, dl, ind, ptrs), exp_ind, exp_ptr
