"""Independent decoders of the four file layouts, used as invariant checkers:
expanding an encoded file must give back exactly the input voxel set.

They follow the READ side of the format (the GLSL traversal's fetch
functions, shaders/octree_dda.frag.glsl:149-345): mirrored reads flip the
child slot / brick coordinate on each axis whose accumulated mirror bit is
set, and a child's mirror bits XOR into the accumulated mask.  Vectorised
level by level with numpy (paths, not nodes, are expanded).
"""
import struct

import numpy as np


def _header(buf: bytes):
    bmin = struct.unpack_from("<3f", buf, 0)
    bmax = struct.unpack_from("<3f", buf, 12)
    root_side, = struct.unpack_from("<f", buf, 24)
    levels, n_nodes = struct.unpack_from("<II", buf, 28)
    return bmin, bmax, root_side, levels, n_nodes


def parse_words(buf: bytes):
    bmin, bmax, rs, levels, n_nodes = _header(buf)
    first_leaf, count = struct.unpack_from("<II", buf, 36)
    words = np.frombuffer(buf, np.uint32, count=count, offset=44)
    assert 44 + 4 * count == len(buf), "trailing bytes"
    return dict(levels=levels, n_nodes=n_nodes, first_leaf=first_leaf, words=words, bmin=bmin, bmax=bmax,
                root_side=rs)


def parse_ssvdag(buf: bytes):
    bmin, bmax, rs, levels, n_nodes = _header(buf)
    n_inner, = struct.unpack_from("<I", buf, 36)
    inner = np.frombuffer(buf, np.uint16, count=n_inner, offset=40)
    o = 40 + 2 * n_inner
    n_leaf, = struct.unpack_from("<I", buf, o)
    leaves = np.frombuffer(buf, np.uint8, count=n_leaf, offset=o + 4)
    o2 = o + 4 + n_leaf
    n_off, = struct.unpack_from("<I", buf, o2)
    offs = np.frombuffer(buf, np.uint32, count=n_off, offset=o2 + 4)
    assert o2 + 4 + 4 * n_off == len(buf), "trailing bytes"
    return dict(levels=levels, n_nodes=n_nodes, inner=inner, leaves=leaves, offsets=offs, bmin=bmin, bmax=bmax,
                root_side=rs)


def _popc8(x):
    x = x.astype(np.uint32) & 0xFF
    return ((x * np.uint32(0x08040201)) >> np.uint32(3) & np.uint32(0x11111111)) % np.uint32(15)


def decode_words(buf: bytes, mirrored: bool) -> np.ndarray:
    """.svdag (mirrored=False) or .ussvdag (mirrored=True) -> sorted Morton codes."""
    f = parse_words(buf)
    L, w = f["levels"], f["words"].astype(np.uint64)
    pos = np.zeros(1, np.uint64)        # word offset of each path's node
    pre = np.zeros(1, np.uint64)        # Morton prefix of each path
    acc = np.zeros(1, np.uint64)        # accumulated flip (4mX|2mY|mZ)
    for lev in range(L):
        hdr = w[pos.astype(np.int64)]
        mask = hdr & np.uint64(0xFF)
        new_pos, new_pre, new_acc = [], [], []
        for g in range(8):                   # geometric child slot
            c = np.uint64(g) ^ acc           # stored slot
            has = ((mask >> c) & np.uint64(1)).astype(bool)
            if not has.any():
                continue
            if lev == L - 1:
                new_pre.append(((pre[has] << np.uint64(3)) | np.uint64(g)))
                continue
            above = mask[has] >> (c[has] + np.uint64(1))            # children with slot > c
            k = _popc8(above).astype(np.uint64)
            ptr = w[(pos[has] + np.uint64(1) + k).astype(np.int64)]
            a = acc[has]
            if mirrored:
                ch = c[has]
                mx = (hdr[has] >> (ch + np.uint64(8))) & np.uint64(1)
                my = (hdr[has] >> (ch + np.uint64(16))) & np.uint64(1)
                mz = (hdr[has] >> (ch + np.uint64(24))) & np.uint64(1)
                a = a ^ ((mx << np.uint64(2)) | (my << np.uint64(1)) | mz)
            new_pos.append(ptr)
            new_pre.append((pre[has] << np.uint64(3)) | np.uint64(g))
            new_acc.append(a)
        if lev == L - 1:
            return np.sort(np.concatenate(new_pre)) if new_pre else np.zeros(0, np.uint64)
        if not new_pos:
            return np.zeros(0, np.uint64)
        pos, pre, acc = np.concatenate(new_pos), np.concatenate(new_pre), np.concatenate(new_acc)
    return np.zeros(0, np.uint64)


def decode_ssvdag(buf: bytes) -> np.ndarray:
    """.ssvdag / .esvdag -> sorted Morton codes."""
    f = parse_ssvdag(buf)
    L, inner, offs = f["levels"], f["inner"].astype(np.uint64), f["offsets"].astype(np.uint64)
    leaves = f["leaves"].reshape(-1, 8)
    bricks = np.zeros(len(leaves), np.uint64)
    for k in range(8):
        bricks |= leaves[:, k].astype(np.uint64) << np.uint64(8 * k)
    idx = np.zeros(1, np.uint64)
    pre = np.zeros(1, np.uint64)
    acc = np.zeros(1, np.uint64)
    for lev in range(L - 2):                    # inner levels 0 .. L-3
        hdr = inner[(idx + offs[lev]).astype(np.int64)]
        new_idx, new_pre, new_acc = [], [], []
        for g in range(8):
            c = np.uint64(g) ^ acc
            cm = (hdr >> (c * np.uint64(2))) & np.uint64(3)
            has = cm != 0
            if not has.any():
                continue
            h, ch, cmh = hdr[has], c[has], cm[has]
            off = np.ones(h.shape, np.uint64)
            for i in range(7, -1, -1):
                sel = np.uint64(i) > ch
                m = (h >> np.uint64(2 * i)) & np.uint64(3)
                off += np.where(sel, np.minimum(m, np.uint64(2)), np.uint64(0))
            at = (offs[lev] + idx[has] + off).astype(np.int64)
            x = inner[at]
            mir = ((x >> np.uint64(13)) & np.uint64(1)) << np.uint64(2) | ((x >> np.uint64(14)) & np.uint64(1)) << np.uint64(1) \
                | ((x >> np.uint64(15)) & np.uint64(1))
            x = x & np.uint64(0x1FFF)
            big = cmh > 1
            if big.any():
                y = inner[np.minimum(at + 1, len(inner) - 1)]
                x = np.where(big, (x << np.uint64(16)) | y | ((cmh & np.uint64(1)) << np.uint64(29)), x)
            new_idx.append(x)
            new_pre.append((pre[has] << np.uint64(3)) | np.uint64(g))
            new_acc.append(acc[has] ^ mir)
        if not new_idx:
            return np.zeros(0, np.uint64)
        idx, pre, acc = np.concatenate(new_idx), np.concatenate(new_pre), np.concatenate(new_acc)
    # 4^3 bricks: stored bit z + 4y + 16x (stored coordinates), mirrored axes read 3 - c
    out = []
    b = bricks[idx.astype(np.int64)]
    mx = (acc >> np.uint64(2)) & np.uint64(1)
    my = (acc >> np.uint64(1)) & np.uint64(1)
    mz = acc & np.uint64(1)
    for gx in range(4):
        for gy in range(4):
            for gz in range(4):
                sx = np.where(mx == 1, np.uint64(3 - gx), np.uint64(gx))
                sy = np.where(my == 1, np.uint64(3 - gy), np.uint64(gy))
                sz = np.where(mz == 1, np.uint64(3 - gz), np.uint64(gz))
                bit = sz + np.uint64(4) * sy + np.uint64(16) * sx
                on = ((b >> bit) & np.uint64(1)).astype(bool)
                if on.any():
                    # geometric position inside the brick -> two Morton levels
                    hi = np.uint64(4 * (gx >> 1) + 2 * (gy >> 1) + (gz >> 1))
                    lo = np.uint64(4 * (gx & 1) + 2 * (gy & 1) + (gz & 1))
                    out.append((pre[on] << np.uint64(6)) | (hi << np.uint64(3)) | lo)
    return np.sort(np.concatenate(out)) if out else np.zeros(0, np.uint64)
