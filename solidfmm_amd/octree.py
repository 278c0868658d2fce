"""Synthetic adaptive FMM workload: Plummer points, an adaptive octree and
the translation lists (M2M, M2L, L2L) of its cells.

The reference library has no tree (its README plans one); this module
exists to feed the serving loop with realistic, mixed-order translation
requests.  Everything is seeded and vectorised with numpy.

Points.  Plummer sphere with scale radius a (Plummer 1911; sampling as in
Aarseth, Henon and Wielen 1974): r = a / sqrt(u^(-2/3) - 1) for u ~ U(0,1),
isotropic directions; radii are capped at 100 a so the root cube stays
finite.

Tree.  Morton keys of 21 bits per axis on the bounding cube; a cell is split
while it holds more than `leaf_size` points (depth <= 21).  Cells are stored
level by level with integer coordinates (ix, iy, iz) at their level.

Lists.  M2M: child -> parent, one item per non-root cell.  L2L: parent ->
child.  M2L: the V-list of each cell b of level >= 2, i.e. the cells c of
the same level with |c - b|_inf >= 2 whose parents are neighbours of b's
parent (|parent(c) - parent(b)|_inf <= 1).  Shifts are new centre minus old
centre.  Orders are set per level: P(level) falls linearly from p_root at
the root to p_leaf at the deepest level.
"""
import numpy as np

from .workload import ITEM_DTYPE, M2L, M2M, L2L, uniform01

MORTON_BITS = 21


def plummer(n, seed, a=1.0, rmax=100.0):
    """n points of a Plummer sphere of scale radius a (seeded)."""
    u = np.maximum(uniform01(seed, 100, n), 1e-300)
    r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    r = np.minimum(r, rmax * a)
    z = 2.0 * uniform01(seed, 101, n) - 1.0
    phi = 2.0 * np.pi * uniform01(seed, 102, n)
    s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([r * s * np.cos(phi), r * s * np.sin(phi), r * z], axis=1)


def _spread(v):
    """Insert two zero bits between the low 21 bits of v (uint64)."""
    v = v.astype(np.uint64) & np.uint64(0x1FFFFF)
    v = (v | (v << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
    v = (v | (v << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
    v = (v | (v << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
    v = (v | (v << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
    v = (v | (v << np.uint64(2))) & np.uint64(0x1249249249249249)
    return v


def _compact(v):
    v = v.astype(np.uint64) & np.uint64(0x1249249249249249)
    v = (v ^ (v >> np.uint64(2))) & np.uint64(0x10C30C30C30C30C3)
    v = (v ^ (v >> np.uint64(4))) & np.uint64(0x100F00F00F00F00F)
    v = (v ^ (v >> np.uint64(8))) & np.uint64(0x1F0000FF0000FF)
    v = (v ^ (v >> np.uint64(16))) & np.uint64(0x1F00000000FFFF)
    v = (v ^ (v >> np.uint64(32))) & np.uint64(0x1FFFFF)
    return v


class Octree:
    """Adaptive octree over `points` with at most `leaf_size` points per leaf.

    Cell arrays (length = number of cells, root first, then level by
    level): level, ix/iy/iz (integer coordinates at the level), parent
    (-1 for the root), center (x,y,z), first/count (range in the
    Morton-sorted point order), leaf flag.
    """

    def __init__(self, points, leaf_size=64):
        pts = np.asarray(points, dtype=np.float64)
        lo = pts.min(axis=0)
        hi = pts.max(axis=0)
        self.side = float((hi - lo).max()) * (1.0 + 1e-12) or 1.0
        self.origin = (lo + hi) / 2.0 - self.side / 2.0
        scale = (1 << MORTON_BITS) / self.side
        q = np.clip(((pts - self.origin) * scale).astype(np.int64), 0, (1 << MORTON_BITS) - 1)
        key = _spread(q[:, 0]) | (_spread(q[:, 1]) << np.uint64(1)) | (_spread(q[:, 2]) << np.uint64(2))
        order = np.argsort(key, kind="stable")
        self.order = order
        self.keys = key[order]
        n = len(pts)

        levels, prefixes, parents, firsts, counts, leaf = [], [], [], [], [], []
        # level 0: the root
        cur_first = np.array([0], dtype=np.int64)
        cur_count = np.array([n], dtype=np.int64)
        cur_prefix = np.array([0], dtype=np.uint64)
        cur_index = np.array([0], dtype=np.int64)
        levels.append(np.zeros(1, dtype=np.int64))
        prefixes.append(cur_prefix)
        parents.append(np.array([-1], dtype=np.int64))
        firsts.append(cur_first)
        counts.append(cur_count)
        ncells = 1
        split_flags = []
        for lev in range(MORTON_BITS):
            split = cur_count > leaf_size
            split_flags.append(split)
            if not split.any():
                break
            # points of the split cells, with their prefix at level lev+1
            sel_first = cur_first[split]
            sel_count = cur_count[split]
            sel_index = cur_index[split]
            idx = np.repeat(sel_first, sel_count) + (np.arange(sel_count.sum()) -
                                                     np.repeat(np.cumsum(sel_count) - sel_count, sel_count))
            shift = np.uint64(3 * (MORTON_BITS - lev - 1))
            pre = self.keys[idx] >> shift
            par = np.repeat(sel_index, sel_count)
            # children: runs of equal prefix (points are sorted, so runs are contiguous)
            start = np.concatenate([[True], pre[1:] != pre[:-1]])
            pos = np.nonzero(start)[0]
            ch_first = idx[pos]
            ch_count = np.diff(np.concatenate([pos, [len(pre)]]))
            ch_prefix = pre[pos]
            ch_parent = par[pos]
            k = len(pos)
            ch_index = ncells + np.arange(k, dtype=np.int64)
            ncells += k
            levels.append(np.full(k, lev + 1, dtype=np.int64))
            prefixes.append(ch_prefix)
            parents.append(ch_parent)
            firsts.append(ch_first)
            counts.append(ch_count)
            cur_first, cur_count, cur_prefix, cur_index = ch_first, ch_count, ch_prefix, ch_index
        self.level = np.concatenate(levels)
        prefix = np.concatenate(prefixes)
        self.parent = np.concatenate(parents)
        self.first = np.concatenate(firsts)
        self.count = np.concatenate(counts)
        self.ix = _compact(prefix).astype(np.int64)
        self.iy = _compact(prefix >> np.uint64(1)).astype(np.int64)
        self.iz = _compact(prefix >> np.uint64(2)).astype(np.int64)
        has_child = np.zeros(len(self.level), dtype=bool)
        has_child[self.parent[self.parent >= 0]] = True
        self.leaf = ~has_child
        width = self.side / (2.0 ** self.level)
        self.center = np.stack([self.origin[d] + (c + 0.5) * width for d, c in
                                enumerate((self.ix, self.iy, self.iz))], axis=1)
        self.depth = int(self.level.max())

    def __len__(self):
        return len(self.level)

    def orders(self, p_root=20, p_leaf=6):
        """Order per cell: linear in the level from p_root (root) to p_leaf
        (deepest level)."""
        if self.depth == 0:
            return np.full(len(self), p_root, dtype=np.int64)
        t = self.level / self.depth
        return np.rint(p_root + (p_leaf - p_root) * t).astype(np.int64)

    def v_list(self):
        """(targets, sources) of the M2L interactions (V-lists), as arrays of
        cell indices."""
        tg, sr = [], []
        for lev in range(2, self.depth + 1):
            cells = np.nonzero(self.level == lev)[0]
            if len(cells) == 0:
                continue
            span = np.int64(1) << np.int64(lev)
            code = (self.ix[cells] * span + self.iy[cells]) * span + self.iz[cells]
            srt = np.argsort(code)
            code_sorted = code[srt]
            cells_sorted = cells[srt]
            bx, by, bz = self.ix[cells], self.iy[cells], self.iz[cells]
            for dx in range(-3, 4):
                for dy in range(-3, 4):
                    for dz in range(-3, 4):
                        if max(abs(dx), abs(dy), abs(dz)) < 2:
                            continue
                        cx, cy, cz = bx + dx, by + dy, bz + dz
                        ok = (cx >= 0) & (cy >= 0) & (cz >= 0) & (cx < span) & (cy < span) & (cz < span)
                        # parents adjacent
                        ok &= (np.abs((cx >> 1) - (bx >> 1)) <= 1) & (np.abs((cy >> 1) - (by >> 1)) <= 1) & \
                              (np.abs((cz >> 1) - (bz >> 1)) <= 1)
                        if not ok.any():
                            continue
                        want = (cx * span + cy) * span + cz
                        pos = np.searchsorted(code_sorted, want[ok])
                        pos = np.minimum(pos, len(code_sorted) - 1)
                        hit = code_sorted[pos] == want[ok]
                        tg.append(cells[ok][hit])
                        sr.append(cells_sorted[pos[hit]])
        if not tg:
            return np.zeros(0, dtype=np.int64), np.zeros(0, dtype=np.int64)
        return np.concatenate(tg), np.concatenate(sr)


def translation_lists(tree, p_root=20, p_leaf=6):
    """Item arrays of the FMM passes over `tree`, with pool offsets.

    Returns (orders, offsets, pool_len, passes): orders and offsets per cell
    (each cell owns one block of P(P+1) reals in the multipole pool and one
    at the same offset in the local pool), and passes, a list of (op,
    items, level) in FMM order: M2M bottom-up level by level, then M2L, then
    L2L top-down level by level.  Items of each pass are sorted by target.
    """
    P = tree.orders(p_root, p_leaf)
    size = P * (P + 1)
    off = np.concatenate([[0], np.cumsum(size)[:-1]]).astype(np.int64)
    pool_len = int(size.sum())
    child = np.nonzero(tree.parent >= 0)[0]
    par = tree.parent[child]

    def items_for(src, dst, shift):
        it = np.zeros(len(src), dtype=ITEM_DTYPE)
        it["src"] = off[src]
        it["dst"] = off[dst]
        it["p_in"] = P[src]
        it["p_out"] = P[dst]
        it["x"] = shift[:, 0]
        it["y"] = shift[:, 1]
        it["z"] = shift[:, 2]
        o = np.argsort(dst, kind="stable")
        return it[o], src[o], dst[o]

    passes = []
    for lev in range(tree.depth, 0, -1):  # M2M: child (lev) -> parent (lev-1)
        sel = tree.level[child] == lev
        s, d = child[sel], par[sel]
        it, s, d = items_for(s, d, tree.center[d] - tree.center[s])
        passes.append((M2M, it, s, d, lev))
    tg, sr = tree.v_list()
    it, s, d = items_for(sr, tg, tree.center[tg] - tree.center[sr])
    passes.append((M2L, it, s, d, -1))
    for lev in range(1, tree.depth + 1):  # L2L: parent (lev-1) -> child (lev)
        sel = tree.level[child] == lev
        s, d = par[sel], child[sel]
        it, s, d = items_for(s, d, tree.center[d] - tree.center[s])
        passes.append((L2L, it, s, d, lev))
    return P, off, pool_len, passes
