"""A complete far-field FMM evaluation on the GPU around the translations.

The reference library stops at the operators (its README plans a full FMM);
this module builds what feeds them and what consumes their results, so a
Plummer-sphere potential and force can be evaluated end to end:

  tree      solidfmm_amd.octree.Octree (adaptive, at most `leaf_size` points
            per leaf);
  centres   every cell's expansion centre and radius are the smallest ball
            around its points (sfmm_bounding_spheres_of_ranges, the
            reference's bounding_sphere, sphere.cpp:546-587);
  lists     a dual tree walk with the multipole acceptance criterion
            r_a + r_b < theta |c_a - c_b|: accepted cell pairs become M2L
            items in both directions, leaf pairs that fail it (and every
            leaf with itself) are near-field pairs, and a failed pair is
            split at its larger cell.  Every ordered pair of distinct
            points is covered exactly once, by one M2L or one near pair;
  passes    P2M at the leaves, M2M up level by level, M2L, L2L down level
            by level, L2P at the leaves -- all on the GPU through the
            C ABI (plans and cell sets are built once and reused).

The near field (direct sums over the near pairs) is not part of solidfmm and
is not computed here; `near_pairs` returns it for callers and tests.
"""
import numpy as np

from . import _native as N
from .octree import Octree
from .workload import ITEM_DTYPE, L2L, M2L, M2M


def children_csr(tree):
    """(ptr, kids): the children of cell c are kids[ptr[c]:ptr[c+1]]."""
    n = len(tree)
    child = np.nonzero(tree.parent >= 0)[0]
    par = tree.parent[child]
    o = np.argsort(par, kind="stable")
    kids = child[o]
    ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(ptr, par + 1, 1)
    return np.cumsum(ptr), kids


def cell_balls(tree, points):
    """Smallest ball (x, y, z, r) around the points of every cell; `points`
    in the caller's order, (n, >= 3)."""
    srt = np.ascontiguousarray(np.asarray(points, dtype=np.float64)[tree.order, :3])
    return N.bounding_spheres(srt, tree.first, tree.count)


def _expand(cells, ptr, kids):
    """Children of each cell in `cells`, and the index into `cells` each
    came from."""
    cnt = ptr[cells + 1] - ptr[cells]
    src = np.repeat(np.arange(len(cells)), cnt)
    start = np.repeat(ptr[cells], cnt)
    off = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    return kids[start + off], src


def mac_lists(tree, balls, theta=0.5):
    """Dual tree walk from (root, root).  Returns (m2l_src, m2l_dst, near_src,
    near_dst): directed cell pairs; the near pairs are between leaves and
    include every leaf with itself."""
    ptr, kids = children_csr(tree)
    leaf = tree.leaf
    c = balls[:, :3]
    r = balls[:, 3]
    m2l_s, m2l_t, near_s, near_t = [], [], [], []
    selfs = np.array([0], dtype=np.int64)
    A = np.zeros(0, dtype=np.int64)
    B = np.zeros(0, dtype=np.int64)
    while len(selfs) or len(A):
        if len(selfs):
            lf = leaf[selfs]
            near_s.append(selfs[lf])
            near_t.append(selfs[lf])
            inner = selfs[~lf]
            nk = ptr[inner + 1] - ptr[inner]
            pa, pb = [A], [B]
            for k in np.unique(nk):
                cells = inner[nk == k]
                mat = kids[ptr[cells][:, None] + np.arange(k)[None, :]]
                iu = np.triu_indices(k, 1)
                pa.append(mat[:, iu[0]].ravel())
                pb.append(mat[:, iu[1]].ravel())
            selfs = _expand(inner, ptr, kids)[0]
            A = np.concatenate(pa)
            B = np.concatenate(pb)
            continue
        d = np.sqrt(((c[A] - c[B]) ** 2).sum(1))
        ok = r[A] + r[B] < theta * d
        m2l_s += [A[ok], B[ok]]
        m2l_t += [B[ok], A[ok]]
        A, B = A[~ok], B[~ok]
        both = leaf[A] & leaf[B]
        near_s += [A[both], B[both]]
        near_t += [B[both], A[both]]
        A, B = A[~both], B[~both]
        split_a = ~leaf[A] & (leaf[B] | (r[A] >= r[B]))
        ka, ia = _expand(A[split_a], ptr, kids)
        kb, ib = _expand(B[~split_a], ptr, kids)
        A = np.concatenate([ka, A[~split_a][ib]])
        B = np.concatenate([B[split_a][ia], kb])
    cat = lambda xs: np.concatenate(xs) if xs else np.zeros(0, dtype=np.int64)
    return cat(m2l_s), cat(m2l_t), cat(near_s), cat(near_t)


def cell_orders(tree, p_root, p_leaf):
    """Order per cell, linear in the level from p_root to p_leaf."""
    return tree.orders(p_root, p_leaf)


def fmm_passes(tree, balls, orders, m2l_src, m2l_dst, dim=1):
    """(offsets, pool_len, passes) like octree.translation_lists, with the
    expansion centres at the cells' balls and the M2L items of the MAC
    lists.  passes: [(op, items, src, dst, level)], M2M bottom-up, M2L, L2L
    top-down; items of a pass sorted by target.  dim > 1: every cell holds
    dim blocks (solid<real>(P, dim)) and every cell pair gives one item per
    dimension, as the reference queues them (translations.cpp:654-658)."""
    P = np.asarray(orders, dtype=np.int64)
    size = P * (P + 1)
    off = np.concatenate([[0], np.cumsum(dim * size)[:-1]]).astype(np.int64)
    pool_len = int(dim * size.sum())
    ctr = balls[:, :3]
    child = np.nonzero(tree.parent >= 0)[0]
    par = tree.parent[child]

    def items_for(src, dst):
        o = np.argsort(dst, kind="stable")
        src, dst = src[o], dst[o]
        it = np.zeros((len(src), dim), dtype=ITEM_DTYPE)
        d = np.arange(dim)[None, :]
        it["src"] = off[src][:, None] + d * size[src][:, None]
        it["dst"] = off[dst][:, None] + d * size[dst][:, None]
        it["p_in"] = P[src][:, None]
        it["p_out"] = P[dst][:, None]
        sh = ctr[dst] - ctr[src]
        it["x"], it["y"], it["z"] = sh[:, 0, None], sh[:, 1, None], sh[:, 2, None]
        return it.reshape(-1), np.repeat(src, dim), np.repeat(dst, dim)

    passes = []
    for lev in range(tree.depth, 0, -1):
        sel = tree.level[child] == lev
        passes.append((M2M, *items_for(child[sel], par[sel]), lev))
    passes.append((M2L, *items_for(np.asarray(m2l_src), np.asarray(m2l_dst)), -1))
    for lev in range(1, tree.depth + 1):
        sel = tree.level[child] == lev
        passes.append((L2L, *items_for(par[sel], child[sel]), lev))
    return off, pool_len, passes


def mac_translation_lists(tree, points, theta=0.5, p_root=20, p_leaf=6):
    """The FMM passes of `tree` with MAC lists, in the format of
    octree.translation_lists: (orders, offsets, pool_len, passes)."""
    balls = cell_balls(tree, points)
    s, t, _, _ = mac_lists(tree, balls, theta)
    P = cell_orders(tree, p_root, p_leaf)
    off, pool_len, passes = fmm_passes(tree, balls, P, s, t)
    return P, off, pool_len, passes


class FMM:
    """Far field of the potentials phi_d(x) = sum_j q_{j,d} / |x - x_j| and
    their gradients at every point, on one GPU.

    points: (n, 3 + dim) array of x, y, z and the dim charges (dim 1..4; a
    vortex method's vector potential has dim 3).  The tree, balls, lists,
    plans and cell sets are built once; `evaluate()` runs P2M, M2M, M2L, L2L
    and L2P and returns (phi, grad) as device tensors in the caller's point
    order, of shapes (n,) and (n, 3) for dim 1, else (n, dim) and
    (n, dim, 3), holding the contributions of every source outside the
    point's near pairs."""

    def __init__(self, ctx, points, leaf_size=64, theta=0.5, p_root=None, p_leaf=None, autotune=False):
        import torch
        pts = np.ascontiguousarray(points, dtype=np.float64)
        self.n = len(pts)
        self.dim = dim = pts.shape[1] - 3
        self.tree = tree = Octree(pts[:, :3], leaf_size)
        self.balls = cell_balls(tree, pts)
        self.m2l_src, self.m2l_dst, self.near_src, self.near_dst = mac_lists(tree, self.balls, theta)
        p_root = ctx.P if p_root is None else p_root
        p_leaf = p_root if p_leaf is None else p_leaf
        self.orders = cell_orders(tree, p_root, p_leaf)
        self.off, self.pool_len, self.passes = fmm_passes(tree, self.balls, self.orders, self.m2l_src, self.m2l_dst,
                                                          dim)
        self.ctx = ctx
        if autotune:  # pick the kernel family per order bucket on this tree's own lists
            for op in (M2M, M2L, L2L):
                its = [it for (o, it, _, _, _) in self.passes if o == op and len(it)]
                if its:
                    ctx.autotune(op, np.concatenate(its))
        self.plans = [(op, N.Plan(ctx, op, it)) for (op, it, _, _, _) in self.passes if len(it)]
        leaves = np.nonzero(tree.leaf)[0]
        cells = np.zeros(len(leaves), dtype=N.CELL_DTYPE)
        cells["x"], cells["y"], cells["z"] = (self.balls[leaves, d] for d in range(3))
        cells["offset"] = self.off[leaves]
        cells["first"] = tree.first[leaves]
        cells["count"] = tree.count[leaves]
        cells["order"] = self.orders[leaves]
        cells["dim"] = dim
        self.leaves = leaves
        self.cellset = N.CellSet(ctx, cells, self.n)
        dev = f"cuda:{ctx.device}" if hasattr(ctx, "device") else "cuda"
        self.sorted_points = torch.from_numpy(pts[tree.order]).to(dev)
        self.M = torch.zeros(self.pool_len, dtype=torch.float64, device=dev)
        self.L = torch.zeros(self.pool_len, dtype=torch.float64, device=dev)
        shape = (self.n,) if dim == 1 else (self.n, dim)
        self.phi = torch.zeros(shape, dtype=torch.float64, device=dev)
        self.grad = torch.zeros(shape + (3,), dtype=torch.float64, device=dev)
        self.inverse = torch.from_numpy(np.argsort(tree.order)).to(dev)

    @property
    def translations(self):
        return int(sum(len(p[1]) for p in self.passes))

    def evaluate(self, stream=None, mark=None):
        """Run every pass; `mark(name)`, if given, is called after each phase
        (P2M, M2M, M2L, L2L, L2P) on the launching thread, e.g. to record
        events for phase timings."""
        self.M.zero_()
        self.L.zero_()
        self.phi.zero_()
        self.grad.zero_()
        mark = mark or (lambda name: None)
        mark("start")
        self.cellset.p2m(self.sorted_points, self.M, stream)
        mark("P2M")
        last = None
        for op, plan in self.plans:
            if last is not None and op != last:
                mark({M2M: "M2M", M2L: "M2L", L2L: "L2L"}[last])
            if op == M2M:
                plan.execute(self.M, self.M, stream)
            elif op == M2L:
                plan.execute(self.M, self.L, stream)
            else:
                plan.execute(self.L, self.L, stream)
            last = op
        if last is not None:
            mark({M2M: "M2M", M2L: "M2L", L2L: "L2L"}[last])
        self.cellset.l2p(self.sorted_points, self.L, self.phi, self.grad, stream)
        mark("L2P")
        return self.phi[self.inverse], self.grad[self.inverse]

    def near_field_at(self, points, i):
        """Host direct sum over the near pairs of point i (caller's order):
        (potential, gradient).  For checks on samples, not for speed."""
        tree = self.tree
        if not hasattr(self, "_leaf_of"):
            leaf_of = np.empty(self.n, dtype=np.int64)
            for c in np.nonzero(tree.leaf)[0]:
                leaf_of[tree.first[c]:tree.first[c] + tree.count[c]] = c
            self._leaf_of = leaf_of
            self._rank = np.argsort(tree.order)
            self._srt = np.asarray(points, dtype=np.float64)[tree.order]
        near = self.near_src[self.near_dst == self._leaf_of[self._rank[i]]]
        idx = np.concatenate([np.arange(tree.first[s], tree.first[s] + tree.count[s]) for s in near])
        d = np.asarray(points[i, :3], dtype=np.float64) - self._srt[idx, :3]
        r = np.linalg.norm(d, axis=1)
        r[r == 0] = np.inf
        q = self._srt[idx, 3:]  # (k, dim)
        phi = (q / r[:, None]).sum(0)
        grad = -(q[:, :, None] * (d / r[:, None] ** 3)[:, None, :]).sum(0)
        if self.dim == 1:
            return phi[0], grad[0]
        return phi, grad

    def near_pairs(self):
        """Directed near-field leaf pairs (source, target), self pairs included."""
        return self.near_src, self.near_dst
