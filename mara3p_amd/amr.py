"""Quad-tree block meshes with level jumps (host-side topology).

The tree is a set of leaf blocks (level, i, j) covering the periodic unit
square.  The reference builds it with bsp::mesh_tree and balances it with
amr::valid_mesh_tree (src/mesh_amr.hpp:407-449) so that edge neighbours differ
by at most one level; here the same 2:1 rule is applied to a leaf set.

Each leaf gets a guard-fill recipe per face direction (-x, +x, -y, +y),
following amr::get_or_create_block (src/mesh_amr.hpp:300-328) applied to the
next_on / prev_on neighbour (src/mesh_block_index.hpp:334-335):
  SAME    the neighbour index is a leaf: copy its zones
  COARSE  its parent is a leaf: refine_block = upsample + split_block
          (piecewise-constant injection from the parent's orthant)
  FINE    it is a node whose four children are leaves: coarsen_blocks =
          tile_blocks + downsample (2 x 2 average)
Recipe row: [kind, leaf, child00, child10, child01, child11, orthant_i, orthant_j]
(children in tile order (0,0), (1,0), (0,1), (1,1); orthant = neighbour index
relative to its parent).
"""
import numpy as np

SAME, COARSE, FINE = 0, 1, 2
# direction d: axis = d // 2, side = -1 for even d, +1 for odd d
DIRECTIONS = ((0, -1), (0, 1), (1, -1), (1, 1))


def _wrap(level, c):
    n = 1 << level
    return (c + n) % n


def neighbour_index(block, axis, side):
    level, i, j = block
    if axis == 0:
        return (level, _wrap(level, i + side), j)
    return (level, i, _wrap(level, j + side))


class QuadTree:
    def __init__(self, leaves):
        self.leaves = sorted(set(leaves))
        self.index = {b: k for k, b in enumerate(self.leaves)}

    @classmethod
    def from_predicate(cls, refine, depth):
        """bsp::mesh_tree (src/core_bqo_tree.hpp:350-362): refine a block while
        refine(block) holds and level < depth."""
        leaves, stack = [], [(0, 0, 0)]
        while stack:
            b = stack.pop()
            if b[0] < depth and refine(b):
                l, i, j = b
                stack.extend([(l + 1, 2 * i + di, 2 * j + dj) for di in (0, 1) for dj in (0, 1)])
            else:
                leaves.append(b)
        return cls(leaves).balanced()

    def contains_leaf(self, b):
        return b in self.index

    def is_node(self, b):
        """True when b is an interior node (a leaf lies strictly below it)."""
        level = b[0]
        return any(l > level and (i >> (l - level), j >> (l - level)) == (b[1], b[2]) for l, i, j in self.leaves)

    def balanced(self):
        """amr::valid_mesh_tree: split leaves whose edge neighbours are refined
        more than one level deeper, until none is."""
        leaves = set(self.leaves)
        while True:
            by_level = {}
            for l, i, j in leaves:
                by_level.setdefault(l, set()).add((i, j))
            maxl = max(by_level)
            split = set()
            for b in leaves:
                l, i, j = b
                for axis, side in DIRECTIONS:
                    nl, ni, nj = neighbour_index(b, axis, side)
                    # any leaf at level >= l + 2 inside the neighbour block?
                    for L in range(l + 2, maxl + 1):
                        s = L - l
                        if any((a >> s, c >> s) == (ni, nj) for a, c in by_level.get(L, ())):
                            split.add(b)
                            break
                    if b in split:
                        break
            if not split:
                return QuadTree(leaves)
            for b in split:
                leaves.remove(b)
                l, i, j = b
                leaves.update((l + 1, 2 * i + di, 2 * j + dj) for di in (0, 1) for dj in (0, 1))

    def recipes(self):
        """[nleaf][4][8] int32 guard-fill recipe (see module docstring)."""
        out = np.full((len(self.leaves), 4, 8), -1, dtype=np.int32)
        for k, b in enumerate(self.leaves):
            for d, (axis, side) in enumerate(DIRECTIONS):
                nb = neighbour_index(b, axis, side)
                l, i, j = nb
                if nb in self.index:
                    out[k, d, :2] = (SAME, self.index[nb])
                elif l > 0 and (l - 1, i >> 1, j >> 1) in self.index:
                    out[k, d, :2] = (COARSE, self.index[(l - 1, i >> 1, j >> 1)])
                    out[k, d, 6:] = (i & 1, j & 1)
                else:
                    kids = [(l + 1, 2 * i + di, 2 * j + dj) for dj in (0, 1) for di in (0, 1)]
                    if not all(c in self.index for c in kids):
                        raise ValueError("quad-tree is not 2:1 balanced at %s" % (nb,))
                    out[k, d, 0] = FINE
                    out[k, d, 2:6] = [self.index[c] for c in kids]
        return out

    def levels(self):
        return np.array([b[0] for b in self.leaves], dtype=np.int32)

    def coordinates(self):
        return np.array([(b[1], b[2]) for b in self.leaves], dtype=np.int32)


def nested_tree(depth, radius=0.5):
    """The reference's 'nested' topology (src/mara.cpp:169-177): refine blocks
    whose centroid lies within radius / level of the centre of [-1/2, 1/2]^2."""
    def refine(b):
        l, i, j = b
        if l == 0:
            return True
        dl = float(1 << l)
        cx = -0.5 + (i + 0.5) / dl
        cy = -0.5 + (j + 0.5) / dl
        return np.hypot(cx, cy) < radius / l
    return QuadTree.from_predicate(refine, depth)
