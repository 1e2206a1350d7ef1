"""Block meshes: neighbour tables consumed by the fused guard fill.

A uniform quad-tree level of ``2**depth`` x ``2**depth`` blocks
(bsp::uniform_mesh_tree<2>, src/core_bqo_tree.hpp:333-344) is stored
block-major, block ``b = i * nb + j`` for block coordinates ``(i, j)``.  Its
neighbours wrap periodically like next_on/prev_on
(src/mesh_block_index.hpp:334-335).
"""
import numpy as np


class UniformMesh2D:
    def __init__(self, depth, block_size):
        self.depth = int(depth)
        self.nb = 1 << self.depth
        self.n = int(block_size)
        self.nblk = self.nb * self.nb

    def block_coordinates(self):
        b = np.arange(self.nblk)
        return np.stack([b // self.nb, b % self.nb], axis=1).astype(np.int32)

    def neighbours(self):
        """[nblk][4] = {prev on axis 0, next on axis 0, prev on axis 1, next on axis 1}."""
        nb = self.nb
        ij = self.block_coordinates()
        i, j = ij[:, 0], ij[:, 1]
        prev_i = ((i + nb - 1) % nb) * nb + j
        next_i = ((i + 1) % nb) * nb + j
        prev_j = i * nb + (j + nb - 1) % nb
        next_j = i * nb + (j + 1) % nb
        return np.stack([prev_i, next_i, prev_j, next_j], axis=1).astype(np.int32)

    def zones(self):
        return self.nblk * self.n * self.n


class ChainMesh1D:
    """A 1D mesh of ``nblk`` blocks of ``n`` zones.  With ``outflow`` the first
    block's lower and the last block's upper guard zones copy the edge zone
    (src/mara.cpp:123-143); a single block gets only the lower outflow, as in
    the reference (its lower branch returns first)."""

    def __init__(self, nblk, block_size, outflow=True):
        self.nblk = int(nblk)
        self.n = int(block_size)
        self.outflow = bool(outflow)

    def neighbours(self):
        b = np.arange(self.nblk)
        prev = (b - 1) % self.nblk
        nxt = (b + 1) % self.nblk
        if self.outflow:
            prev[0] = -1
            if self.nblk > 1:
                nxt[-1] = -1
        return np.stack([prev, nxt], axis=1).astype(np.int32)

    def zones(self):
        return self.nblk * self.n


def to_device_layout(u_aos):
    """Reference logical layout [nblk, n, n, nc] -> device SoA [nblk, nc, n, n]."""
    return np.ascontiguousarray(np.moveaxis(u_aos, -1, 1))


def to_reference_layout(u_soa):
    """Device SoA [nblk, nc, n(, n)] -> reference logical layout [nblk, n(, n), nc]."""
    return np.ascontiguousarray(np.moveaxis(u_soa, 1, -1))


class EnsembleMesh1D:
    """Many independent 1D problems in one block list.  Problem k has
    ``lengths[k]`` zones (a multiple of the block size ``n``) on [0, 1], so its
    cell spacing is 1 / lengths[k]; its blocks are chained with outflow ends
    (src/mara.cpp:123-143).  Blocks are stored problem after problem."""

    def __init__(self, lengths, block_size):
        self.n = int(block_size)
        self.lengths = np.asarray(lengths, dtype=np.int64)
        if np.any(self.lengths % self.n) or np.any(self.lengths <= 0):
            raise ValueError("every problem length must be a positive multiple of the block size")
        self.blocks_per_problem = self.lengths // self.n
        self.first_block = np.concatenate([[0], np.cumsum(self.blocks_per_problem)[:-1]])
        self.nblk = int(self.blocks_per_problem.sum())

    def neighbours(self):
        nbr = np.empty((self.nblk, 2), dtype=np.int32)
        for f, m in zip(self.first_block, self.blocks_per_problem):
            b = np.arange(f, f + m)
            nbr[f:f + m, 0] = b - 1
            nbr[f:f + m, 1] = b + 1
            nbr[f, 0] = -1
            if m > 1:
                nbr[f + m - 1, 1] = -1
            else:
                nbr[f, 1] = f    # one block: lower outflow only, as a 1-block chain
        return nbr

    def cell_spacing(self):
        return np.repeat(1.0 / self.lengths.astype(np.float64), self.blocks_per_problem)

    def problem_slice(self, k):
        return slice(int(self.first_block[k]), int(self.first_block[k] + self.blocks_per_problem[k]))

    def zones(self):
        return int(self.lengths.sum())
