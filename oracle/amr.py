"""Oracle: guard fill and euler2d stages on the leaves of a quad-tree with
level jumps.  TEST INFRASTRUCTURE ONLY: tests load this as the parity reference; the
product path never imports it.

The neighbour of each leaf is assembled as a whole virtual block, in the
order amr::get_or_create_block takes (src/mesh_amr.hpp:300-328):

  * the tree holds a leaf there: that leaf's primitives;
  * the parent is a leaf: refine_block (:258-268) = upsample (:140-166,
    a(i/2, j/2)) then split_block (:204-252) to the orthant
    to_integral(orthant(relative_to_parent(block))) (bit 0 = axis 0,
    src/core_numeric_array.hpp:262-315);
  * the block is a node: collapse (:287-294, recursive) = coarsen_blocks
    (:171-197) = tile_blocks (:47-92) then downsample (:98-137),
    0.25 * (((a(2i,2j) + a(2i,2j+1)) + a(2i+1,2j)) + a(2i+1,2j+1)).

The extended block is then cut as amr::extend_block (:360-393) does: the
n x n leaf plus two guard zones on each side from the edge neighbours
(next_on / prev_on, src/mesh_block_index.hpp:334-335).  Corner guard zones
are set to NaN here: the euler2d update never reads them, and the NaN makes
any read visible.

Pinning: tests/test_amr.py checks the assembled guard zones against the
area average of the piecewise-constant finest-level image of the leaves
(an independent characterisation of injection + 2x2 averaging, exact for
integer-valued data), and the stage against the uniform-mesh oracle on a
tree with no level jumps.
"""
import ctypes

import numpy as np

from . import orc, P

DIRS = ((0, -1), (0, 1), (1, -1), (1, 1))


def _neighbour(block, axis, side):
    l, i, j = block
    m = 1 << l
    return (l, (i + side) % m, j) if axis == 0 else (l, i, (j + side) % m)


def virtual_block(prim, index, block, n):
    """Primitive array [n][n][nc] of `block` per get_or_create_block; `prim`
    holds the leaves' primitives [nleaf][n][n][nc], `index` maps leaf -> row."""
    if block in index:
        return prim[index[block]]
    l, i, j = block
    parent = (l - 1, i >> 1, j >> 1)
    if l > 0 and parent in index:
        up = np.repeat(np.repeat(prim[index[parent]], 2, axis=0), 2, axis=1)
        oi, oj = i & 1, j & 1
        return up[oi * n:(oi + 1) * n, oj * n:(oj + 1) * n]
    kids = {(a, b): virtual_block(prim, index, (l + 1, 2 * i + a, 2 * j + b), n) for a in (0, 1) for b in (0, 1)}
    tile = np.empty((2 * n, 2 * n) + prim.shape[3:])
    for (a, b), v in kids.items():
        tile[a * n:(a + 1) * n, b * n:(b + 1) * n] = v
    return 0.25 * (((tile[0::2, 0::2] + tile[0::2, 1::2]) + tile[1::2, 0::2]) + tile[1::2, 1::2])


def recover_primitives(u_ref, gamma):
    """euler::recover_primitive per zone (oracle C), [..., 5] -> [..., 5]."""
    flat = np.ascontiguousarray(u_ref.reshape(-1, 5))
    out = np.empty_like(flat)
    f = orc().orc_euler_recover_primitive
    for k in range(flat.shape[0]):
        if f(P(flat[k]), gamma, P(out[k])) != 0:
            raise ValueError("negative density in zone %d" % k)
    return out.reshape(u_ref.shape)


def extend_leaves(prim, leaves, n):
    """Extended primitive blocks [nleaf][n+4][n+4][nc] (corners NaN)."""
    index = {b: k for k, b in enumerate(leaves)}
    nc = prim.shape[-1]
    pe = np.full((len(leaves), n + 4, n + 4, nc), np.nan)
    for k, b in enumerate(leaves):
        pe[k, 2:n + 2, 2:n + 2] = prim[k]
        for axis, side in DIRS:
            v = virtual_block(prim, index, _neighbour(b, axis, side), n)
            if axis == 0:
                if side < 0:
                    pe[k, 0:2, 2:n + 2] = v[n - 2:n]
                else:
                    pe[k, n + 2:n + 4, 2:n + 2] = v[0:2]
            else:
                if side < 0:
                    pe[k, 2:n + 2, 0:2] = v[:, n - 2:n]
                else:
                    pe[k, 2:n + 2, n + 2:n + 4] = v[:, 0:2]
    return pe


def stage(u_ref, leaves, levels, n, dt, theta, gamma, hllc=0, domain_size=1.0):
    """One euler2d explicit-Euler stage over every leaf (reference layout
    [nleaf][n][n][5], conserved); each leaf uses dl = L / n / 2^level
    (src/mesh_geometry.cpp:168-171)."""
    prim = recover_primitives(u_ref, gamma)
    pe = extend_leaves(prim, leaves, n)
    out = np.empty_like(u_ref)
    f = orc().orc_euler2d_updated_conserved
    for k in range(len(leaves)):
        dl = domain_size / float(n) / float(1 << int(levels[k]))
        uc = np.ascontiguousarray(u_ref[k])
        ext = np.ascontiguousarray(pe[k])
        o = np.empty_like(uc)
        st = f(P(uc), P(ext), n, n, dt, dl, theta, gamma, int(hllc), P(o))
        if st != 0:
            raise ValueError("oracle euler2d update failed with status %d on leaf %d" % (st, k))
        out[k] = o
    return out


def rk2(u_ref, leaves, levels, n, dt, theta, gamma, hllc=0, steps=1):
    """RK2 as control::advance_runge_kutta (src/app_control.hpp:169-195):
    u1 = L(u0); u = u0 * 0.5 + L(u1) * (1 - 0.5)."""
    u = u_ref
    for _ in range(steps):
        u1 = stage(u, leaves, levels, n, dt, theta, gamma, hllc)
        u2 = stage(u1, leaves, levels, n, dt, theta, gamma, hllc)
        u = u * 0.5 + u2 * (1.0 - 0.5)
    return u


# --------------------------------------------------------------------------- conservative mode
def edge_fluxes(pe_k, n, theta, gamma, hllc=0):
    """Face fluxes on the four edges of one leaf from its extended primitive
    block [n+4][n+4][5]: [4][n][5] for the edges -x (face row 0), +x (face
    row n), -y (face column 0), +y (face column n), each face from the PLM
    states lower + g/2 and upper - g/2 (src/module_euler.cpp:274-295)."""
    riemann = orc().orc_euler_riemann_hllc if hllc else orc().orc_euler_riemann_hlle
    plm = orc().orc_plm
    out = np.zeros((4, n, 5))

    def face(c0, c1, c2, c3, axis):
        pl = np.array([c1[q] + 0.5 * plm(c0[q], c1[q], c2[q], theta) for q in range(5)])
        pr = np.array([c2[q] - 0.5 * plm(c1[q], c2[q], c3[q], theta) for q in range(5)])
        f = np.zeros(5)
        riemann(P(pl), P(pr), axis, gamma, P(f))
        return f

    for i in range(n):
        e = i + 2
        out[0, i] = face(pe_k[0, e], pe_k[1, e], pe_k[2, e], pe_k[3, e], 1)
        out[1, i] = face(pe_k[n, e], pe_k[n + 1, e], pe_k[n + 2, e], pe_k[n + 3, e], 1)
        out[2, i] = face(pe_k[e, 0], pe_k[e, 1], pe_k[e, 2], pe_k[e, 3], 2)
        out[3, i] = face(pe_k[e, n], pe_k[e, n + 1], pe_k[e, n + 2], pe_k[e, n + 3], 2)
    return out


def apply_flux_correction(out, fluxes, leaves, levels, n, dt, w, domain_size=1.0):
    """For every leaf edge bordering finer leaves, replace the leaf's flux
    through each edge face by the mean of the two finer face fluxes covering
    it: the zone next to the edge changes by s dt / dl (mean - own) * w,
    s = +1 on the lower edges, -1 on the upper; x edges first, then y edges.
    out: [nleaf][n][n][5] (modified); fluxes: [nleaf][4][n][5]."""
    index = {b: k for k, b in enumerate(leaves)}
    for axis in (0, 1):
        for k, b in enumerate(leaves):
            dl = domain_size / float(n) / float(1 << int(levels[k]))
            for side in (-1, 1):
                nb = _neighbour(b, axis, side)
                l, i, j = nb
                if nb in index or (l > 0 and (l - 1, i >> 1, j >> 1) in index):
                    continue
                e = 2 * axis + (side > 0)
                across = 1 if side < 0 else 0          # the children's position facing this leaf
                for m in range(n):
                    along = (2 * m) // n
                    fi = 2 * m - along * n
                    kid = (l + 1, 2 * i + across, 2 * j + along) if axis == 0 else \
                          (l + 1, 2 * i + along, 2 * j + across)
                    fe = 2 * axis + (side < 0)         # the child's edge facing this leaf
                    fr = fluxes[index[kid], fe]
                    mean = 0.5 * (fr[fi] + fr[fi + 1])
                    own = fluxes[k, e, m]
                    corr = (1.0 if side < 0 else -1.0) * ((mean - own) * dt / dl)
                    r, c = (0 if side < 0 else n - 1, m) if axis == 0 else (m, 0 if side < 0 else n - 1)
                    out[k, r, c] = out[k, r, c] + corr * w
    return out


def stage_conservative(u_ref, leaves, levels, n, dt, theta, gamma, hllc=0, domain_size=1.0):
    """stage() followed by the flux correction at level jumps."""
    out = stage(u_ref, leaves, levels, n, dt, theta, gamma, hllc, domain_size)
    pe = extend_leaves(recover_primitives(u_ref, gamma), leaves, n)
    fl = np.array([edge_fluxes(pe[k], n, theta, gamma, hllc) for k in range(len(leaves))])
    return apply_flux_correction(out, fl, leaves, levels, n, dt, 1.0, domain_size)


def rk2_conservative(u_ref, leaves, levels, n, dt, theta, gamma, hllc=0, steps=1):
    """RK2 of the conservative stages; the second stage's correction is
    applied after the combine u * 0.5 + L(u1) * (1 - 0.5), weighted 1 - 0.5."""
    u = u_ref
    for _ in range(steps):
        u1 = stage_conservative(u, leaves, levels, n, dt, theta, gamma, hllc)
        u2 = stage(u1, leaves, levels, n, dt, theta, gamma, hllc)
        comb = u * 0.5 + u2 * (1.0 - 0.5)
        pe = extend_leaves(recover_primitives(u1, gamma), leaves, n)
        fl = np.array([edge_fluxes(pe[k], n, theta, gamma, hllc) for k in range(len(leaves))])
        u = apply_flux_correction(comb, fl, leaves, levels, n, dt, 1.0 - 0.5)
    return u
