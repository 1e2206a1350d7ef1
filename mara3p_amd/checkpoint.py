"""Checkpoints in the layout of the reference's HDF5 checkpoints.

The reference writes chkpt.NNNN.h5 (mara::problem_base_t::write_checkpoint,
src/app_problem.cpp:55-78) holding "git_commit", "run_config", "schedule",
"module" and a "solution" group with "iteration", "time" and one dataset per
block under "solution/conserved", named by mesh::format_block_index
(src/mesh_block_index.hpp:353-375) and holding the block's conserved tuples
in logical component order (h5::write of the solution,
src/module_euler_io.cpp:81-102).

HDF5 is not available to this package, so the same hierarchy is stored as a
NumPy .npz archive whose keys are the HDF5 paths ("solution/time",
"solution/conserved/4:03-12", "run_config/cfl", ...).  Block datasets use
the reference's logical layout [n, n, n_c] (1D: [n, n_c]).  A converter to
HDF5 only has to create one dataset per key.
"""
import math

import numpy as np

FORMAT_VERSION = 1


def format_block_index(level, *coords):
    """mesh::format_block_index: "level:i-j" with every coordinate zero-padded
    to the width 1 + log10(2^level) (truncated), src/mesh_block_index.hpp:353-375."""
    width = int(1 + math.log10(1 << level))
    return "%d:" % level + "-".join("%0*d" % (width, c) for c in coords)


def read_block_index(text, rank):
    """mesh::read_block_index (src/mesh_block_index.hpp:388-408)."""
    if text.count("-") != rank - 1:
        raise ValueError("read_block_index (string has wrong rank)")
    level, rest = text.split(":", 1)
    coords = tuple(int(x) for x in rest.split("-"))
    return (int(level),) + coords


def write(path, blocks, arrays, iteration, time, run_config=None, schedule=None, module="", git_commit=""):
    """Write one checkpoint.  blocks: block indices (level, i[, j]); arrays:
    the conserved state in the device layout [nblk, n_c, n(, n)]."""
    arrays = np.asarray(arrays, dtype=np.float64)
    if len(blocks) != arrays.shape[0]:
        raise ValueError("checkpoint: one block index per block array")
    rank = arrays.ndim - 2
    out = {"format_version": np.int64(FORMAT_VERSION), "git_commit": np.str_(git_commit), "module": np.str_(module),
           "solution/iteration": np.int64(iteration), "solution/time": np.float64(time)}
    for index, a in zip(blocks, arrays):
        if len(index) != rank + 1:
            raise ValueError("checkpoint: block index rank does not match the arrays")
        out["solution/conserved/" + format_block_index(*index)] = np.ascontiguousarray(np.moveaxis(a, 0, -1))
    for k, v in (run_config or {}).items():
        out["run_config/" + k] = np.asarray(v)
    for k, v in (schedule or {}).items():
        out["schedule/" + k] = np.asarray(v)
    np.savez(path, **out)


def read(path, rank):
    """Read a checkpoint: returns dict with iteration, time, module,
    run_config, schedule and conserved = {block index: [n_c, n(, n)] array}."""
    with np.load(path, allow_pickle=False) as z:
        res = {"iteration": int(z["solution/iteration"]), "time": float(z["solution/time"]),
               "module": str(z["module"]), "git_commit": str(z["git_commit"]),
               "run_config": {}, "schedule": {}, "conserved": {}}
        for key in z.files:
            if key.startswith("solution/conserved/"):
                idx = read_block_index(key[len("solution/conserved/"):], rank)
                res["conserved"][idx] = np.ascontiguousarray(np.moveaxis(z[key], -1, 0))
            elif key.startswith("run_config/"):
                res["run_config"][key[len("run_config/"):]] = z[key][()]
            elif key.startswith("schedule/"):
                res["schedule"][key[len("schedule/"):]] = z[key][()]
    return res


def solver_blocks(solver):
    """Block indices of a device-resident solver, in its storage order."""
    if hasattr(solver, "tree"):
        return list(solver.tree.leaves)
    mesh = solver.mesh
    if hasattr(mesh, "block_coordinates"):
        return [(mesh.depth, int(i), int(j)) for i, j in mesh.block_coordinates()]
    if hasattr(mesh, "lengths"):
        raise ValueError("checkpoint: an ensemble of independent problems has no single block tree")
    depth = int(round(math.log2(mesh.nblk)))
    if 1 << depth != mesh.nblk:
        raise ValueError("checkpoint: a 1D chain must have 2^depth blocks")
    return [(depth, i) for i in range(mesh.nblk)]


def save_solver(path, solver, iteration, time, run_config=None, schedule=None, module=""):
    write(path, solver_blocks(solver), solver.state(), iteration, time, run_config, schedule, module)


def restore_solver(path, solver):
    """Load a checkpoint into a solver with the same block layout; returns
    (iteration, time)."""
    blocks = solver_blocks(solver)
    ck = read(path, len(blocks[0]) - 1)
    missing = [b for b in blocks if b not in ck["conserved"]]
    if missing or len(ck["conserved"]) != len(blocks):
        raise ValueError("checkpoint: block layout differs from the solver's")
    solver.set_state(np.stack([ck["conserved"][b] for b in blocks]))
    if hasattr(solver, "time"):
        solver.time = ck["time"]
    return ck["iteration"], ck["time"]
