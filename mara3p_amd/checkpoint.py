"""Checkpoints: HDF5 files in the layout of the reference's checkpoints.

The reference writes chkpt.NNNN.h5 (mara::problem_base_t::write_checkpoint,
src/app_problem.cpp:55-78): a string "git_commit", a "run_config" group with
one dataset per parameter (int, double or string, src/app_hdf5_config.hpp),
a "schedule" group ({checkpoint, time_series} x {next_due, count},
src/app_problem.cpp:165-185), a string "module", and a "solution" group
(src/module_euler_io.cpp:81-102) with "iteration" (the rational
{num: long, den: unsigned long} compound, src/app_hdf5_rational.hpp:38-47),
"time" (double) and one dataset per block under "solution/conserved", named
by mesh::format_block_index (src/mesh_block_index.hpp:353-375), of shape
[n, n] (1D: [n]) whose elements are arrays of the block's n_c conserved
doubles in logical component order (src/app_hdf5_ndarray_dimensional.hpp:41-56).
The files are written and read by mara3p_amd.hdf5, with no HDF5 library.
"""
import math

import numpy as np

from . import hdf5

RATIONAL = hdf5.compound([("num", hdf5.integer(8, True)), ("den", hdf5.integer(8, False))])


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


def _config_value(v):
    """A run_config value as the reference stores it: int (32-bit), double or string."""
    if isinstance(v, (bool, np.bool_)):
        return np.int32(int(v))
    if isinstance(v, (int, np.integer)):
        return np.int32(v)
    if isinstance(v, (float, np.floating)):
        return np.float64(v)
    if isinstance(v, str):
        return v
    raise TypeError("checkpoint: run_config values are int, float or str (got %r)" % (v,))


def _side_effect(v):
    v = v or {}
    return {"next_due": np.float64(v.get("next_due", 0.0)), "count": np.int32(v.get("count", 0))}


def write(path, blocks, arrays, iteration, time, run_config=None, schedule=None, module="", git_commit=""):
    """Write one checkpoint.  blocks: block indices (level, i[, j]); arrays:
    the conserved state in the device layout [nblk, n_c, n(, n)]; iteration:
    an integer or a (num, den) pair; schedule: {"checkpoint": {"next_due",
    "count"}, "time_series": {...}}."""
    arrays = np.asarray(arrays, dtype=np.float64)
    if len(blocks) != arrays.shape[0]:
        raise ValueError("checkpoint: one block index per block array")
    rank = arrays.ndim - 2
    nc = arrays.shape[1]
    num, den = iteration if isinstance(iteration, tuple) else (int(iteration), 1)
    conserved = {}
    for index, a in zip(blocks, arrays):
        if len(index) != rank + 1:
            raise ValueError("checkpoint: block index rank does not match the arrays")
        conserved[format_block_index(*index)] = (np.moveaxis(a, 0, -1), hdf5.array(hdf5.float64(), nc))
    tree = {
        "git_commit": git_commit,
        "run_config": {k: _config_value(v) for k, v in (run_config or {}).items()},
        "schedule": {"checkpoint": _side_effect((schedule or {}).get("checkpoint")),
                     "time_series": _side_effect((schedule or {}).get("time_series"))},
        "module": module,
        "solution": {"iteration": (np.array((num, den), dtype=RATIONAL.dtype), RATIONAL),
                     "time": np.float64(time),
                     "conserved": conserved},
    }
    hdf5.write_file(path, tree)


def _text(a):
    return a.item().decode("ascii") if a.dtype.kind == "S" else str(a.item())


def read(path, rank):
    """Read a checkpoint: returns dict with iteration (an int, or (num, den)
    when not whole), time, module, git_commit, run_config, schedule and
    conserved = {block index: [n_c, n(, n)] array}."""
    f = hdf5.read_file(path)
    sol = f["solution"]
    it = sol["iteration"]
    num, den = int(it["num"]), int(it["den"])
    res = {"iteration": num // den if den and num % den == 0 else (num, den), "time": float(sol["time"]),
           "module": _text(f["module"]), "git_commit": _text(f["git_commit"]),
           "run_config": {k: (_text(v) if v.dtype.kind == "S" else v.item()) for k, v in f["run_config"].items()},
           "schedule": {k: {q: v.item() for q, v in g.items()} for k, g in f["schedule"].items()},
           "conserved": {}}
    for name, a in sol["conserved"].items():
        res["conserved"][read_block_index(name, rank)] = np.ascontiguousarray(np.moveaxis(a, -1, 0))
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
