"""Reader for the legacy-VTK text files of the block-LB field output (unstructured grid of voxels
with cell data), as the reference's VTK text writer lays them out
(tarch/plotter/griddata/unstructured/vtk/VTKTextFileWriter.cpp:49-83 and its vertex, cell and
cell-data writers).  Returns the arrays and the section header lines, so that tests can compare
both the format and the values."""
import numpy as np


def _numbers(lines, start, count, conv):
    vals, j = [], start
    while len(vals) < count:
        if lines[j]:
            vals += [conv(v) for v in lines[j].split()]
        j += 1
    return vals, j


def read_vtk(path):
    lines = [ln.strip() for ln in open(path).read().split("\n")]
    out = {"headers": [ln for ln in lines if ln[:1].isalpha() and not ln[0].isdigit()]}
    k = next(i for i, ln in enumerate(lines) if ln.startswith("POINTS "))
    npts = int(lines[k].split()[1])
    vals, _ = _numbers(lines, k + 1, 3 * npts, float)
    out["points"] = np.array(vals, float).reshape(npts, 3)
    k = next(i for i, ln in enumerate(lines) if ln.startswith("CELLS "))
    nc, nent = (int(v) for v in lines[k].split()[1:3])
    vals, _ = _numbers(lines, k + 1, nent, int)
    cells = np.array(vals, np.int64).reshape(nc, nent // nc)
    assert (cells[:, 0] == 8).all()
    out["cells"] = cells[:, 1:]
    k = next(i for i, ln in enumerate(lines) if ln.startswith("CELL_TYPES "))
    vals, _ = _numbers(lines, k + 1, nc, int)
    out["types"] = np.array(vals, np.int64)
    for name, rec in (("density", 1), ("velocity", 3), ("block-id", 1)):
        k = next(i for i, ln in enumerate(lines) if ln.split()[:2] in (["SCALARS", name], ["VECTORS", name]))
        start = k + (2 if lines[k].startswith("SCALARS") else 1)
        vals, _ = _numbers(lines, start, rec * nc, float)
        out[name] = np.array(vals, float).reshape(nc, rec) if rec > 1 else np.array(vals, float)
    return out


def cell_index(d, dx, offset=(0.0, 0.0, 0.0)):
    """Integer lattice coordinates (x, y, z) of every plotted cell from its voxel's lowest corner."""
    lo = d["points"][d["cells"]].min(axis=1) - np.asarray(offset)
    return np.rint(lo / dx).astype(np.int64)


def fixture_file(d, k):
    """The parsed reference file of plotted iteration k from a vtk_*.npz fixture."""
    return {"points": d[f"points_{k}"], "cells": d[f"cells_{k}"], "types": d[f"types_{k}"],
            "density": d[f"density_{k}"], "velocity": d[f"velocity_{k}"], "block-id": d[f"block_id_{k}"]}


def assert_same_vtk(mine, ref, headers, field_rtol=0.0, field_atol=0.0):
    """Same sections, voxels, cell order, block ids and (printed) fields as the reference file.
    Points are printed to six significant digits from differently rounded formulas: 1e-6 abs."""
    assert mine["headers"][1:] == [h for h in headers[2:]], (mine["headers"], headers)
    assert mine["points"].shape == ref["points"].shape
    assert np.max(np.abs(mine["points"] - ref["points"])) <= 1e-6
    assert np.array_equal(mine["cells"], ref["cells"])
    assert np.array_equal(mine["types"], ref["types"])
    assert np.array_equal(mine["block-id"], ref["block-id"])
    for key in ("density", "velocity"):
        a, b = mine[key], ref[key]
        if field_rtol == 0.0 and field_atol == 0.0:
            assert np.array_equal(a, b), (key, np.abs(a - b).max())
        else:
            assert np.all(np.abs(a - b) <= field_rtol * np.abs(b) + field_atol), (key, np.abs(a - b).max())
