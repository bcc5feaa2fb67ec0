"""Offline post-processing of saved states (phasefield5d/solver/post_process.py).

Host-only consumers of the ``data/step_*.npz`` files that ``solver.io`` writes:
iteration over a run's states, parsing of run-directory names, and export of a
3D composition field to legacy ASCII VTK (structured points) or to VTK XML
ImageData with base64 binary arrays, for ParaView.  The dependent component
(Fe) is written first as 1 - sum(X).  Points are written with x varying
fastest, the VTK ordering, so the (Nx, Ny, Nz) array is traversed z-major.
"""
import base64
import os
import re
import struct
from pathlib import Path

import numpy as np

DEFAULT_NAMES = ("Fe", "Mn", "Ni", "Co", "Cu")


def iter_npz_states(data_path):
    """post_process.py:14-22: (timestep, time, dt, composition) per ``.npz`` in name order."""
    names = sorted(n for n in os.listdir(data_path) if n.endswith(".npz"))
    for n in names:
        with np.load(os.path.join(data_path, n), allow_pickle=False) as z:
            yield int(z["timestep"]), float(z["time"]), float(z["dt"]), z["current_composition"]


_NUM = r"[0-9eE+\-.]+"
_RUN_NAME = re.compile(rf"cells(?P<cells>\d+)_dx(?P<dx>{_NUM})_fluc(?P<fluc>{_NUM})_kappa(?P<kappa>{_NUM})"
                       rf"_steps(?P<steps>\d+)_dt(?P<dt>{_NUM})s_(?P<date>\d{{8}})_(?P<time>\d{{6}})")


def parse_run_name(name):
    """post_process.py:29-57: the fields of a run directory name, or None if it is not one.

    Only a single-number cell count matches (as in the reference), so the names of
    2D / 3D runs (``cells128x128...``) give None."""
    m = _RUN_NAME.match(name)
    if m is None:
        return None
    g = m.groupdict()
    out = {"name": name}
    for k in ("cells", "steps"):
        out[k] = int(g[k])
    for k in ("dx", "fluc", "kappa", "dt"):
        out[k] = float(g[k])
    out["date"], out["time"] = g["date"], g["time"]
    return out


def _with_dependent(X):
    """(..., n) -> (..., n + 1) with the dependent component 1 - sum first."""
    X = np.asarray(X, dtype=np.float64)
    return np.concatenate([1.0 - X.sum(axis=-1, keepdims=True), X], axis=-1)


def _vtk_order(field3d):
    """Scalar field (Nx, Ny, Nz) as float32 with x fastest (VTK point order)."""
    return np.ascontiguousarray(np.transpose(np.asarray(field3d), (2, 1, 0)), dtype=np.float32).ravel()


def npz_to_vtk(npz_path, vtk_path=None, spacing=(0.01, 0.01, 0.01), origin=(0.0, 0.0, 0.0),
               component_names=DEFAULT_NAMES):
    """post_process.py:63-101: legacy ASCII VTK STRUCTURED_POINTS, nine values per line."""
    npz_path = Path(npz_path)
    with np.load(npz_path, allow_pickle=False) as z:
        keys = list(z.files)
        key = keys[0] if len(keys) == 1 else ("composition" if "composition" in keys else keys[0])
        X = z[key]
    if X.ndim != 4 or X.shape[-1] != 4:
        raise ValueError(f"Expected shape (Nx,Ny,Nz,4), got {X.shape}")
    nx, ny, nz, _ = X.shape
    full = _with_dependent(X)
    out = Path(vtk_path) if vtk_path is not None else npz_path.with_suffix(".vtk")
    npts = nx * ny * nz
    lines = ["# vtk DataFile Version 3.0", f"CH 5-component data from {npz_path.name}", "ASCII",
             "DATASET STRUCTURED_POINTS", f"DIMENSIONS {nx} {ny} {nz}",
             f"ORIGIN {origin[0]} {origin[1]} {origin[2]}", f"SPACING {spacing[0]} {spacing[1]} {spacing[2]}",
             f"POINT_DATA {npts}"]
    with out.open("w") as f:
        f.write("\n".join(lines) + "\n")
        for c, name in enumerate(component_names):
            f.write(f"SCALARS {name} float 1\nLOOKUP_TABLE default\n")
            v = _vtk_order(full[..., c])
            for s in range(0, npts, 9):
                f.write(" ".join("%.7e" % x for x in v[s:s + 9]) + "\n")
    return str(out)


def _b64_block(values):
    """VTK XML binary payload: a little-endian uint32 byte count, then the bytes, base64."""
    raw = values.astype("<f4").tobytes()
    return base64.b64encode(struct.pack("<I", len(raw)) + raw).decode("ascii")


def npz_to_vti(npz_path, vti_path=None, spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0),
               component_names=DEFAULT_NAMES):
    """post_process.py:104-155: VTK XML ImageData, one Float32 point array per component."""
    npz_path = Path(npz_path)
    with np.load(npz_path, allow_pickle=False) as z:
        keys = list(z.files)
        key = keys[0] if len(keys) == 1 else ("current_composition" if "current_composition" in keys else keys[0])
        X = z[key]
    if X.ndim != 4:
        raise ValueError(f"Expected 4D array (Nx,Ny,Nz,n_comp), got {X.shape}")
    nx, ny, nz, _ = X.shape
    full = _with_dependent(X)
    out = Path(vti_path) if vti_path is not None else npz_path.with_suffix(".vti")
    ext = f"0 {nx - 1} 0 {ny - 1} 0 {nz - 1}"
    parts = ['<?xml version="1.0"?>',
             '<VTKFile type="ImageData" version="0.1" byte_order="LittleEndian">',
             f'  <ImageData WholeExtent="{ext}" Origin="{origin[0]} {origin[1]} {origin[2]}" '
             f'Spacing="{spacing[0]} {spacing[1]} {spacing[2]}">',
             f'    <Piece Extent="{ext}">', "      <PointData>"]
    for c in range(full.shape[-1]):
        parts.append(f'        <DataArray type="Float32" Name="{component_names[c]}" '
                     f'NumberOfComponents="1" format="binary">')
        parts.append("          " + _b64_block(_vtk_order(full[..., c])))
        parts.append("        </DataArray>")
    parts += ["      </PointData>", "      <CellData/>", "    </Piece>", "  </ImageData>", "</VTKFile>"]
    with out.open("w") as f:
        f.write("\n".join(parts) + "\n")
    return str(out)


def batch_npz_to_vti(input_dir, output_dir=None, spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0),
                     component_names=DEFAULT_NAMES, verbose=True):
    """post_process.py:158-193: every ``.npz`` of a directory to ``.vti``, in name order."""
    src = Path(input_dir)
    dst = Path(output_dir) if output_dir is not None else src
    dst.mkdir(parents=True, exist_ok=True)
    files = sorted(src.glob("*.npz"))
    if verbose:
        print(f"Found {len(files)} .npz files in {src}")
    done = []
    for fp in files:
        target = dst / (fp.stem + ".vti")
        if verbose:
            print(f"Converting {fp.name} -> {target.name}")
        done.append(npz_to_vti(fp, target, spacing=spacing, origin=origin, component_names=component_names))
    if verbose:
        print(f"Finished writing {len(done)} .vti files.")
    return done
