"""The reference's output files from device-resident fields.

Mirrors ``vtk_writer_t`` (src/module_vtk_writer.f90) for one rank:

* ``HistoryWriter`` -- ``<case>_history.csv``: the header of ``vtk_init``
  (:46-85) and one row per ``write`` with the quantities of
  ``vtk_write_diagnostics`` (:399-531), which the GPU reduces
  (``covof_history_d``); the row format is the reference's
  ``(I0,",",ES14.6, ...)``.
* ``write_step`` -- ``<case>_p0_<step>.vtr`` and ``<case>_<step>.pvtr`` of
  ``vtk_write_step`` (:90-394): ASCII RectilinearGrid with the reference's
  arrays, order and ``ES16.8`` number format; divergence and cell-centred
  velocity come from the GPU (``covof_cell_output_d``).

Fields are float64 CUDA tensors in the reference's halo layout.  Only the
text formatting runs on the host.
"""
from __future__ import annotations

import os

import numpy as np

HISTORY_HEADER = ("step,time,dt,center_of_mass_y,rise_velocity,bubble_volume,kinetic_energy,max_u,max_div,"
                  "avg_rho_gas,avg_p_gas")


def fortran_es(x: float, width: int, digits: int) -> str:
    """Fortran ``ESw.d`` edit of one value: d.ddd...E+xx right-justified in
    ``width``; a three-digit exponent drops the ``E`` (d.ddd...+xxx)."""
    s = "%.*E" % (digits, x)
    if "E" not in s:  # inf / nan
        return s.rjust(width)
    mant, ex = s.split("E")
    e = int(ex)
    sign = "+" if e >= 0 else "-"
    out = f"{mant}E{sign}{abs(e):02d}" if abs(e) <= 99 else f"{mant}{sign}{abs(e):03d}"
    return out.rjust(width) if len(out) <= width else "*" * width


def _es_array(a: np.ndarray, width: int, digits: int) -> np.ndarray:
    """fortran_es over an array (vectorised; the rare three-digit exponents
    are fixed one by one)."""
    a = np.asarray(a, dtype=np.float64)
    s = np.char.mod(f"%{width}.{digits}E", a)
    odd = (np.abs(a) >= 1e100) | ((a != 0.0) & (np.abs(a) < 1e-99)) | ~np.isfinite(a)
    for idx in zip(*np.nonzero(odd)):
        s[idx] = fortran_es(float(a[idx]), width, digits)
    return s


def _rows(a: np.ndarray) -> str:
    s = _es_array(a, 16, 8)
    return "".join("".join(v + " " for v in row) + "\n" for row in s)


class HistoryWriter:
    """``<out_dir>/<case>_history.csv`` with the reference's header and rows."""

    def __init__(self, ctx, out_dir: str, case_name: str):
        self.ctx = ctx
        os.makedirs(out_dir, exist_ok=True)
        self.path = os.path.join(out_dir, f"{case_name}_history.csv")
        with open(self.path, "w") as f:
            f.write(HISTORY_HEADER + "\n")

    def write(self, step: int, time: float, dt: float, C, u, v, p, rho, T=None):
        vals = self.ctx.history_row(C, u, v, p, rho, T)
        line = f"{int(step)}," + ",".join(fortran_es(x, 14, 6) for x in (time, dt, *vals))
        with open(self.path, "a") as f:
            f.write(line + "\n")
        return vals


def vtr_text(nx: int, ny: int, dx: float, dy: float, arrays: dict, uc: np.ndarray, vc: np.ndarray) -> str:
    """The .vtr piece of vtk_write_step from host (ny, nx) arrays: arrays
    holds C, C_s, p, rho, mu, div, T, curvature; uc, vc the cell-centred
    velocity."""
    ext = f"0 {nx} 0 {ny} 0 0"
    out = ['<?xml version="1.0"?>\n',
           '<VTKFile type="RectilinearGrid" version="0.1" byte_order="LittleEndian">\n',
           f'  <RectilinearGrid WholeExtent="{ext}">\n', f'    <Piece Extent="{ext}">\n',
           '      <CellData Scalars="PhaseIndicator,Pressure,Density,Viscosity,Divergence" Vectors="Velocity">\n']
    for name, key in VTR_ARRAYS:
        out.append(f'        <DataArray type="Float64" Name="{name}" format="ascii">\n')
        out.append(_rows(arrays[key]))
        out.append("        </DataArray>\n")
    out.append('        <DataArray type="Float64" Name="Velocity" NumberOfComponents="3" format="ascii">\n')
    su, sv = _es_array(uc, 16, 8), _es_array(vc, 16, 8)
    out.append("".join("".join(a + " " + b + " 0.0 " for a, b in zip(ru, rv)) + "\n" for ru, rv in zip(su, sv)))
    out.append("        </DataArray>\n")
    out.append("      </CellData>\n      <Coordinates>\n")
    xs = np.array([float(i) * dx for i in range(nx + 1)])
    ys = np.array([float(j) * dy for j in range(ny + 1)])
    for name, c in (("XCoordinates", xs), ("YCoordinates", ys)):
        out.append(f'        <DataArray type="Float64" Name="{name}" format="ascii">\n')
        out.append(_rows(c[None, :]))
        out.append("        </DataArray>\n")
    out.append('        <DataArray type="Float64" Name="ZCoordinates" format="ascii">\n          0.0\n'
               "        </DataArray>\n")
    out.append("      </Coordinates>\n    </Piece>\n  </RectilinearGrid>\n</VTKFile>\n")
    return "".join(out)


def pvtr_text(nx: int, ny: int, case_name: str, step: int) -> str:
    """The .pvtr index of vtk_write_step for one rank's piece."""
    stepstr = f"{int(step):06d}"
    lines = ['<?xml version="1.0"?>',
             '<VTKFile type="PRectilinearGrid" version="0.1" byte_order="LittleEndian">',
             f'  <PRectilinearGrid WholeExtent="0 {nx} 0 {ny} 0 0" GhostLevel="0">',
             '    <PCellData Scalars="PhaseIndicator,Pressure,Density,Viscosity,Divergence" Vectors="Velocity">']
    lines += [f'      <PDataArray type="Float64" Name="{name}"/>' for name, _ in VTR_ARRAYS]
    lines += ['      <PDataArray type="Float64" Name="Velocity" NumberOfComponents="3"/>', "    </PCellData>",
              "    <PCoordinates>", '      <PDataArray type="Float64" Name="XCoordinates"/>',
              '      <PDataArray type="Float64" Name="YCoordinates"/>',
              '      <PDataArray type="Float64" Name="ZCoordinates"/>', "    </PCoordinates>",
              f'    <Piece Extent="0 {nx} 0 {ny} 0 0" Source="{case_name}_p0_{stepstr}.vtr"/>',
              "  </PRectilinearGrid>", "</VTKFile>"]
    return "\n".join(lines) + "\n"


# (name in the file, key of the arrays dict), the reference's order
VTR_ARRAYS = (("PhaseIndicator", "C"), ("PhaseIndicator_Smoothed", "C_s"), ("Pressure", "p"), ("Density", "rho"),
              ("Viscosity", "mu"), ("Divergence", "div"), ("Temperature", "T"), ("Curvature", "curvature"))


def write_step(ctx, out_dir: str, case_name: str, step: int, fields: dict) -> tuple[str, str]:
    """Writes the .vtr piece and the .pvtr index of one step.  ``fields``: C,
    C_s, p, rho, mu, T, curvature, u, v (halo tensors; C_s, T and curvature
    may be absent: the reference's reset values 1, 300 and 0 are written)."""
    nx, ny = ctx.nx, ctx.ny
    os.makedirs(out_dir, exist_ok=True)
    div, uc, vc = ctx.cell_output(fields["u"], fields["v"])
    defaults = {"C": 1.0, "C_s": 1.0, "p": 0.0, "rho": 1.0, "mu": 1.0e-3, "T": 300.0, "curvature": 0.0}
    arrays = {k: (fields[k][1:ny + 1, 1:nx + 1].cpu().numpy() if fields.get(k) is not None
                  else np.full((ny, nx), fill)) for k, fill in defaults.items()}
    arrays["div"] = div.cpu().numpy()
    stepstr = f"{int(step):06d}"
    vtr = os.path.join(out_dir, f"{case_name}_p0_{stepstr}.vtr")
    with open(vtr, "w") as f:
        f.write(vtr_text(nx, ny, ctx.lx / nx, ctx.ly / ny, arrays, uc.cpu().numpy(), vc.cpu().numpy()))
    pvtr = os.path.join(out_dir, f"{case_name}_{stepstr}.pvtr")
    with open(pvtr, "w") as f:
        f.write(pvtr_text(nx, ny, case_name, step))
    return vtr, pvtr


__all__ = ["HistoryWriter", "write_step", "vtr_text", "pvtr_text", "fortran_es", "HISTORY_HEADER"]
