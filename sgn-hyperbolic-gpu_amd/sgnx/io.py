"""Output of element means in the reference's legacy VTK format
(src/io/vtk2d.h: STRUCTURED_POINTS, one point per element at its centre,
fields eta, u, v, p, H as "%.5e " rows, ey outer, ex inner)."""
from __future__ import annotations


def write_vtk(solver, path):
    """One VTK frame of the solver's resident state (element means formed on
    the device, written on the host)."""
    m = solver.elementMeans()
    Nx, Ny = solver.Nx, solver.Ny
    hx, hy = solver.Lx / Nx, solver.Ly / Ny
    with open(path, "w") as f:
        f.write("# vtk DataFile Version 3.0\nSGN2D element means\nASCII\n")
        f.write(f"DATASET STRUCTURED_POINTS\nDIMENSIONS {Nx} {Ny} 1\n")
        f.write(f"ORIGIN {0.5 * hx:.6f} {0.5 * hy:.6f} 0\nSPACING {hx:.6f} {hy:.6f} 1\n")
        f.write(f"POINT_DATA {Nx * Ny}\n")
        for k, name in enumerate(("eta", "u", "v", "p", "H")):
            f.write(f"SCALARS {name} float 1\nLOOKUP_TABLE default\n")
            for ey in range(Ny):
                f.write("".join(f"{v:.5e} " for v in m[k, ey]) + "\n")
