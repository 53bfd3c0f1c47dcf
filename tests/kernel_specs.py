"""Argument lists of the hydro kernels, shared by the oracle-vs-reference
tests (CPU) and the HIP-vs-oracle tests (GPU).

Each kernel is described by its arguments in reference order.  Kinds:
  i                  integer scalar by value (bounds, flags)
  d                  double scalar by value
  C V X Y Z          3-D arrays: cell, vertex, x-face, y-face, z-face
  c1x c1y c1z        1-D cell-centred coordinate arrays (n+4)
  v1x v1y v1z        1-D vertex arrays (n+5)
  od oi              output scalar (double*, int*)
  f19 n6             int[19] field flags, int[6] neighbours
The `fill` tag picks the random distribution so every division stays finite.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Tuple

import numpy as np

DIMS = ["x_min", "x_max", "y_min", "y_max", "z_min", "z_max"]


def _d(names):
    return [(n, "i") for n in names]


SPECS: Dict[str, List[Tuple[str, str, str]]] = {}


def spec(name, args):
    SPECS[name] = [(a[0], a[1], a[2] if len(a) > 2 else "") for a in args]


D = [(n, "i") for n in DIMS]

spec("ideal_gas_kernel", D + [("density", "C", "rho"), ("energy", "C", "e"), ("pressure", "C", "any"),
                               ("soundspeed", "C", "any")])
spec("viscosity_kernel", D + [("xarea", "X", "area"), ("yarea", "Y", "area"), ("zarea", "Z", "area"),
                               ("celldx", "c1x", "dx"), ("celldy", "c1y", "dx"), ("celldz", "c1z", "dx"),
                               ("density0", "C", "rho"), ("pressure", "C", "p"), ("viscosity", "C", "any"),
                               ("xvel0", "V", "vel"), ("yvel0", "V", "vel"), ("zvel0", "V", "vel")])
spec("calc_dt_kernel", D + [("g_small", "d"), ("g_big", "d"), ("dtmin", "d"), ("dtc_safe", "d"),
                             ("dtu_safe", "d"), ("dtv_safe", "d"), ("dtw_safe", "d"), ("dtdiv_safe", "d"),
                             ("xarea", "X", "area"), ("yarea", "Y", "area"), ("zarea", "Z", "area"),
                             ("cellx", "c1x", "pos"), ("celly", "c1y", "pos"), ("cellz", "c1z", "pos"),
                             ("celldx", "c1x", "dx"), ("celldy", "c1y", "dx"), ("celldz", "c1z", "dx"),
                             ("volume", "C", "vol"), ("density0", "C", "rho"), ("energy0", "C", "e"),
                             ("pressure", "C", "p"), ("viscosity_a", "C", "q"), ("soundspeed", "C", "c"),
                             ("xvel0", "V", "vel"), ("yvel0", "V", "vel"), ("zvel0", "V", "vel"),
                             ("dt_min", "V", "any"), ("dt_min_val", "od"), ("dtl_control", "oi"),
                             ("xl_pos", "od"), ("yl_pos", "od"), ("zl_pos", "od"), ("jldt", "oi"),
                             ("kldt", "oi"), ("lldt", "oi"), ("small", "oi")])
spec("PdV_kernel", [("predict", "i")] + D + [("dt", "d"), ("xarea", "X", "area"), ("yarea", "Y", "area"),
                                              ("zarea", "Z", "area"), ("volume", "C", "vol"),
                                              ("density0", "C", "rho"), ("density1", "C", "any"),
                                              ("energy0", "C", "e"), ("energy1", "C", "any"),
                                              ("pressure", "C", "p"), ("viscosity", "C", "q"),
                                              ("xvel0", "V", "vel"), ("xvel1", "V", "vel"),
                                              ("yvel0", "V", "vel"), ("yvel1", "V", "vel"),
                                              ("zvel0", "V", "vel"), ("zvel1", "V", "vel")])
spec("revert_kernel", D + [("density0", "C", "rho"), ("density1", "C", "any"), ("energy0", "C", "e"),
                           ("energy1", "C", "any")])
spec("accelerate_kernel", D + [("dt", "d"), ("xarea", "X", "area"), ("yarea", "Y", "area"),
                               ("zarea", "Z", "area"), ("volume", "C", "vol"), ("density0", "C", "rho"),
                               ("pressure", "C", "p"), ("viscosity", "C", "q"), ("xvel0", "V", "vel"),
                               ("yvel0", "V", "vel"), ("zvel0", "V", "vel"), ("xvel1", "V", "any"),
                               ("yvel1", "V", "any"), ("zvel1", "V", "any")])
spec("flux_calc_kernel", D + [("dt", "d"), ("xarea", "X", "area"), ("yarea", "Y", "area"),
                              ("zarea", "Z", "area"), ("xvel0", "V", "vel"), ("yvel0", "V", "vel"),
                              ("zvel0", "V", "vel"), ("xvel1", "V", "vel"), ("yvel1", "V", "vel"),
                              ("zvel1", "V", "vel"), ("vol_flux_x", "X", "any"), ("vol_flux_y", "Y", "any"),
                              ("vol_flux_z", "Z", "any")])
spec("advec_cell_kernel", D + [("advect_x", "i"), ("dir", "i"), ("sweep_number", "i"),
                               ("vertexdx", "v1x", "dx"), ("vertexdy", "v1y", "dx"), ("vertexdz", "v1z", "dx"),
                               ("volume", "C", "vol"), ("density1", "C", "rho"), ("energy1", "C", "e"),
                               ("mass_flux_x", "X", "any"), ("vol_flux_x", "X", "vflux"),
                               ("mass_flux_y", "Y", "any"), ("vol_flux_y", "Y", "vflux"),
                               ("mass_flux_z", "Z", "any"), ("vol_flux_z", "Z", "vflux"),
                               ("pre_vol", "V", "any"), ("post_vol", "V", "any"), ("pre_mass", "V", "any"),
                               ("post_mass", "V", "any"), ("advec_vol", "V", "any"), ("post_ener", "V", "any"),
                               ("ener_flux", "V", "any")])
spec("advec_mom_kernel", D + [("vel1", "V", "vel"), ("mass_flux_x", "X", "mflux"), ("vol_flux_x", "X", "vflux"),
                              ("mass_flux_y", "Y", "mflux"), ("vol_flux_y", "Y", "vflux"),
                              ("mass_flux_z", "Z", "mflux"), ("vol_flux_z", "Z", "vflux"),
                              ("volume", "C", "vol"), ("density1", "C", "rho"),
                              ("node_flux", "V", "any"), ("node_mass_post", "V", "any"),
                              ("node_mass_pre", "V", "any"), ("advec_vel", "V", "any"), ("mom_flux", "V", "any"),
                              ("pre_vol", "V", "any"), ("post_vol", "V", "any"),
                              ("celldx", "c1x", "dx"), ("celldy", "c1y", "dx"), ("celldz", "c1z", "dx"),
                              ("advect_x", "i"), ("which_vel", "i"), ("sweep_number", "i"), ("direction", "i")])
spec("reset_field_kernel", D + [("density0", "C", "any"), ("density1", "C", "rho"), ("energy0", "C", "any"),
                                ("energy1", "C", "e"), ("xvel0", "V", "any"), ("xvel1", "V", "vel"),
                                ("yvel0", "V", "any"), ("yvel1", "V", "vel"), ("zvel0", "V", "any"),
                                ("zvel1", "V", "vel")])
HALO_FIELDS = [("density0", "C"), ("energy0", "C"), ("pressure", "C"), ("viscosity", "C"), ("soundspeed", "C"),
               ("density1", "C"), ("energy1", "C"), ("xvel0", "V"), ("yvel0", "V"), ("zvel0", "V"),
               ("xvel1", "V"), ("yvel1", "V"), ("zvel1", "V"), ("vol_flux_x", "X"), ("vol_flux_y", "Y"),
               ("vol_flux_z", "Z"), ("mass_flux_x", "X"), ("mass_flux_y", "Y"), ("mass_flux_z", "Z")]
spec("update_halo_kernel", D + [("chunk_neighbours", "n6"), ("tile_neighbours", "n6")] +
     [(n, k, "any") for n, k in HALO_FIELDS] + [("fields", "f19"), ("depth", "i")])
spec("field_summary_kernel", D + [("volume", "C", "vol"), ("density0", "C", "rho"), ("energy0", "C", "e"),
                                  ("pressure", "C", "p"), ("xvel0", "V", "vel"), ("yvel0", "V", "vel"),
                                  ("zvel0", "V", "vel"), ("vol", "od"), ("mass", "od"), ("ie", "od"),
                                  ("ke", "od"), ("press", "od")])
for _face in ("left", "right", "bottom", "top", "back", "front"):
    spec(f"update_tile_halo_{_face}_kernel",
         D + [(n, k, "any") for n, k in HALO_FIELDS] +
         [("n_" + d, "i") for d in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")] +
         [("n_" + n, k + "n", "any") for n, k in HALO_FIELDS] + [("fields", "f19"), ("depth", "i")])

FACE_ID = {"left": 1, "right": 2, "bottom": 3, "top": 4, "back": 5, "front": 6}


def extents(kind: str, nx: int, ny: int, nz: int) -> Tuple[int, ...]:
    """Fortran extents (x, y, z) of an array class (build_field.f90:41-147)."""
    k = kind.rstrip("n")
    return {
        "C": (nx + 4, ny + 4, nz + 4),
        "V": (nx + 5, ny + 5, nz + 5),
        "X": (nx + 5, ny + 4, nz + 4),
        "Y": (nx + 4, ny + 5, nz + 4),
        "Z": (nx + 4, ny + 4, nz + 5),
        "c1x": (nx + 4,), "c1y": (ny + 4,), "c1z": (nz + 4,),
        "v1x": (nx + 5,), "v1y": (ny + 5,), "v1z": (nz + 5,),
    }[k]


def fill(rng: np.random.Generator, tag: str, shape) -> np.ndarray:
    u = lambda a, b: rng.uniform(a, b, size=shape)
    if tag == "rho":
        return u(0.1, 2.0)
    if tag == "e":
        return u(0.5, 3.0)
    if tag == "p":
        return u(0.05, 2.0)
    if tag == "q":
        return np.where(rng.uniform(size=shape) < 0.5, 0.0, u(0.0, 0.1))
    if tag == "c":
        return u(0.3, 3.0)
    if tag == "vol":
        return u(0.8, 1.2)
    if tag == "area":
        return u(0.8, 1.2)
    if tag == "dx":
        return u(0.8, 1.2)
    if tag == "pos":
        return u(0.0, 10.0)
    if tag == "vel":
        v = u(-1.0, 1.0)
        v[rng.uniform(size=shape) < 0.1] = 0.0  # exact zeros exercise the sign branches
        return v
    if tag == "vflux":
        v = u(-0.05, 0.05)
        v[rng.uniform(size=shape) < 0.1] = 0.0
        return v
    if tag == "mflux":
        return u(-0.05, 0.05)
    return u(-1.0, 1.0)


def make_args(name: str, dims: Tuple[int, int, int], rng: np.random.Generator, ints: Dict[str, int],
              doubles: Dict[str, float], ndims: Tuple[int, int, int] = None,
              fields: np.ndarray = None, neighbours: Dict[str, np.ndarray] = None):
    """Build a dict of numpy arrays / python scalars for one call."""
    nx, ny, nz = dims
    vals = {}
    for aname, kind, tag in SPECS[name]:
        if kind == "i":
            if aname in ("x_min", "y_min", "z_min"):
                vals[aname] = 1
            elif aname == "x_max":
                vals[aname] = nx
            elif aname == "y_max":
                vals[aname] = ny
            elif aname == "z_max":
                vals[aname] = nz
            elif aname in ("n_xmin", "n_ymin", "n_zmin"):
                vals[aname] = 1
            elif aname == "n_xmax":
                vals[aname] = ndims[0]
            elif aname == "n_ymax":
                vals[aname] = ndims[1]
            elif aname == "n_zmax":
                vals[aname] = ndims[2]
            else:
                vals[aname] = ints[aname]
        elif kind == "d":
            vals[aname] = doubles[aname]
        elif kind in ("od",):
            vals[aname] = np.zeros(1, dtype=np.float64)
        elif kind == "oi":
            vals[aname] = np.zeros(1, dtype=np.int32)
        elif kind == "f19":
            vals[aname] = np.ascontiguousarray(fields, dtype=np.int32)
        elif kind == "n6":
            vals[aname] = np.ascontiguousarray(neighbours[aname], dtype=np.int32)
        else:
            d3 = ndims if kind.endswith("n") else dims
            shp = extents(kind, *d3)
            vals[aname] = np.asfortranarray(fill(rng, tag, shp))
    return vals


def call(fn, name: str, vals: dict, by_value_scalars: bool = True):
    """Call a C-ABI kernel `fn` with the arrays/scalars in `vals`."""
    args = []
    for aname, kind, _ in SPECS[name]:
        v = vals[aname]
        if kind == "i":
            args.append(ctypes.c_int(int(v)))
        elif kind == "d":
            args.append(ctypes.c_double(float(v)))
        else:
            args.append(ctypes.c_void_p(v.ctypes.data))
    return fn(*args)


def array_names(name: str) -> List[str]:
    return [a for a, k, _ in SPECS[name] if k not in ("i", "d")]


def copy_vals(vals: dict) -> dict:
    return {k: (v.copy(order="K") if isinstance(v, np.ndarray) else v) for k, v in vals.items()}
