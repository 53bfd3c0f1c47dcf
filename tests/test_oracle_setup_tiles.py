"""Pins the oracle's set-up kernels, tile decomposition and tiled halo driver.

* initialise_chunk_kernel and generate_chunk_kernel: the C restatement
  against the reference's own kernels (compiled from their sources into
  oracle/_ref by oracle/Makefile), bit for bit, at odd shapes and tile
  offsets, for cuboid, sphere and point states with velocities, state edges
  on vertex coordinates and states painted over one another.
* clover_decompose_tile (clover.f90:258-447): the split the oracle (and the
  product's host code) picks equals a restatement of clover.f90's metric in
  IEEE single precision -- the default REAL of `1.0*x_cells` -- on random
  decks, including ones where the single-precision rounding breaks a tie
  that double precision breaks the other way.
* update_halo over tiles (update_halo.f90:27-88, the passes of
  update_tile_halo.f90:47-423, then each tile's reflective pass): after the
  tiled exchange every tile's halo band holds exactly what the untiled
  exchange leaves at the same global cells.  The tile halos that lie inside
  the chunk start poisoned, so a missed or misordered copy shows.  The only
  differences are the reference's two documented quirks, which the tiled
  exchange must reproduce: the right-face ZVEL0 branch copies yvel0
  (update_tile_halo_kernel.f90:579-589), and energy1 is copied on the
  left/right/bottom/top faces only when density1 is flagged.
* Whole runs (the reference's kernels through the oracle driver): a split
  along z leaves every field and dt bit-identical to the untiled run; a
  split along y departs after a few steps at one corner that the face-data
  mirror does not reach; splits along x depart too.  These are answers of
  the reference with tiles, which the oracle's tiled driver reproduces.
"""
from __future__ import annotations

import ctypes
import zlib

import numpy as np
import pytest

import oracle_lib as ol
from cloverleaf3d_ref_amd.deck import parse_deck

needs_ref = pytest.mark.skipif(not ol.ref_available(), reason="oracle/_ref not built (needs the reference tree)")

P = ctypes.c_void_p


def _ptr(a):
    return P(a.ctypes.data)


def _bitwise(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


# ----------------------------------------------------------------- set-up kernels

def _mesh(fn, nx, ny, nz, xmin, ymin, zmin, dx, dy, dz, fill=np.nan):
    """Call an initialise_chunk_kernel; returns its 16 output arrays."""
    v = lambda n: np.full(n + 5, fill)
    c = lambda n: np.full(n + 4, fill)
    out = dict(vertexx=v(nx), vertexdx=v(nx), vertexy=v(ny), vertexdy=v(ny), vertexz=v(nz), vertexdz=v(nz),
               cellx=c(nx), celldx=c(nx), celly=c(ny), celldy=c(ny), cellz=c(nz), celldz=c(nz),
               volume=np.full((nx + 4, ny + 4, nz + 4), fill, order="F"),
               xarea=np.full((nx + 5, ny + 4, nz + 4), fill, order="F"),
               yarea=np.full((nx + 4, ny + 5, nz + 4), fill, order="F"),
               zarea=np.full((nx + 4, ny + 4, nz + 5), fill, order="F"))
    fn.restype = None
    fn(1, nx, 1, ny, 1, nz, *[ctypes.c_double(x) for x in (xmin, ymin, zmin, dx, dy, dz)],
       *[_ptr(out[k]) for k in ("vertexx", "vertexdx", "vertexy", "vertexdy", "vertexz", "vertexdz", "cellx",
                                "celldx", "celly", "celldy", "cellz", "celldz", "volume", "xarea", "yarea",
                                "zarea")])
    return out


def _tile_origin(deck_lo, d, first_cell):
    # initialise_chunk.f90:37-44: xmin = deck xmin + dx*float(left-1), default REAL conversion
    return deck_lo + d * float(np.float32(first_cell - 1))


MESH_CASES = [
    # (nx, ny, nz, first cells (left, bottom, back), extent, cells)
    ((7, 5, 6), (1, 1, 1), (10.0, 10.0, 10.0), (7, 5, 6)),
    ((4, 9, 3), (5, 1, 9), (10.0, 7.0, 3.5), (11, 9, 13)),
    ((12, 3, 5), (13, 4, 1), (1.0, 0.3, 100.0), (24, 9, 5)),
    ((6, 6, 6), (33, 17, 65), (10.0, 10.0, 10.0), (96, 40, 72)),
]


def _mesh_args(case):
    (nx, ny, nz), (lf, bt, bk), ext, cells = case
    d = [(ext[i] - 0.0) / float(np.float32(cells[i])) for i in range(3)]  # start.f90:60-62
    lo = [_tile_origin(0.0, d[i], (lf, bt, bk)[i]) for i in range(3)]
    return (nx, ny, nz, *lo, *d)


@needs_ref
@pytest.mark.parametrize("case", MESH_CASES)
def test_initialise_chunk_vs_ref(case):
    args = _mesh_args(case)
    a = _mesh(ol.lib().oracle_initialise_chunk_kernel, *args)
    b = _mesh(ol.ref_lib().ref_initialise_chunk_kernel, *args)
    for k in a:
        assert _bitwise(a[k], b[k]), k
    # the area arrays stop at +2 along their normal: the +3 plane is untouched
    assert np.isnan(a["xarea"][-1]).all() and np.isnan(a["zarea"][:, :, -1]).all()


def _states(rng, mesh, kind, n_states):
    """State arrays (background first) for generate_chunk_kernel."""
    vx, vy, vz = mesh["vertexx"], mesh["vertexy"], mesh["vertexz"]
    cx, cy, cz = mesh["cellx"], mesh["celly"], mesh["cellz"]
    S = {k: np.zeros(n_states) for k in ("density", "energy", "xvel", "yvel", "zvel", "xmin", "xmax", "ymin",
                                           "ymax", "zmin", "zmax", "radius")}
    geo = np.zeros(n_states, np.int32)
    S["density"][:] = rng.uniform(0.1, 2.0, n_states)
    S["energy"][:] = rng.uniform(0.5, 3.0, n_states)
    for v in ("xvel", "yvel", "zvel"):
        S[v][:] = rng.uniform(-1.0, 1.0, n_states)
    for s in range(1, n_states):
        g = kind if kind != "mixed" else (1, 2, 3)[s % 3]
        geo[s] = g
        if g == 1:
            # edges on vertex coordinates exercise .GE. and .LT.; others nudged
            for lo, hi, verts in (("xmin", "xmax", vx), ("ymin", "ymax", vy), ("zmin", "zmax", vz)):
                i, j = sorted(rng.choice(len(verts), 2, replace=False))
                a, b = verts[i], verts[j]
                if rng.uniform() < 0.5:
                    a += (b - a) * 0.01
                S[lo][s], S[hi][s] = a, b
        elif g == 2:
            # centre on a cell centre, radius through other cell centres
            S["xmin"][s], S["ymin"][s], S["zmin"][s] = (rng.choice(cx[1:-1]), rng.choice(cy[1:-1]),
                                                        rng.choice(cz[1:-1]))
            S["radius"][s] = abs(rng.choice(cx) - S["xmin"][s]) + rng.choice([0.0, 1e-3])
        else:
            # a point on a vertex (hit) or between vertices (miss)
            S["xmin"][s], S["ymin"][s], S["zmin"][s] = rng.choice(vx), rng.choice(vy), rng.choice(vz)
            if rng.uniform() < 0.25:
                S["xmin"][s] += 1e-9
    return S, geo


def _generate(fn, dims, mesh, S, geo, init):
    nx, ny, nz = dims
    out = {k: v.copy(order="F") for k, v in init.items()}
    fn.restype = None
    fn(1, nx, 1, ny, 1, nz, *[_ptr(mesh[k]) for k in ("vertexx", "vertexy", "vertexz", "cellx", "celly", "cellz")],
       *[_ptr(out[k]) for k in ("density0", "energy0", "xvel0", "yvel0", "zvel0")], ctypes.c_int(len(geo)),
       *[_ptr(S[k]) for k in ("density", "energy", "xvel", "yvel", "zvel", "xmin", "xmax", "ymin", "ymax", "zmin",
                              "zmax", "radius")], _ptr(geo))
    return out


@needs_ref
@pytest.mark.parametrize("case", MESH_CASES)
@pytest.mark.parametrize("kind", [1, 2, 3, "mixed"])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_generate_chunk_vs_ref(case, kind, seed):
    args = _mesh_args(case)
    nx, ny, nz = args[:3]
    mesh = _mesh(ol.ref_lib().ref_initialise_chunk_kernel, *args)
    rng = np.random.default_rng(1000 * seed + zlib.crc32(str(kind).encode()) % 1000)
    S, geo = _states(rng, mesh, kind, 4)
    init = {k: np.asfortranarray(rng.uniform(-5, 5, (nx + 4, ny + 4, nz + 4))) for k in ("density0", "energy0")}
    init.update({k: np.asfortranarray(rng.uniform(-5, 5, (nx + 5, ny + 5, nz + 5)))
                 for k in ("xvel0", "yvel0", "zvel0")})
    a = _generate(ol.lib().oracle_generate_chunk_kernel, (nx, ny, nz), mesh, S, geo, init)
    b = _generate(ol.ref_lib().ref_generate_chunk_kernel, (nx, ny, nz), mesh, S, geo, init)
    for k in a:
        assert _bitwise(a[k], b[k]), (k, kind, seed)
    # the case paints something other than the background (not a vacuous pass)
    if kind in (1, 2, "mixed"):
        assert not np.all(a["density0"] == S["density"][0])


@needs_ref
def test_whole_run_with_reference_setup_kernels():
    """Sphere and point states through start + 12 steps: the oracle's C
    kernels against the reference's kernels (set-up ones included)."""
    deck = parse_deck("""*clover
 state 1 density=0.2 energy=1.0
 state 2 density=1.0 energy=2.5 xvel=0.1 geometry=sphere xmin=3.0 ymin=2.5 zmin=2.0 radius=1.7
 state 3 density=2.0 energy=3.0 zvel=-0.2 geometry=point xmin=5.0 ymin=5.0 zmin=5.0
 state 4 density=0.5 energy=4.0 yvel=0.3 geometry=cuboid xmin=6 xmax=9 ymin=1 ymax=4 zmin=2 zmax=8
 x_cells=13 y_cells=11 z_cells=9
 xmin=0 ymin=0 zmin=0 xmax=10 ymax=10 zmax=10
 initial_timestep=0.04 timestep_rise=1.5 end_step=12 end_time=100 summary_frequency=3
*endclover""")
    a, da = ol.run(deck)
    b, db = ol.run(deck, use_ref=True)
    assert np.array_equal(da.view(np.uint64), db.view(np.uint64))
    for ra, rb in zip(a, b):
        for k in ("vol", "mass", "ie", "ke", "press"):
            assert ra[k] == pytest.approx(rb[k], rel=1e-13, abs=1e-300), (ra["step"], k)
    assert a[0]["ke"] > 0  # the velocities were painted


# ----------------------------------------------------------------- decomposition

def _metric_f32(X, Y, Z, cx, cy, cz):
    f = np.float32
    fx, fy, fz = f(f(1.0) * f(X)) / f(cx), f(f(1.0) * f(Y)) / f(cy), f(f(1.0) * f(Z)) / f(cz)
    surface = f(f(f(fx * fy) * f(2)) + f(f(fx * fz) * f(2))) + f(f(fy * fz) * f(2))
    volume = f(f(fx * fy) * fz)
    return float(surface) / float(volume)


def _split_f32(X, Y, Z, T, dims):
    """clover.f90:276-343 with default-REAL surface and volume."""
    best = _metric_f32(X, Y, Z, 1, 1, T)
    chunk = (1, 1, T)
    x_bound, y_bound = {3: (T, T), 2: (1, T), 1: (1, 1)}[dims]
    for t in range(1, x_bound + 1):
        if T % t:
            continue
        div1 = min(y_bound, T // t)
        for j in range(1, div1 + 1):
            if div1 % j or T % (t * j):
                continue
            m = _metric_f32(X, Y, Z, t, j, T // (t * j))
            if m < best:
                best, chunk = m, (t, j, T // (t * j))
    return chunk


def _oracle_split(X, Y, Z, T, dims):
    n = [np.zeros(T, np.int32) for _ in range(6)]
    nb = np.zeros(6 * T, np.int32)
    nt = ol.lib().oracle_decompose_tiles(X, Y, Z, T, dims, *[_ptr(a) for a in n], _ptr(nb))
    assert nt == T
    tiles = [dict(n=(int(n[0][i]), int(n[1][i]), int(n[2][i])), off=(int(n[3][i]), int(n[4][i]), int(n[5][i])),
                  nbr=nb[6 * i:6 * i + 6].tolist()) for i in range(T)]
    cx = len({t["off"][0] for t in tiles})
    cy = len({t["off"][1] for t in tiles})
    cz = len({t["off"][2] for t in tiles})
    return (cx, cy, cz), tiles


def test_decomposition_single_precision_tie():
    # 1x3x1 and 1x1x3 tie exactly; single precision picks 1x3x1 (double: 1x1x3)
    assert _split_f32(109, 29, 29, 3, 2) == (1, 3, 1)
    assert _oracle_split(109, 29, 29, 3, 2)[0] == (1, 3, 1)


def test_decomposition_matches_single_precision_metric():
    rng = np.random.default_rng(5)
    flips = 0
    for _ in range(3000):
        X, Y, Z = (int(v) for v in rng.integers(32, 200, 3))  # at least one cell per tile
        T = int(rng.choice([2, 3, 4, 5, 6, 8, 9, 12, 16, 18, 24, 27, 32]))
        dims = int(rng.choice([1, 2, 3]))
        want = _split_f32(X, Y, Z, T, dims)
        got, tiles = _oracle_split(X, Y, Z, T, dims)
        assert got == want, (X, Y, Z, T, dims)
        # cells add up and every neighbour pair is mutual
        assert sum(t["n"][0] * t["n"][1] * t["n"][2] for t in tiles) == X * Y * Z
        for i, t in enumerate(tiles):
            for f, o in ((0, 1), (1, 0), (2, 3), (3, 2), (4, 5), (5, 4)):
                if t["nbr"][f] != -1:
                    assert tiles[t["nbr"][f] - 1]["nbr"][o] == i + 1
        flips += want != _split_f64(X, Y, Z, T, dims)
    assert flips > 0  # the sample holds decks where the precision decides


def test_product_decomposition_matches_oracle():
    """The product's host code (clover_decompose_tiles, the split
    clover_chunk_create uses) gives the same tiles as the oracle on random
    decks, the single-precision tie deck included.  Host code only: no GPU."""
    from cloverleaf3d_ref_amd import _lib
    L = _lib.lib()
    fn = L.clover_decompose_tiles
    fn.restype = ctypes.c_int
    rng = np.random.default_rng(9)
    decks = [(109, 29, 29, 3, 2)] + [(int(rng.integers(32, 300)), int(rng.integers(32, 300)),
                                      int(rng.integers(32, 300)), int(rng.choice([2, 3, 4, 6, 8, 12, 16, 27, 32])),
                                      int(rng.choice([1, 2, 3]))) for _ in range(1500)]
    for X, Y, Z, T, dims in decks:
        d3, o3, nb = np.zeros(3 * T, np.int32), np.zeros(3 * T, np.int32), np.zeros(6 * T, np.int32)
        assert fn(X, Y, Z, T, dims, T, _ptr(d3), _ptr(o3), _ptr(nb)) == T
        _, tiles = _oracle_split(X, Y, Z, T, dims)
        for i, t in enumerate(tiles):
            assert tuple(d3[3 * i:3 * i + 3]) == t["n"], (X, Y, Z, T, dims)
            assert tuple(o3[3 * i:3 * i + 3]) == t["off"]
            assert nb[6 * i:6 * i + 6].tolist() == t["nbr"]
    assert fn(10, 10, 10, 4, 3, 2, None, None, None) < 0  # more tiles than the caller's room


def _split_f64(X, Y, Z, T, dims):
    def metric(cx, cy, cz):
        fx, fy, fz = X / cx, Y / cy, Z / cz
        return (fx * fy * 2 + fx * fz * 2 + fy * fz * 2) / (fx * fy * fz)
    best, chunk = metric(1, 1, T), (1, 1, T)
    x_bound, y_bound = {3: (T, T), 2: (1, T), 1: (1, 1)}[dims]
    for t in range(1, x_bound + 1):
        if T % t:
            continue
        div1 = min(y_bound, T // t)
        for j in range(1, div1 + 1):
            if div1 % j or T % (t * j):
                continue
            m = metric(t, j, T // (t * j))
            if m < best:
                best, chunk = m, (t, j, T // (t * j))
    return chunk


# ----------------------------------------------------------------- tiled update_halo

# update_tile_halo argument order, with each field's FIELD id (data.f90:55-74)
# and its extent class per axis (0: cell-like n+4, 1: node-like n+5)
HALO = [("density0", 1, (0, 0, 0)), ("energy0", 3, (0, 0, 0)), ("pressure", 5, (0, 0, 0)),
        ("viscosity", 6, (0, 0, 0)), ("soundspeed", 7, (0, 0, 0)), ("density1", 2, (0, 0, 0)),
        ("energy1", 4, (0, 0, 0)), ("xvel0", 8, (1, 1, 1)), ("yvel0", 10, (1, 1, 1)), ("zvel0", 12, (1, 1, 1)),
        ("xvel1", 9, (1, 1, 1)), ("yvel1", 11, (1, 1, 1)), ("zvel1", 13, (1, 1, 1)),
        ("vol_flux_x", 14, (1, 0, 0)), ("vol_flux_y", 15, (0, 1, 0)), ("vol_flux_z", 16, (0, 0, 1)),
        ("mass_flux_x", 17, (1, 0, 0)), ("mass_flux_y", 18, (0, 1, 0)), ("mass_flux_z", 19, (0, 0, 1))]

# the field sets of the reference's update_halo calls (start.f90:121-127,
# timestep.f90:47-66, PdV.f90:99, advection.f90:42-146), plus all fields and
# energy1 without density1 (the gating quirk)
HALO_SETS = {
    "start": [1, 3, 5, 6, 2, 4, 8, 10, 12, 9, 11, 13],
    "timestep": [5, 3, 1, 8, 10],
    "visc": [6],
    "pdv": [5],
    "cell": [4, 2, 14, 15, 16],
    "mom": [2, 4, 9, 11, 13, 17, 18, 19],
    "all": list(range(1, 20)),
    "e1_only": [4, 12],
}

POISON = 7.77e77


def _tiled_vs_untiled(X, Y, Z, T, dims, fset, depth, lib, seed):
    rng = np.random.default_rng(seed)
    _, tiles = _oracle_split(X, Y, Z, T, dims)
    fields = np.zeros(19, np.int32)
    for f in HALO_SETS[fset]:
        fields[f - 1] = 1
    # the untiled chunk
    G = [np.asfortranarray(rng.uniform(-1, 1, tuple(n + 4 + c for n, c in zip((X, Y, Z), cls))))
         for _, _, cls in HALO]
    g_after = [a.copy(order="F") for a in G]
    # each tile: its window of the chunk, with the halo cells that lie inside the chunk poisoned
    tarr, dims3, nbrs, masks = [], [], [], []
    for t in tiles:
        arrs, band_masks = [], []
        for (name, fid, cls), g in zip(HALO, G):
            sl = tuple(slice(o - 1, o - 1 + n + 4 + c) for o, n, c in zip(t["off"], t["n"], cls))
            a = g[sl].copy(order="F")
            # per axis: local index i (from -1), tile interior [1, n+c], chunk interior [1, N+c]
            inside_band, halo_any, chunk_inner = True, False, True
            grids = np.meshgrid(*[np.arange(-1, n + 3 + c) for n, c in zip(t["n"], cls)], indexing="ij")
            inside_band = np.ones(a.shape, bool)
            halo_any = np.zeros(a.shape, bool)
            chunk_inner = np.ones(a.shape, bool)
            for ax in range(3):
                i, n, c, o, N = grids[ax], t["n"][ax], cls[ax], t["off"][ax], (X, Y, Z)[ax]
                inside_band &= (i >= 1 - depth) & (i <= n + c + depth)
                halo_any |= (i < 1) | (i > n + c)
                gi = i + o - 1
                chunk_inner &= (gi >= 1) & (gi <= N + c)
            a[inside_band & halo_any & chunk_inner] = POISON
            arrs.append(a)
            band_masks.append(inside_band)
        tarr.append(arrs)
        dims3 += list(t["n"])
        nbrs += t["nbr"]
        masks.append(band_masks)
    tbefore = [[a.copy(order="F") for a in arrs] for arrs in tarr]
    L = ol.lib()
    L.oracle_update_halo_tiles.restype = ctypes.c_int
    ptrs = (P * (19 * len(tiles)))(*[a.ctypes.data for arrs in tarr for a in arrs])
    rc = L.oracle_update_halo_tiles(len(tiles), _ptr(np.array(dims3, np.int32)), _ptr(np.array(nbrs, np.int32)),
                                    ptrs, _ptr(fields), depth, lib)
    assert rc == 0
    gptrs = (P * 19)(*[a.ctypes.data for a in g_after])
    rc = L.oracle_update_halo_tiles(1, _ptr(np.array([X, Y, Z], np.int32)), _ptr(np.full(6, -1, np.int32)),
                                    gptrs, _ptr(fields), depth, lib)
    assert rc == 0
    # every flagged field's halo band must hold the untiled result; fields not
    # flagged keep every value they had
    bad = {}
    for t, arrs, before, band_masks in zip(tiles, tarr, tbefore, masks):
        for (name, fid, cls), a, a0, g, band in zip(HALO, arrs, before, g_after, band_masks):
            if fields[fid - 1]:
                sl = tuple(slice(o - 1, o - 1 + n + 4 + c) for o, n, c in zip(t["off"], t["n"], cls))
                diff = band & (a.view(np.uint64) != g[sl].view(np.uint64))
            else:
                diff = a.view(np.uint64) != a0.view(np.uint64)
            if diff.any():
                bad.setdefault(name, []).append((t, diff, a))
    return bad, fields


FACE_NORMAL = {"vol_flux_x": 0, "mass_flux_x": 0, "vol_flux_y": 1, "mass_flux_y": 1, "vol_flux_z": 2,
               "mass_flux_z": 2}
EXT_FACES = ((0, 1), (2, 3), (4, 5))  # (low, high) neighbour slots per axis


def _face_corner_ok(name, t, idx, depth):
    """Face data: update_halo_kernel reflects it across the tangential faces
    only up to the (max+depth)-th face along its normal (update_halo_kernel.f90
    1018-1142, 1156-1280, 1287-1417), so in a corner where such a face is
    external, the tile's last band face (n+1+depth) keeps what the tile copy
    wrote, while the untiled chunk reflects the same global face."""
    a = FACE_NORMAL[name]
    if not np.all(idx[:, a] == t["n"][a] + 1 + depth):
        return False
    for row in idx:
        ext = False
        for ax in range(3):
            if ax == a:
                continue
            lo, hi = EXT_FACES[ax]
            if (row[ax] < 1 and t["nbr"][lo] == -1) or (row[ax] > t["n"][ax] and t["nbr"][hi] == -1):
                ext = True
        if not ext:
            return False
    return True


def _quirk_ok(name, entries, fields, depth):
    """The differences are exactly the reference's documented tile-halo quirks."""
    for t, diff, a in entries:
        if name in FACE_NORMAL and _face_corner_ok(name, t, np.argwhere(diff) - 1, depth):
            continue
        if name == "yvel0" and fields[12 - 1] and not fields[10 - 1]:
            # ... and so writes yvel0's right x band even when yvel0 is not flagged
            idx = np.argwhere(diff)
            if t["nbr"][1] == -1 or not np.all(idx[:, 0] - 1 > t["n"][0] + 1):
                return False
            continue
        if name == "zvel0" and fields[12 - 1]:
            # right-face ZVEL0 branch copies yvel0: the right x band keeps its old value
            nx = t["n"][0]
            idx = np.argwhere(diff)
            if t["nbr"][1] == -1 or not np.all(idx[:, 0] - 1 > nx + 1):
                return False
        elif name == "energy1" and fields[4 - 1] and not fields[2 - 1]:
            # left/right/bottom/top copies of energy1 are gated by density1's flag:
            # those bands keep their old value, the back/front copies happen
            idx = np.argwhere(diff) - 1  # local indices
            n = t["n"]
            in_xy_band = (idx[:, 0] < 1) | (idx[:, 0] > n[0]) | (idx[:, 1] < 1) | (idx[:, 1] > n[1])
            if not np.all(in_xy_band):
                return False
        else:
            return False
    return True


TILINGS = [((9, 8, 7), 8, 3), ((11, 7, 6), 6, 3), ((10, 9, 8), 3, 1), ((8, 12, 7), 4, 2), ((13, 6, 5), 12, 3)]


@pytest.mark.parametrize("tiling", TILINGS)
@pytest.mark.parametrize("fset", list(HALO_SETS))
@pytest.mark.parametrize("depth", [1, 2])
@pytest.mark.parametrize("kernels", ["oracle", "reference"])
def test_tiled_update_halo_equals_untiled(tiling, fset, depth, kernels):
    if kernels == "reference" and not ol.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    (X, Y, Z), T, dims = tiling
    lib = str(ol.REF_SO).encode() if kernels == "reference" else None
    bad, fields = _tiled_vs_untiled(X, Y, Z, T, dims, fset, depth, lib, seed=zlib.crc32(repr((tiling, fset, depth)).encode()))
    for name, entries in bad.items():
        assert _quirk_ok(name, entries, fields, depth), (name, fset, depth,
                                                         [int(e[1].sum()) for e in entries])
    if fset in ("all", "start"):
        # the ZVEL0 quirk is real: some tile with an inside right face keeps its band
        _, tiles = _oracle_split(X, Y, Z, T, dims)
        if any(t["nbr"][1] != -1 for t in tiles):
            assert "zvel0" in bad
    if fset == "e1_only":
        _, tiles = _oracle_split(X, Y, Z, T, dims)
        if any(n != -1 for t in tiles for n in t["nbr"][:4]):
            assert "energy1" in bad


# ----------------------------------------------------------------- whole runs over tiles

DUMP_NAMES = ol.DUMP_NAMES


def _run_dump(deck, use_ref):
    rows, dts, fields = ol.run_with_fields(deck, use_ref=use_ref)
    return rows, dts, {k: np.ravel(v, order="F") for k, v in fields.items()}


def _tiled(deck_text, tiles, dims):
    d = parse_deck(deck_text)
    d.tiles_per_chunk = tiles
    d.tile_dims = dims
    return d


YZ_DECK = """*clover
 state 1 density=0.2 energy=1.0
 state 2 density=1.0 energy=2.5 geometry=cuboid xmin=0 xmax=5 ymin=0 ymax=2 zmin=0 zmax=2
 state 3 density=0.6 energy=1.9 xvel=0.05 geometry=sphere xmin=7 ymin=7 zmin=6 radius=2.2
 x_cells=14 y_cells=18 z_cells=16
 xmin=0 ymin=0 zmin=0 xmax=10 ymax=10 zmax=10
 initial_timestep=0.04 timestep_rise=1.5 end_step=20 end_time=100 summary_frequency=5
*endclover"""


@needs_ref
def test_z_split_is_bit_identical_to_untiled():
    """1x1x2 tiles: every field and every dt equal the untiled run's over 20
    steps; only the summary sums differ, by their tile order."""
    base = _tiled(YZ_DECK, 1, 3)
    tiled = _tiled(YZ_DECK, 2, 1)
    assert _oracle_split(tiled.x_cells, tiled.y_cells, tiled.z_cells, 2, 1)[0] == (1, 1, 2)
    ra, da, fa = _run_dump(base, True)
    rb, db, fb = _run_dump(tiled, True)
    assert np.array_equal(da.view(np.uint64), db.view(np.uint64))
    for k in DUMP_NAMES:
        assert _bitwise(fa[k], fb[k]), k
    for x, y in zip(ra, rb):
        for k in ("mass", "ie", "ke", "press"):
            assert x[k] == pytest.approx(y[k], rel=1e-13), k


@needs_ref
@pytest.mark.parametrize("split", [(2, 2, (1, 2, 1)), (4, 2, (1, 2, 2))])
def test_y_split_departs_at_the_face_data_corner(split):
    """y splits: bit-identical for four steps; then the first differences sit
    on the planes next to the external z_max face and within two rows of the
    tile boundary in y.  There the y momentum sweep's node flux reads
    mass_flux_y at the tile's (y_max+3, z_max+1) corner, which update_halo's
    front-face mirror of face data does not reach (it stops at y_max+depth),
    so it holds the tile copy of the neighbour's previous exchange instead of
    the untiled chunk's mirror.  The reference with tiles gives this answer."""
    tiles, dims, want = split
    assert _oracle_split(14, 18, 16, tiles, dims)[0] == want
    for steps, same in ((4, True), (5, False)):
        txt = YZ_DECK.replace("end_step=20", f"end_step={steps}")
        _, da, fa = _run_dump(_tiled(txt, 1, 3), True)
        _, db, fb = _run_dump(_tiled(txt, tiles, dims), True)
        diff = {k: ~np.equal(fa[k].view(np.uint64), fb[k].view(np.uint64)) for k in DUMP_NAMES}
        assert same == (not any(d.any() for d in diff.values())), steps
    boundary = 9  # tile 1 holds y cells 1..9: its top node is node 9 from 0
    for k, d in diff.items():
        node = k in DUMP_NAMES[7:]
        idx = np.argwhere(d.reshape(16 + node, 18 + node, 14 + node))  # (z, y, x) from 0
        assert np.all(idx[:, 0] >= 15 - 1 + node), k                   # the last two planes in z
        assert np.all(np.abs(idx[:, 1] - boundary) <= 2), k


@needs_ref
def test_x_split_changes_the_answer():
    """Tiles along x move apart from the untiled run within a few steps (the
    survey measured 2.5e-7 in KE for 2x1x1 tiles at 64x16x16): the tiled run
    is a different answer of the reference, not a reordering."""
    # 2 tiles over a deck that is longest in x split it along x
    tx = parse_deck(YZ_DECK.replace("x_cells=14 y_cells=18 z_cells=16", "x_cells=20 y_cells=12 z_cells=11"))
    bx = parse_deck(YZ_DECK.replace("x_cells=14 y_cells=18 z_cells=16", "x_cells=20 y_cells=12 z_cells=11"))
    tx.tiles_per_chunk = 2
    assert _oracle_split(tx.x_cells, tx.y_cells, tx.z_cells, 2, 3)[0] == (2, 1, 1)
    _, _, fa = _run_dump(bx, True)
    _, _, fb = _run_dump(tx, True)
    assert not _bitwise(fa["density0"], fb["density0"])
