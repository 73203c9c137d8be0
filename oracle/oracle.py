"""ctypes front end of the CPU oracle (oracle/uw_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker and as the timed reference CPU path. The engine never
imports it. Build with ``make -C oracle``.

Every function here is a thin wrapper; the arithmetic and its reference citations live in
uw_oracle.c.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
MAX_OCTAVES = 16


class UwoParams(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint32),
        ("octaves", ctypes.c_int32),
        ("frequency", ctypes.c_double),
        ("lacunarity", ctypes.c_double),
        ("amplitudes", ctypes.c_double * MAX_OCTAVES),
        ("iso", ctypes.c_float),
        ("bias_kind", ctypes.c_int32),
        ("bias_a", ctypes.c_float),
        ("bias_b", ctypes.c_float),
        ("bias_c", ctypes.c_float),
        ("chunk_extent", ctypes.c_int32),
        ("offset", ctypes.c_double * 3),
    ]


class UwoMeshCounts(ctypes.Structure):
    _fields_ = [("n_verts", ctypes.c_int32), ("n_inds", ctypes.c_int32),
                ("n_active", ctypes.c_int32), ("overflow", ctypes.c_int32)]


class UwoChunkResult(ctypes.Structure):
    _fields_ = [("n_verts", ctypes.c_int32), ("n_inds", ctypes.c_int32),
                ("n_active", ctypes.c_int32), ("early_blank", ctypes.c_int32),
                ("checksum", ctypes.c_uint64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is missing: run `make -C oracle`")
        L = ctypes.CDLL(path)
        P = ctypes.POINTER
        L.uwo_xorshift_stream.argtypes = [ctypes.c_uint32] * 4 + [ctypes.c_int, P(ctypes.c_uint32)]
        L.uwo_gen_index_histogram16.argtypes = [ctypes.c_uint32, P(ctypes.c_uint32)]
        L.uwo_perm.argtypes = [ctypes.c_uint32, P(ctypes.c_uint8)]
        L.uwo_perlin3.argtypes = [P(ctypes.c_uint8), ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.uwo_perlin3.restype = ctypes.c_double
        L.uwo_perlin3_many.argtypes = [P(ctypes.c_uint8), P(ctypes.c_double), ctypes.c_int64, P(ctypes.c_double)]
        L.uwo_fbm.argtypes = [P(ctypes.c_uint8), P(UwoParams), ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.uwo_fbm.restype = ctypes.c_double
        L.uwo_iso_at.argtypes = [P(ctypes.c_uint8), P(UwoParams), ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.uwo_iso_at.restype = ctypes.c_float
        L.uwo_coord.argtypes = [P(UwoParams), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
        L.uwo_coord.restype = ctypes.c_double
        L.uwo_step.argtypes = [P(UwoParams), ctypes.c_int32]
        L.uwo_step.restype = ctypes.c_float
        L.uwo_density_ref.argtypes = [P(UwoParams), P(ctypes.c_int32), ctypes.c_int32, P(ctypes.c_float)]
        L.uwo_density_fp64.argtypes = [P(UwoParams), P(ctypes.c_int32), ctypes.c_int32, P(ctypes.c_double)]
        L.uwo_early_blank.argtypes = [P(ctypes.c_float), ctypes.c_int64, ctypes.c_float]
        L.uwo_early_blank.restype = ctypes.c_int
        L.uwo_mesh_ref.argtypes = [P(UwoParams), P(ctypes.c_int32), ctypes.c_int32, P(ctypes.c_float),
                                   ctypes.c_int, P(ctypes.c_float), ctypes.c_int64, P(ctypes.c_uint32),
                                   ctypes.c_int64, P(ctypes.c_uint8), P(ctypes.c_uint32), P(UwoMeshCounts)]
        L.uwo_mesh_ref.restype = ctypes.c_int
        L.uwo_mesh_ref_rgb.argtypes = [P(UwoParams), P(ctypes.c_int32), ctypes.c_int32, P(ctypes.c_float),
                                       P(ctypes.c_float), P(ctypes.c_float), ctypes.c_int64, P(ctypes.c_uint32),
                                       ctypes.c_int64, P(UwoMeshCounts)]
        L.uwo_mesh_ref_rgb.restype = ctypes.c_int
        L.uwo_mesh_ref_tris.argtypes = [P(UwoParams), P(ctypes.c_int32), ctypes.c_int32, P(ctypes.c_float),
                                        P(ctypes.c_float), ctypes.c_int64, P(ctypes.c_uint32), ctypes.c_int64,
                                        P(ctypes.c_uint32), P(ctypes.c_float), P(UwoMeshCounts)]
        L.uwo_mesh_ref_tris.restype = ctypes.c_int
        L.uwo_tri_normal.argtypes = [P(ctypes.c_float)] * 4
        L.uwo_tri_normal.restype = None
        L.uwo_tris_around.argtypes = [P(ctypes.c_uint32), ctypes.c_int64, ctypes.c_int32, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_float, ctypes.c_int32, P(ctypes.c_uint32),
                                      ctypes.c_int64]
        L.uwo_tris_around.restype = ctypes.c_int64
        L.uwo_mesh_fp64.argtypes = [P(UwoParams), P(ctypes.c_int32), ctypes.c_int32, P(ctypes.c_double),
                                    P(ctypes.c_double), ctypes.c_int64, P(ctypes.c_uint32), ctypes.c_int64,
                                    P(ctypes.c_uint8), P(UwoMeshCounts)]
        L.uwo_mesh_fp64.restype = ctypes.c_int
        L.uwo_build_chunks.argtypes = [P(UwoParams), P(ctypes.c_int32), ctypes.c_int64, ctypes.c_int32,
                                       ctypes.c_int, ctypes.c_int, P(UwoChunkResult)]
        L.uwo_build_chunks.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def make_params(p: dict) -> UwoParams:
    """dict (see underwaterworld_amd.params) -> oracle struct."""
    s = UwoParams()
    s.seed = int(p["seed"]) & 0xFFFFFFFF
    s.octaves = int(p["octaves"])
    s.frequency = float(p["frequency"])
    s.lacunarity = float(p["lacunarity"])
    amps = list(p["amplitudes"])
    for k in range(MAX_OCTAVES):
        s.amplitudes[k] = float(amps[k]) if k < len(amps) else 0.0
    s.iso = float(p["iso"])
    s.bias_kind = int(p["bias_kind"])
    s.bias_a, s.bias_b, s.bias_c = (float(v) for v in p["bias"])
    s.chunk_extent = int(p["chunk_extent"])
    for i, v in enumerate(p.get("offset", (0.0, 0.0, 0.0))):
        s.offset[i] = float(v)
    return s


def xorshift_stream(x, y, z, w, n):
    out = np.zeros(n, np.uint32)
    lib().uwo_xorshift_stream(x, y, z, w, n, _ptr(out, ctypes.c_uint32))
    return out


def gen_index_histogram16(rng):
    out = np.zeros(rng, np.uint32)
    lib().uwo_gen_index_histogram16(rng, _ptr(out, ctypes.c_uint32))
    return out


def perm(seed: int) -> np.ndarray:
    out = np.zeros(256, np.uint8)
    lib().uwo_perm(int(seed) & 0xFFFFFFFF, _ptr(out, ctypes.c_uint8))
    return out


def perlin3(P: np.ndarray, pts: np.ndarray) -> np.ndarray:
    P = np.ascontiguousarray(P, np.uint8)
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    out = np.zeros(len(pts), np.float64)
    lib().uwo_perlin3_many(_ptr(P, ctypes.c_uint8), _ptr(pts, ctypes.c_double), len(pts), _ptr(out, ctypes.c_double))
    return out


def fbm(params: dict, x, y, z) -> float:
    P = perm(params["seed"])
    s = make_params(params)
    return lib().uwo_fbm(_ptr(P, ctypes.c_uint8), ctypes.byref(s), x, y, z)


def coords(params: dict, c: int, n: int) -> np.ndarray:
    s = make_params(params)
    return np.array([lib().uwo_coord(ctypes.byref(s), c, n, i) for i in range(n + 1)])


def step(params: dict, n: int) -> float:
    return lib().uwo_step(ctypes.byref(make_params(params)), n)


def density_ref(params: dict, coord, n: int) -> np.ndarray:
    s = make_params(params)
    c = np.ascontiguousarray(coord, np.int32)
    out = np.zeros((n + 1) ** 3, np.float32)
    lib().uwo_density_ref(ctypes.byref(s), _ptr(c, ctypes.c_int32), n, _ptr(out, ctypes.c_float))
    return out


def density_fp64(params: dict, coord, n: int) -> np.ndarray:
    s = make_params(params)
    c = np.ascontiguousarray(coord, np.int32)
    out = np.zeros((n + 1) ** 3, np.float64)
    lib().uwo_density_fp64(ctypes.byref(s), _ptr(c, ctypes.c_int32), n, _ptr(out, ctypes.c_double))
    return out


def early_blank(dens: np.ndarray, iso: float) -> bool:
    d = np.ascontiguousarray(dens, np.float32)
    return bool(lib().uwo_early_blank(_ptr(d, ctypes.c_float), d.size, iso))


def vert_capacity(n: int) -> int:
    return 5 * n * (n + 1) ** 2 + 16


def mesh_ref(params: dict, coord, n: int, dens: np.ndarray, dedup: str = "hash", with_tri_cells=False):
    """Reference-faithful mesh of one chunk from its f32 density grid.
    Returns dict(verts (V,3) f32, inds (I,) u32, cases (n^3,) u8, n_active, tri_cells)."""
    s = make_params(params)
    c = np.ascontiguousarray(coord, np.int32)
    d = np.ascontiguousarray(dens, np.float32)
    assert d.size == (n + 1) ** 3
    vcap, icap = vert_capacity(n), 15 * n ** 3 + 16
    verts = np.zeros((vcap, 3), np.float32)
    inds = np.zeros(icap, np.uint32)
    cases = np.zeros(n ** 3, np.uint8)
    tri_cells = np.zeros(icap // 3 + 1, np.uint32)
    cnt = UwoMeshCounts()
    rc = lib().uwo_mesh_ref(ctypes.byref(s), _ptr(c, ctypes.c_int32), n, _ptr(d, ctypes.c_float),
                            0 if dedup == "linear" else 1, _ptr(verts, ctypes.c_float), vcap,
                            _ptr(inds, ctypes.c_uint32), icap, _ptr(cases, ctypes.c_uint8),
                            _ptr(tri_cells, ctypes.c_uint32), ctypes.byref(cnt))
    if rc != 0:
        raise RuntimeError("uwo_mesh_ref failed: %d" % rc)
    out = dict(verts=verts[:cnt.n_verts].copy(), inds=inds[:cnt.n_inds].copy(), cases=cases,
               n_active=cnt.n_active)
    if with_tri_cells:
        out["tri_cells"] = tri_cells[:cnt.n_inds // 3].copy()
    return out


def mesh_ref_rgb(params: dict, coord, n: int, dens: np.ndarray):
    """mesh_ref with the reference's VertColor colours: (verts (V,3), colors (V,3), inds)."""
    s = make_params(params)
    c = np.ascontiguousarray(coord, np.int32)
    d = np.ascontiguousarray(dens, np.float32)
    vcap, icap = vert_capacity(n), 15 * n ** 3 + 16
    verts = np.zeros((vcap, 3), np.float32)
    colors = np.zeros((vcap, 3), np.float32)
    inds = np.zeros(icap, np.uint32)
    cnt = UwoMeshCounts()
    rc = lib().uwo_mesh_ref_rgb(ctypes.byref(s), _ptr(c, ctypes.c_int32), n, _ptr(d, ctypes.c_float),
                                _ptr(verts, ctypes.c_float), _ptr(colors, ctypes.c_float), vcap,
                                _ptr(inds, ctypes.c_uint32), icap, ctypes.byref(cnt))
    if rc != 0:
        raise RuntimeError("uwo_mesh_ref_rgb failed: %d" % rc)
    return verts[:cnt.n_verts].copy(), colors[:cnt.n_verts].copy(), inds[:cnt.n_inds].copy()


def mesh_ref_tris(params: dict, coord, n: int, dens: np.ndarray):
    """mesh_ref with Build.tris: (verts, inds, tri_cells (T,) u32, tri_normals (T,3) f32)."""
    s = make_params(params)
    c = np.ascontiguousarray(coord, np.int32)
    d = np.ascontiguousarray(dens, np.float32)
    vcap, icap = vert_capacity(n), 15 * n ** 3 + 16
    verts = np.zeros((vcap, 3), np.float32)
    inds = np.zeros(icap, np.uint32)
    cells = np.zeros(icap // 3 + 1, np.uint32)
    normals = np.zeros((icap // 3 + 1, 3), np.float32)
    cnt = UwoMeshCounts()
    rc = lib().uwo_mesh_ref_tris(ctypes.byref(s), _ptr(c, ctypes.c_int32), n, _ptr(d, ctypes.c_float),
                                 _ptr(verts, ctypes.c_float), vcap, _ptr(inds, ctypes.c_uint32), icap,
                                 _ptr(cells, ctypes.c_uint32), _ptr(normals, ctypes.c_float), ctypes.byref(cnt))
    if rc != 0:
        raise RuntimeError("uwo_mesh_ref_tris failed: %d" % rc)
    nt = cnt.n_inds // 3
    return verts[:cnt.n_verts].copy(), inds[:cnt.n_inds].copy(), cells[:nt].copy(), normals[:nt].copy()


def tri_normal(a, b, c) -> np.ndarray:
    """util::Tri::new's normal of one triangle (f32)."""
    pts = [np.ascontiguousarray(v, np.float32) for v in (a, b, c)]
    out = np.zeros(3, np.float32)
    lib().uwo_tri_normal(*[_ptr(v, ctypes.c_float) for v in pts], _ptr(out, ctypes.c_float))
    return out


def tris_around(tri_cells: np.ndarray, n: int, pct, range_: int) -> np.ndarray:
    """Chunk::tris_around restated literally: triangle ids in the reference's order."""
    tc = np.ascontiguousarray(tri_cells, np.uint32)
    cap = max(len(tc), 1)
    ids = np.zeros(cap, np.uint32)
    k = lib().uwo_tris_around(_ptr(tc, ctypes.c_uint32), len(tc), int(n), float(pct[0]), float(pct[1]),
                              float(pct[2]), int(range_), _ptr(ids, ctypes.c_uint32), cap)
    return ids[:k].copy()


def vertex_color(world_z: float, corner_b: int, chunk_size: float = 16.0) -> np.ndarray:
    """The reference's per-vertex terrain colour (chunk.rs:215-222) for one vertex."""
    L = lib()
    if not hasattr(L, "_color_ready"):
        L.uwo_vertex_color.argtypes = [ctypes.c_float, ctypes.c_int, ctypes.c_float, ctypes.POINTER(ctypes.c_float)]
        L.uwo_vertex_color.restype = None
        L._color_ready = True
    out = np.zeros(3, np.float32)
    L.uwo_vertex_color(float(world_z), int(corner_b), float(chunk_size), _ptr(out, ctypes.c_float))
    return out


def mesh_fp64(params: dict, coord, n: int, dens64: np.ndarray):
    s = make_params(params)
    c = np.ascontiguousarray(coord, np.int32)
    d = np.ascontiguousarray(dens64, np.float64)
    vcap, icap = vert_capacity(n), 15 * n ** 3 + 16
    verts = np.zeros((vcap, 3), np.float64)
    inds = np.zeros(icap, np.uint32)
    cases = np.zeros(n ** 3, np.uint8)
    cnt = UwoMeshCounts()
    rc = lib().uwo_mesh_fp64(ctypes.byref(s), _ptr(c, ctypes.c_int32), n, _ptr(d, ctypes.c_double),
                             _ptr(verts, ctypes.c_double), vcap, _ptr(inds, ctypes.c_uint32), icap,
                             _ptr(cases, ctypes.c_uint8), ctypes.byref(cnt))
    if rc != 0:
        raise RuntimeError("uwo_mesh_fp64 failed: %d" % rc)
    return dict(verts=verts[:cnt.n_verts].copy(), inds=inds[:cnt.n_inds].copy(), cases=cases,
                n_active=cnt.n_active)


def build_chunk(params: dict, coord, n: int, dedup: str = "hash"):
    """Full reference chunk build: density, early blank check, mesh (chunk.rs:266-313)."""
    dens = density_ref(params, coord, n)
    if early_blank(dens, params["iso"]):
        return dict(dens=dens, early_blank=True, verts=np.zeros((0, 3), np.float32),
                    inds=np.zeros(0, np.uint32), cases=None, n_active=0)
    m = mesh_ref(params, coord, n, dens, dedup)
    m.update(dens=dens, early_blank=False)
    return m


def build_chunks(params: dict, coords: np.ndarray, n: int, nthreads: int = 1, dedup: str = "linear"):
    """Bulk reference builds (counts + checksum per chunk); the CPU-baseline workload."""
    s = make_params(params)
    c = np.ascontiguousarray(coords, np.int32).reshape(-1, 3)
    res = (UwoChunkResult * len(c))()
    lib().uwo_build_chunks(ctypes.byref(s), _ptr(c, ctypes.c_int32), len(c), n,
                           0 if dedup == "linear" else 1, nthreads, res)
    return np.array([(r.n_verts, r.n_inds, r.n_active, r.early_blank) for r in res], np.int64)


def _adjoint_lib():
    L = lib()
    if not hasattr(L, "_adjoint_ready"):
        P = ctypes.POINTER
        L.uwo_adjoint_chunk.argtypes = [P(UwoParams), P(ctypes.c_int32), ctypes.c_int32, P(ctypes.c_uint8),
                                        P(ctypes.c_double), ctypes.c_double, P(ctypes.c_double), P(ctypes.c_double),
                                        P(ctypes.c_double), P(ctypes.c_int32)]
        L.uwo_adjoint_chunk.restype = ctypes.c_double
        L.uwo_plane_loss_dens.argtypes = [ctypes.c_int32, ctypes.c_double, ctypes.c_double, P(ctypes.c_double),
                                          P(ctypes.c_uint8), ctypes.c_double, P(ctypes.c_double)]
        L.uwo_plane_loss_dens.restype = ctypes.c_double
        L.uwo_perlin3_grad.argtypes = [P(ctypes.c_uint8), ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       P(ctypes.c_double)]
        L.uwo_perlin3_grad.restype = ctypes.c_double
        L._adjoint_ready = True
    return L


def perlin3_grad(P, x, y, z):
    g = np.zeros(3, np.float64)
    P = np.ascontiguousarray(P, np.uint8)
    v = _adjoint_lib().uwo_perlin3_grad(_ptr(P, ctypes.c_uint8), x, y, z, _ptr(g, ctypes.c_double))
    return v, g


def adjoint_chunk(params: dict, coord, n: int, target_z: float, cases=None, grads: bool = True, rho=None):
    """FP64 plane loss of one chunk, sum over vertices of (v_z_local - target_z)^2, and its analytic
    gradients w.r.t. the densities, octave amplitudes and the noise-space offset. `cases` freezes
    the topology (n^3 u8); None classifies the FP64 densities."""
    L = _adjoint_lib()
    s = make_params(params)
    c = np.ascontiguousarray(coord, np.int32)
    S = (n + 1) ** 3
    gd = np.zeros(S, np.float64)
    ga = np.zeros(params["octaves"], np.float64)
    go = np.zeros(3, np.float64)
    nv = ctypes.c_int32()
    cs = None if cases is None else np.ascontiguousarray(cases, np.uint8)
    rr = None if rho is None else np.ascontiguousarray(rho, np.float64)
    loss = L.uwo_adjoint_chunk(ctypes.byref(s), _ptr(c, ctypes.c_int32), n,
                               _ptr(cs, ctypes.c_uint8) if cs is not None else None,
                               _ptr(rr, ctypes.c_double) if rr is not None else None, target_z,
                               _ptr(gd, ctypes.c_double) if grads else None,
                               _ptr(ga, ctypes.c_double) if grads else None,
                               _ptr(go, ctypes.c_double) if grads else None, ctypes.byref(nv))
    return dict(loss=loss, nverts=nv.value, g_dens=gd, g_amp=ga, g_off=go)


def plane_loss_dens(n: int, step: float, iso: float, rho64: np.ndarray, cases: np.ndarray, target_z: float,
                    grads: bool = False):
    L = _adjoint_lib()
    r = np.ascontiguousarray(rho64, np.float64)
    cs = np.ascontiguousarray(cases, np.uint8)
    gd = np.zeros(r.size, np.float64)
    loss = L.uwo_plane_loss_dens(n, step, iso, _ptr(r, ctypes.c_double), _ptr(cs, ctypes.c_uint8), target_z,
                                 _ptr(gd, ctypes.c_double) if grads else None)
    return (loss, gd) if grads else loss


def iso_at(params: dict, pts: np.ndarray) -> np.ndarray:
    """perlin_util::iso_at at noise-space points (N, 3): the boids' point query."""
    L = lib()
    P = perm(params["seed"])
    s = make_params(params)
    pts = np.asarray(pts, np.float64).reshape(-1, 3)
    f = params["frequency"]
    off = [float(v) * f for v in params.get("offset", (0.0, 0.0, 0.0))]
    Pp = _ptr(P, ctypes.c_uint8)
    return np.array([L.uwo_iso_at(Pp, ctypes.byref(s), x + off[0], y + off[1], z + off[2]) for x, y, z in pts],
                    np.float32)
