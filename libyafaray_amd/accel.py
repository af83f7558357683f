"""Python handle on the "mi355x-lbvh" accelerator (C ABI in include/ymx_accel.h).

Mirrors the query side of libYafaRay's Accelerator
(include/accelerator/accelerator.h:37-87): `intersect` is the batched
Accelerator::intersect(ray, t_max), `intersect_shadow` the batched
Accelerator::intersectShadow(ray, t_max), `bound` is getBound().  PyTorch only
supplies device memory and the stream; all work runs in libymx.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native
from .native import YMX_NO_HIT, YmxAccelInfo, YmxAccelParams, check

RAY_FLOATS = 8  # org xyz, tmin, dir xyz, tmax
HIT_WORDS = 4   # prim, t, u, v


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise native.YmxError("the mi355x-lbvh accelerator needs a HIP device")
    return torch


class GpuAccelerator:
    """Immutable BVH over a primitive list; queries are batched on the device.

    tris: (n, 3) triangles, or (n, 4) faces whose 4th index is YMX_NO_VERTEX
    for a triangle and a vertex for a quad (one primitive per face, tested as
    the reference's ShapePolygon<float, 4>)."""

    type_name = "mi355x-lbvh"

    def __init__(self, verts: np.ndarray, tris: np.ndarray, vis: np.ndarray | None = None,
                 max_leaf_size: int = 0, builder: str = "lbvh", treelet_passes: int = 0, ray_reorder: int = 0,
                 reorder_cell_bits: int = 0):
        self._lib = native.lib()
        verts = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
        tris = np.ascontiguousarray(tris, dtype=np.uint32)
        faces = tris.ndim == 2 and tris.shape[1] == 4
        tris = tris.reshape(-1, 4 if faces else 3)
        if vis is not None:
            vis = np.ascontiguousarray(vis, dtype=np.uint8).reshape(-1)
            if vis.shape[0] != tris.shape[0]:
                raise ValueError("one visibility mask per triangle is required")
        if builder not in native.BUILDERS:
            raise ValueError(f"unknown BVH builder {builder!r} (one of {sorted(native.BUILDERS)})")
        params = YmxAccelParams(C.sizeof(YmxAccelParams), int(max_leaf_size), native.BUILDERS[builder],
                                int(treelet_passes), int(ray_reorder), int(reorder_cell_bits))
        handle = C.c_void_p()
        create = self._lib.ymx_accel_create_faces if faces else self._lib.ymx_accel_create
        rc = create(
            verts.ctypes.data_as(C.c_void_p), verts.shape[0],
            tris.ctypes.data_as(C.c_void_p), None if vis is None else vis.ctypes.data_as(C.c_void_p),
            tris.shape[0], C.byref(params), C.byref(handle))
        check(rc, "ymx_accel_create")
        self._h = handle
        self.n_tris = int(tris.shape[0])

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.ymx_accel_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def bound(self) -> tuple[np.ndarray, np.ndarray]:
        a = (C.c_float * 3)()
        g = (C.c_float * 3)()
        check(self._lib.ymx_accel_bound(self._h, a, g), "ymx_accel_bound")
        return np.array(a[:], dtype=np.float32), np.array(g[:], dtype=np.float32)

    def info(self) -> dict:
        inf = YmxAccelInfo()
        inf.struct_size = C.sizeof(YmxAccelInfo)
        check(self._lib.ymx_accel_get_info(self._h, C.byref(inf)), "ymx_accel_get_info")
        return {
            "n_tris": inf.n_tris, "n_nodes": inf.n_nodes, "max_depth": inf.max_depth,
            "node_bytes": inf.node_bytes, "tri_bytes": inf.tri_bytes, "build_ms": inf.build_ms,
            "bound_min": list(inf.bound_min), "bound_max": list(inf.bound_max), "n_quads": inf.n_quads,
            "treelet_passes": inf.treelet_passes, "treelets_rewritten": inf.treelets_rewritten,
        }

    # --- batched queries on device tensors -------------------------------
    @staticmethod
    def _stream_ptr(stream) -> int:
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream()
        return s.cuda_stream

    def intersect_device(self, rays, hits=None, stream=None):
        """rays: (n, 8) float32 CUDA tensor; returns (n, 4) int32 view of ymx_hit."""
        torch = _torch()
        if rays.dtype != torch.float32 or rays.dim() != 2 or rays.shape[1] != RAY_FLOATS or not rays.is_contiguous():
            raise ValueError("rays must be a contiguous (n, 8) float32 tensor")
        n = rays.shape[0]
        if hits is None:
            hits = torch.empty((n, HIT_WORDS), dtype=torch.int32, device=rays.device)
        rc = self._lib.ymx_accel_intersect(self._h, rays.data_ptr(), hits.data_ptr(), n, self._stream_ptr(stream))
        check(rc, "ymx_accel_intersect")
        return hits

    def intersect_shadow_device(self, rays, occluded=None, prim=None, stream=None):
        torch = _torch()
        if rays.dtype != torch.float32 or rays.dim() != 2 or rays.shape[1] != RAY_FLOATS or not rays.is_contiguous():
            raise ValueError("rays must be a contiguous (n, 8) float32 tensor")
        n = rays.shape[0]
        if occluded is None:
            occluded = torch.empty(((n + 31) // 32,), dtype=torch.int32, device=rays.device)
        rc = self._lib.ymx_accel_intersect_shadow(
            self._h, rays.data_ptr(), occluded.data_ptr(), 0 if prim is None else prim.data_ptr(), n,
            self._stream_ptr(stream))
        check(rc, "ymx_accel_intersect_shadow")
        return occluded

    def traversal_stats(self, rays, query: int) -> tuple[int, int]:
        nf = C.c_uint64()
        nt = C.c_uint64()
        rc = self._lib.ymx_accel_traversal_stats(self._h, rays.data_ptr(), rays.shape[0], int(query),
                                                 C.byref(nf), C.byref(nt), self._stream_ptr(None))
        check(rc, "ymx_accel_traversal_stats")
        return int(nf.value), int(nt.value)

    # --- host convenience (copies) ----------------------------------------
    def intersect(self, rays: np.ndarray) -> dict:
        """Closest hit for host rays (n, 8); returns prim/t/u/v numpy arrays."""
        torch = _torch()
        r = torch.from_numpy(np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, RAY_FLOATS)).cuda()
        h = self.intersect_device(r)
        torch.cuda.synchronize()
        return unpack_hits(h.cpu().numpy())

    def intersect_transparent(self, rays: np.ndarray, transparent: np.ndarray, max_depth: int):
        """Transparent-shadow query for host rays; transparent: one flag per
        primitive.  Returns (occluded bool[n], counts[n], hits (n, max_depth, 2)
        of (primitive, t))."""
        torch = _torch()
        r = torch.from_numpy(np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, RAY_FLOATS)).cuda()
        n = r.shape[0]
        mask = torch.from_numpy(transparent_mask(transparent).view(np.int32)).cuda()
        occ = torch.empty(((n + 31) // 32,), dtype=torch.int32, device=r.device)
        counts = torch.empty((n,), dtype=torch.int32, device=r.device)
        hits = torch.zeros((max(1, n * max_depth), 2), dtype=torch.int32, device=r.device)
        rc = self._lib.ymx_accel_intersect_transparent(self._h, r.data_ptr(), mask.data_ptr(), int(max_depth),
                                                       occ.data_ptr(), counts.data_ptr(), hits.data_ptr(), n,
                                                       self._stream_ptr(None))
        check(rc, "ymx_accel_intersect_transparent")
        torch.cuda.synchronize()
        h = hits.cpu().numpy().view(np.uint32)[: n * max_depth].reshape(n, max_depth, 2)
        return unpack_bits(occ.cpu().numpy(), n), counts.cpu().numpy().view(np.uint32), h

    def intersect_shadow(self, rays: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        torch = _torch()
        r = torch.from_numpy(np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, RAY_FLOATS)).cuda()
        n = r.shape[0]
        prim = torch.empty((n,), dtype=torch.int32, device=r.device)
        occ = self.intersect_shadow_device(r, prim=prim)
        torch.cuda.synchronize()
        return unpack_bits(occ.cpu().numpy(), n), prim.cpu().numpy().view(np.uint32)


def transparent_mask(flags: np.ndarray) -> np.ndarray:
    """Per-primitive transparency flags -> the query's bitmask words (uint32)."""
    f = np.ascontiguousarray(flags, dtype=bool).reshape(-1)
    words = np.zeros((len(f) + 31) // 32 + 1, dtype=np.uint32)
    idx = np.nonzero(f)[0]
    np.bitwise_or.at(words, idx >> 5, (np.uint32(1) << (idx & 31).astype(np.uint32)))
    return words


def unpack_hits(h: np.ndarray) -> dict:
    h = np.ascontiguousarray(h).reshape(-1, 4)
    return {
        "prim": h[:, 0].view(np.uint32).copy(),
        "t": h[:, 1].view(np.float32).copy(),
        "u": h[:, 2].view(np.float32).copy(),
        "v": h[:, 3].view(np.float32).copy(),
    }


def unpack_bits(words: np.ndarray, n: int) -> np.ndarray:
    w = np.ascontiguousarray(words).view(np.uint32)
    bits = (w[:, None] >> np.arange(32, dtype=np.uint32)[None, :]) & 1
    return bits.reshape(-1)[:n].astype(np.uint8)


__all__ = ["GpuAccelerator", "unpack_hits", "unpack_bits", "YMX_NO_HIT"]
