"""Python host bindings for libogjk.so (the C ABI in include/opengjk/ogjk.h).

This is plumbing for tests and the benchmark: the library itself is C++/HIP.
Device buffers for the resident-scene entry points come from torch (ROCm);
the host entry points take numpy arrays.  Every call raises OgjkError on a
non-zero status; there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libogjk.so")

GJK, EPA, GJK_EPA = 1, 2, 3
F64_RECORD = 24
CONTACT_FLOATS = 10  # ogjk_contact: distance, normal[3], witness[2][3]

SIMPLEX_DTYPE = np.dtype({
    "names": ["nvrtx", "vrtx", "vrtx_idx", "witnesses"],
    "formats": [np.int32, (np.float32, (4, 3)), (np.int32, (4, 2)), (np.float32, (2, 3))],
    "offsets": [0, 4, 52, 84],
    "itemsize": 108,
})
POLYTOPE_DTYPE = np.dtype({
    "names": ["numpoints", "s", "s_idx", "coord"],
    "formats": [np.int32, (np.float32, 3), np.int32, np.uint64],
    "offsets": [0, 4, 16, 24],
    "itemsize": 32,
})
PAIR_DTYPE = np.dtype([("idx1", np.int32), ("idx2", np.int32)])

# Every symbol include/opengjk/ogjk.h declares (checked by the ABI test).
C_ABI_SYMBOLS = [
    "ogjk_version", "ogjk_last_error", "ogjk_device_count", "ogjk_sync",
    "ogjk_set_config", "ogjk_set_buckets", "ogjk_workspace_bytes", "ogjk_collide_scene_ws", "ogjk_pack_scene",
    "ogjk_collide_scene", "ogjk_pack_scene_f64", "ogjk_collide_scene_f64",
    "ogjk_collide_polytopes", "ogjk_collide_polytopes_indexed", "ogjk_collide_host",
    "ogjk_collide_host_indexed", "ogjk_cluster_slots", "ogjk_cluster_supers", "ogjk_build_clusters",
    "ogjk_graph_capture_scene", "ogjk_graph_launch", "ogjk_graph_destroy", "ogjk_collide_posed",
    "ogjk_collide_posed_bodies",
    "ogjk_server_create", "ogjk_server_submit", "ogjk_server_wait", "ogjk_server_poll",
    "ogjk_server_info", "ogjk_server_destroy", "ogjk_broad_phase_workspace", "ogjk_broad_phase",
    "ogjk_expand_pairs_workspace", "ogjk_expand_pairs", "ogjk_collide_scene_contacts",
    "ogjk_collide_posed_bodies_contacts", "ogjk_build_hulls", "ogjk_hull_moments",
    "ogjk_center_points", "ogjk_collide_scene_noalloc", "ogjk_pack_scene_f64_aos",
    "ogjk_set_block_waves",
]


class OgjkError(RuntimeError):
    pass


class _Clusters(C.Structure):
    _fields_ = [("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("orig", C.c_void_p),
                ("body_slot", C.c_void_p), ("bounds", C.c_void_p), ("n_slots", C.c_int64),
                ("min_verts", C.c_int32), ("super_bounds", C.c_void_p),
                ("body_super", C.c_void_p), ("n_supers", C.c_int64),
                ("n_supers_used", C.c_int64)]


class _Scene(C.Structure):
    _fields_ = [("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("body_off", C.c_void_p),
                ("body_cnt", C.c_void_p), ("body_r1", C.c_void_p), ("n_bodies", C.c_int32),
                ("clusters", C.POINTER(_Clusters))]


class _ServerConfig(C.Structure):
    _fields_ = [("max_batch_pairs", C.c_int32), ("slots", C.c_int32), ("memory_budget", C.c_int64)]


class _Grid(C.Structure):
    _fields_ = [("cell_size", C.c_float), ("boundary", C.c_float), ("grid_size", C.c_int32),
                ("max_per_cell", C.c_int32)]


class _Scene64(C.Structure):
    _fields_ = [("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("body_off", C.c_void_p),
                ("body_cnt", C.c_void_p), ("n_bodies", C.c_int32), ("xyz", C.c_void_p)]


_lib = None


def lib():
    """Load libogjk.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OgjkError(f"{LIB_PATH} missing: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        sig = {
            "ogjk_version": (C.c_int, []),
            "ogjk_last_error": (C.c_char_p, []),
            "ogjk_device_count": (C.c_int, []),
            "ogjk_sync": (C.c_int, [vp]),
            "ogjk_set_config": (C.c_int, [C.c_int, C.c_int, C.c_int]),
            "ogjk_set_buckets": (C.c_int, [C.c_int, vp, vp, vp]),
            "ogjk_set_block_waves": (C.c_int, [C.c_int, vp, vp]),
            "ogjk_workspace_bytes": (C.c_size_t, [i32]),
            "ogjk_collide_scene_ws": (C.c_int, [C.c_int, C.POINTER(_Scene), vp, i32, vp, vp, vp, vp,
                                                vp, C.c_size_t, vp]),
            "ogjk_pack_scene": (C.c_int, [vp, vp, vp, i32, i64, vp, vp, vp, vp, vp]),
            "ogjk_pack_scene_f64": (C.c_int, [vp, i64, vp, vp, vp, vp]),
            "ogjk_pack_scene_f64_aos": (C.c_int, [vp, i64, vp, vp]),
            "ogjk_cluster_slots": (i64, [i64, i32]),
            "ogjk_cluster_supers": (i64, [i64]),
            "ogjk_graph_capture_scene": (C.c_int, [C.c_int, C.POINTER(_Scene), vp, i32, vp, vp, vp,
                                                   vp, C.POINTER(vp)]),
            "ogjk_graph_launch": (C.c_int, [vp, vp]),
            "ogjk_collide_posed": (C.c_int, [C.c_int, C.POINTER(_Scene), vp, vp, i32, vp, vp, vp,
                                             vp, vp, C.c_size_t, vp]),
            "ogjk_collide_posed_bodies": (C.c_int, [C.c_int, C.POINTER(_Scene), vp, vp, i32, vp,
                                                    vp, vp, vp, vp, C.c_size_t, vp]),
            "ogjk_server_create": (C.c_int, [C.POINTER(_Scene), C.POINTER(_ServerConfig),
                                             C.POINTER(vp)]),
            "ogjk_server_submit": (C.c_int, [vp, vp, vp, i32, vp, vp, vp, vp, C.POINTER(i64)]),
            "ogjk_server_wait": (C.c_int, [vp, i64, C.POINTER(C.c_double), C.POINTER(i64)]),
            "ogjk_server_poll": (C.c_int, [vp, i64, C.POINTER(C.c_int)]),
            "ogjk_server_info": (C.c_int, [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i64),
                                           C.POINTER(i64)]),
            "ogjk_server_destroy": (C.c_int, [vp]),
            "ogjk_graph_destroy": (C.c_int, [vp]),
            "ogjk_build_clusters": (C.c_int, [C.POINTER(_Scene), C.POINTER(_Clusters), vp]),
            "ogjk_collide_scene": (C.c_int, [C.c_int, C.POINTER(_Scene), vp, i32, vp, vp, vp, vp, vp]),
            "ogjk_collide_scene_noalloc": (C.c_int, [C.c_int, C.POINTER(_Scene), vp, i32, vp, vp,
                                                     vp, vp, vp]),
            "ogjk_collide_scene_f64": (C.c_int, [C.c_int, C.POINTER(_Scene64), vp, i32, vp, vp, vp, vp, vp]),
            "ogjk_collide_polytopes": (C.c_int, [C.c_int, vp, vp, i32, vp, vp, vp, vp]),
            "ogjk_collide_polytopes_indexed": (C.c_int, [C.c_int, vp, i32, vp, i32, vp, vp, vp, vp]),
            "ogjk_collide_host": (C.c_int, [C.c_int, i32, vp, vp, vp, vp, vp]),
            "ogjk_collide_host_indexed": (C.c_int, [C.c_int, i32, i32, vp, vp, vp, vp, vp]),
            "ogjk_collide_scene_contacts": (C.c_int, [C.c_int, C.POINTER(_Scene), vp, i32, vp,
                                                      vp, vp, C.c_size_t, vp]),
            "ogjk_collide_posed_bodies_contacts": (C.c_int, [C.c_int, C.POINTER(_Scene), vp, vp,
                                                             i32, vp, vp, vp, C.c_size_t, vp]),
            "ogjk_build_hulls": (C.c_int, [vp, vp, vp, i32, vp, i64, vp, vp, vp, C.c_size_t,
                                           vp]),
            "ogjk_hull_moments": (C.c_int, [vp, vp, vp, vp, vp, i32, vp, vp, vp, vp]),
            "ogjk_center_points": (C.c_int, [vp, vp, i64, vp, vp]),
            "ogjk_broad_phase_workspace": (C.c_int, [i32, C.POINTER(_Grid),
                                                     C.POINTER(C.c_size_t)]),
            "ogjk_broad_phase": (C.c_int, [vp, i32, C.POINTER(_Grid), vp, i32, C.POINTER(i32),
                                           C.POINTER(i64), vp, C.c_size_t, vp]),
            "ogjk_expand_pairs_workspace": (C.c_int, [i32, C.POINTER(C.c_size_t)]),
            "ogjk_expand_pairs": (C.c_int, [vp, i32, vp, vp, i32, vp, i32, C.POINTER(i32),
                                            C.POINTER(i64), vp, C.c_size_t, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(status: int) -> None:
    if status != 0:
        msg = lib().ogjk_last_error().decode(errors="replace")
        raise OgjkError(f"ogjk status {status}: {msg}")


def _ptr(a):
    return None if a is None else a.ctypes.data


PIPELINE, FUSED, SPLIT = 0, 1, 2  # launch shapes of ogjk_set_config


def set_config(gjk_lanes: int = 16, epa_lanes: int = 32, fused: int = PIPELINE) -> None:
    """Launch shape (see ogjk_set_config in include/opengjk/ogjk.h): fused is
    PIPELINE (0), FUSED (1) or SPLIT (2, the plain-FP64 path in the
    reference's two-launch shape)."""
    _check(lib().ogjk_set_config(int(gjk_lanes), int(epa_lanes), int(fused)))


def set_buckets(max_verts, gjk_lanes, epa_lanes) -> None:
    """Size-bucketed launch shape (see ogjk_set_buckets in include/opengjk/ogjk.h)."""
    nb = len(gjk_lanes)
    mv = np.array(list(max_verts) + [0] * (nb - len(max_verts)), dtype=np.int32)
    gl = np.array(gjk_lanes, dtype=np.int32)
    el = np.array(epa_lanes, dtype=np.int32)
    _check(lib().ogjk_set_buckets(nb, _ptr(mv), _ptr(gl), _ptr(el)))


def set_block_waves(gjk_waves, epa_waves) -> None:
    """Waves per block of the GJK and EPA kernels per size bucket (see
    ogjk_set_block_waves in include/opengjk/ogjk.h)."""
    gw = np.array(gjk_waves, dtype=np.int32)
    ew = np.array(epa_waves, dtype=np.int32)
    _check(lib().ogjk_set_block_waves(len(gw), _ptr(gw), _ptr(ew)))


def device_count() -> int:
    return int(lib().ogjk_device_count())


# ------------------------------------------------------------------ host API

def polytopes(coords: np.ndarray, off, cnt, bodies) -> np.ndarray:
    """gkPolytope records (FP32 ABI) pointing into an AoS float32 pool."""
    assert coords.dtype == np.float32 and coords.flags.c_contiguous
    base = coords.ctypes.data
    bodies = np.asarray(bodies)
    p = np.zeros(len(bodies), dtype=POLYTOPE_DTYPE)
    p["numpoints"] = np.asarray(cnt)[bodies]
    p["coord"] = base + 12 * np.asarray(off, dtype=np.int64)[bodies].astype(np.uint64)
    return p


def collide_host(mode: int, batch, simplices=None, distances=None):
    """ogjk_collide_host over a generator Batch: pair i = (pa[i], pb[i])."""
    n = batch.n_pairs
    bd1 = polytopes(batch.coords, batch.off, batch.cnt, batch.pa)
    bd2 = polytopes(batch.coords, batch.off, batch.cnt, batch.pb)
    simp = np.zeros(n, SIMPLEX_DTYPE) if simplices is None else np.array(simplices, SIMPLEX_DTYPE)
    dist = np.zeros(n, np.float32) if distances is None else np.array(distances, np.float32)
    nrm = np.zeros((n, 3), np.float32)
    _check(lib().ogjk_collide_host(mode, n, _ptr(bd1), _ptr(bd2), _ptr(simp), _ptr(dist), _ptr(nrm)))
    return simp, dist, nrm


def collide_host_indexed(mode: int, coords, off, cnt, pairs, simplices=None, distances=None):
    pool = polytopes(coords, off, cnt, np.arange(len(cnt)))
    pr = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
    n = len(pr)
    simp = np.zeros(n, SIMPLEX_DTYPE) if simplices is None else np.array(simplices, SIMPLEX_DTYPE)
    dist = np.zeros(n, np.float32) if distances is None else np.array(distances, np.float32)
    nrm = np.zeros((n, 3), np.float32)
    _check(lib().ogjk_collide_host_indexed(mode, len(cnt), n, _ptr(pool), _ptr(pr), _ptr(simp),
                                           _ptr(dist), _ptr(nrm)))
    return simp, dist, nrm


# ------------------------------------------------------------ device scenes

class DeviceScene:
    """A batch resident in device memory as an SoA scene (torch tensors).
    f64: also keep the plain-FP64 comparison scene, SoA (True), one AoS pool
    ("aos") or both ("both")."""

    def __init__(self, batch, device="cuda", f64: bool = False, stream=None,
                 clusters: bool = False, cluster_min: int = 512):
        import torch
        self.torch = torch
        self.n_bodies = len(batch.cnt)
        self.n_verts = int(batch.coords.shape[0])
        dev = torch.device(device)
        aos = torch.from_numpy(np.ascontiguousarray(batch.coords).reshape(-1)).to(dev)
        self.off = torch.from_numpy(np.ascontiguousarray(batch.off, dtype=np.int64)).to(dev)
        self.cnt = torch.from_numpy(np.ascontiguousarray(batch.cnt, dtype=np.int32)).to(dev)
        nv = max(self.n_verts, 1)
        self.x = torch.empty(nv, dtype=torch.float32, device=dev)
        self.y = torch.empty(nv, dtype=torch.float32, device=dev)
        self.z = torch.empty(nv, dtype=torch.float32, device=dev)
        self.r1 = torch.empty(max(self.n_bodies, 1), dtype=torch.float32, device=dev)
        st = self._stream(stream)
        _check(lib().ogjk_pack_scene(aos.data_ptr(), self.off.data_ptr(), self.cnt.data_ptr(),
                                     self.n_bodies, self.n_verts, self.x.data_ptr(),
                                     self.y.data_ptr(), self.z.data_ptr(), self.r1.data_ptr(), st))
        self.f64 = self.f64_aos = None
        if f64 in ("aos", "both"):
            # plain FP64 in the reference's gkPolytope layout: one AoS pool
            self.xyz64 = torch.empty(3 * nv, dtype=torch.float64, device=dev)
            _check(lib().ogjk_pack_scene_f64_aos(aos.data_ptr(), self.n_verts,
                                                 self.xyz64.data_ptr(), st))
            self.f64_aos = _Scene64(None, None, None, self.off.data_ptr(), self.cnt.data_ptr(),
                                    self.n_bodies, self.xyz64.data_ptr())
        if f64 and f64 != "aos":
            self.x64 = torch.empty(nv, dtype=torch.float64, device=dev)
            self.y64 = torch.empty(nv, dtype=torch.float64, device=dev)
            self.z64 = torch.empty(nv, dtype=torch.float64, device=dev)
            _check(lib().ogjk_pack_scene_f64(aos.data_ptr(), self.n_verts, self.x64.data_ptr(),
                                             self.y64.data_ptr(), self.z64.data_ptr(), st))
            self.f64 = _Scene64(self.x64.data_ptr(), self.y64.data_ptr(), self.z64.data_ptr(),
                                self.off.data_ptr(), self.cnt.data_ptr(), self.n_bodies, None)
        torch.cuda.synchronize()
        del aos
        self.scene = _Scene(self.x.data_ptr(), self.y.data_ptr(), self.z.data_ptr(),
                            self.off.data_ptr(), self.cnt.data_ptr(), self.r1.data_ptr(),
                            self.n_bodies, None)
        self.clusters = None
        if clusters:
            self.build_clusters(cluster_min, stream)

    def build_clusters(self, min_verts: int = 512, stream=None):
        """Build the vertex clusters (ogjk_build_clusters) of the bodies of
        at least min_verts vertices and use them for those bodies."""
        t = self.torch
        dev = self.x.device
        slots = int(lib().ogjk_cluster_slots(self.n_verts, self.n_bodies))
        ns = max(slots, 32)
        # x, y, z and orig as views of one allocation (one large device
        # allocation took a third of the time of four), each 256-byte aligned
        stride = -(-ns // 64) * 64
        buf = t.empty(4 * stride, dtype=t.float32, device=dev)
        self.cx, self.cy, self.cz = (buf[k * stride:k * stride + ns] for k in range(3))
        self.corig = buf[3 * stride:3 * stride + ns].view(t.int32)
        self.cslot = t.empty(max(self.n_bodies, 1), dtype=t.int64, device=dev)
        self.cbounds = t.empty((ns // 32 + 1) * 8, dtype=t.float32, device=dev)
        nsup = int(lib().ogjk_cluster_supers(ns))
        self.csbounds = t.empty(max(nsup, 1) * 8, dtype=t.float32, device=dev)
        self.csuper = t.empty(max(self.n_bodies, 1), dtype=t.int64, device=dev)
        self.clusters = _Clusters(self.cx.data_ptr(), self.cy.data_ptr(), self.cz.data_ptr(),
                                  self.corig.data_ptr(), self.cslot.data_ptr(),
                                  self.cbounds.data_ptr(), ns, int(min_verts),
                                  self.csbounds.data_ptr(), self.csuper.data_ptr(), nsup, 0)
        _check(lib().ogjk_build_clusters(C.byref(self.scene), C.byref(self.clusters),
                                         self._stream(stream)))
        self.scene.clusters = C.pointer(self.clusters)

    def set_cluster_min(self, min_verts: int) -> None:
        """Culling threshold in vertices (needs build_clusters).  Bodies
        below the threshold the clusters were built with have none, so a
        lower value culls no further bodies."""
        if self.clusters is None:
            raise OgjkError("no clusters built")
        self.clusters.min_verts = int(min_verts)

    def collide_posed(self, mode: int, pairs, poses, out=None, f64_record: bool = False,
                      stream=None):
        """Scene bodies under per-pair poses (ogjk_collide_posed): pairs (n, 2)
        int32 and poses (n, 14) float32 device tensors; asynchronous."""
        n = int(pairs.shape[0])
        out = out or self.outputs(n, f64_record)
        need = int(lib().ogjk_workspace_bytes(n))
        ws = self.torch.empty(max(need, 1), dtype=self.torch.uint8, device=self.x.device)
        f = out["f64"].data_ptr() if out.get("f64") is not None else None
        _check(lib().ogjk_collide_posed(mode, C.byref(self.scene), pairs.data_ptr(),
                                        poses.data_ptr(), n, out["simplices"].data_ptr(),
                                        out["distances"].data_ptr(), out["normals"].data_ptr(), f,
                                        ws.data_ptr(), need, self._stream(stream)))
        out["_workspace"] = ws  # alive until the caller drops the outputs
        return out

    def collide_posed_bodies(self, mode: int, pairs, body_poses, out=None,
                             f64_record: bool = False, workspace=None, stream=None):
        """Scene bodies under one pose per body (ogjk_collide_posed_bodies):
        pairs (n, 2) int32 (a broad-phase list) and body_poses (n_bodies, 7)
        float32 device tensors; asynchronous.  workspace: a uint8 device
        tensor of at least ogjk_workspace_bytes(n) bytes, else allocated."""
        n = int(pairs.shape[0])
        out = out or self.outputs(n, f64_record)
        need = int(lib().ogjk_workspace_bytes(n))
        ws = workspace
        if ws is None or ws.numel() < need:
            ws = self.torch.empty(max(need, 1), dtype=self.torch.uint8, device=self.x.device)
        f = out["f64"].data_ptr() if out.get("f64") is not None else None
        _check(lib().ogjk_collide_posed_bodies(
            mode, C.byref(self.scene), pairs.data_ptr(), body_poses.data_ptr(), n,
            out["simplices"].data_ptr(), out["distances"].data_ptr(), out["normals"].data_ptr(), f,
            ws.data_ptr(), ws.numel(), self._stream(stream)))
        out["_workspace"] = ws
        return out

    def contacts(self, n: int, f64: bool = False):
        """Output buffers for the contact-record entry points."""
        t = self.torch
        dev = self.x.device
        return {"contacts": t.empty((max(n, 1), CONTACT_FLOATS), dtype=t.float32, device=dev),
                "f64": t.empty((max(n, 1), F64_RECORD), dtype=t.float64, device=dev)
                if f64 else None}

    def _ws(self, n, workspace):
        need = int(lib().ogjk_workspace_bytes(n))
        if workspace is None or workspace.numel() < need:
            workspace = self.torch.empty(max(need, 1), dtype=self.torch.uint8, device=self.x.device)
        return workspace

    def collide_contacts(self, mode: int, n_pairs: int, pairs=None, out=None,
                         f64_record: bool = False, workspace=None, stream=None):
        """ogjk_collide_scene_contacts: pairs (n, 2) int32 device tensor or
        None for (2i, 2i+1); out["contacts"] is (n, 10) float32."""
        out = out or self.contacts(n_pairs, f64_record)
        ws = self._ws(n_pairs, workspace)
        f = out["f64"].data_ptr() if out.get("f64") is not None else None
        _check(lib().ogjk_collide_scene_contacts(
            mode, C.byref(self.scene), None if pairs is None else pairs.data_ptr(), n_pairs,
            out["contacts"].data_ptr(), f, ws.data_ptr(), ws.numel(), self._stream(stream)))
        out["_workspace"] = ws
        return out

    def collide_posed_bodies_contacts(self, mode: int, pairs, body_poses, out=None,
                                      f64_record: bool = False, workspace=None, stream=None):
        """ogjk_collide_posed_bodies_contacts (see collide_posed_bodies)."""
        n = int(pairs.shape[0])
        out = out or self.contacts(n, f64_record)
        ws = self._ws(n, workspace)
        f = out["f64"].data_ptr() if out.get("f64") is not None else None
        _check(lib().ogjk_collide_posed_bodies_contacts(
            mode, C.byref(self.scene), pairs.data_ptr(), body_poses.data_ptr(), n,
            out["contacts"].data_ptr(), f, ws.data_ptr(), ws.numel(), self._stream(stream)))
        out["_workspace"] = ws
        return out

    def capture(self, mode: int, n_pairs: int, pairs=None, out=None, f64_record: bool = False):
        """Capture one launch over this scene as a HIP graph (see Graph)."""
        out = out or self.outputs(n_pairs, f64_record)
        return Graph(self, mode, n_pairs, pairs, out), out

    def drop_clusters(self) -> None:
        """Sweep every body without culling (the clusters stay built)."""
        self.scene.clusters = None

    def use_clusters(self) -> None:
        """Cull again through the clusters built by build_clusters."""
        if self.clusters is None:
            raise OgjkError("no clusters built")
        self.scene.clusters = C.pointer(self.clusters)

    def _stream(self, stream):
        if stream is None:
            return self.torch.cuda.current_stream().cuda_stream
        return stream.cuda_stream if hasattr(stream, "cuda_stream") else stream

    def outputs(self, n: int, f64: bool = False):
        t = self.torch
        dev = self.x.device
        return {
            "simplices": t.zeros((n, 27), dtype=t.int32, device=dev),  # 108 B records
            "distances": t.zeros(n, dtype=t.float32, device=dev),
            "normals": t.zeros((n, 3), dtype=t.float32, device=dev),
            "f64": t.zeros((n, F64_RECORD), dtype=t.float64, device=dev) if f64 else None,
        }

    def collide(self, mode: int, n_pairs: int, pairs=None, out=None, plain_f64=False,
                f64_record: bool = False, stream=None):
        """Launch on `stream` (default: torch's current stream); asynchronous.
        plain_f64: run the plain-FP64 comparison path over the SoA (True) or
        AoS ("aos") FP64 copy."""
        out = out or self.outputs(n_pairs, f64_record)
        pp = None if pairs is None else pairs.data_ptr()
        f = out["f64"].data_ptr() if out.get("f64") is not None else None
        st = self._stream(stream)
        if plain_f64:
            sc64 = self.f64_aos if plain_f64 == "aos" else self.f64
            if sc64 is None:
                raise OgjkError(f"scene built without its plain-FP64 copy ({plain_f64!r})")
            _check(lib().ogjk_collide_scene_f64(mode, C.byref(sc64), pp, n_pairs,
                                                out["simplices"].data_ptr(),
                                                out["distances"].data_ptr(),
                                                out["normals"].data_ptr(), f, st))
        else:
            _check(lib().ogjk_collide_scene(mode, C.byref(self.scene), pp, n_pairs,
                                            out["simplices"].data_ptr(), out["distances"].data_ptr(),
                                            out["normals"].data_ptr(), f, st))
        return out


class Server:
    """Serving loop over a resident scene (ogjk_server_*): requests of posed
    pairs are batched as they arrive and answered into numpy arrays."""

    def __init__(self, scene, max_batch_pairs: int = 65536, slots: int = 4, memory_budget: int = 0):
        if scene.scene.clusters:
            raise OgjkError("serving uses scenes without clusters")
        self._scene = scene
        cfg = _ServerConfig(max_batch_pairs, slots, memory_budget)
        h = C.c_void_p()
        _check(lib().ogjk_server_create(C.byref(scene.scene), C.byref(cfg), C.byref(h)))
        self._h = h
        self._live = {}

    def submit(self, pairs, poses, f64: bool = False) -> int:
        """pairs: (n, 2) int32 body ids; poses: (n, 14) float32 (quaternion
        w, x, y, z and translation of each body).  Returns a ticket."""
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        poses = np.ascontiguousarray(poses, dtype=np.float32).reshape(-1, 14)
        n = len(pairs)
        out = {"simplices": np.zeros(n, SIMPLEX_DTYPE), "distances": np.zeros(n, np.float32),
               "normals": np.zeros((n, 3), np.float32),
               "f64": np.zeros((n, F64_RECORD), np.float64) if f64 else None}
        t = C.c_int64()
        _check(lib().ogjk_server_submit(self._h, _ptr(pairs), _ptr(poses), n,
                                        _ptr(out["simplices"]), _ptr(out["distances"]),
                                        _ptr(out["normals"]), _ptr(out["f64"]), C.byref(t)))
        self._live[t.value] = (pairs, poses, out)
        return t.value

    def wait(self, ticket: int):
        """Blocks until answered; returns (outputs, latency_s, batch id)."""
        lat, batch = C.c_double(), C.c_int64()
        st = lib().ogjk_server_wait(self._h, int(ticket), C.byref(lat), C.byref(batch))
        _, _, out = self._live.pop(ticket)
        _check(st)
        return out, lat.value, batch.value

    def info(self) -> dict:
        sl, cap, la, pd = C.c_int32(), C.c_int32(), C.c_int64(), C.c_int64()
        _check(lib().ogjk_server_info(self._h, C.byref(sl), C.byref(cap), C.byref(la), C.byref(pd)))
        return {"slots": sl.value, "capacity": cap.value, "launches": la.value,
                "pairs_done": pd.value}

    def close(self) -> None:
        if self._h is not None and self._h.value:
            _check(lib().ogjk_server_destroy(self._h))
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """A captured scene launch (ogjk_graph_capture_scene); replay with launch()."""

    def __init__(self, scene, mode: int, n_pairs: int, pairs, out):
        self._scene, self._out, self._pairs = scene, out, pairs  # keep buffers alive
        h = C.c_void_p()
        f = out["f64"].data_ptr() if out.get("f64") is not None else None
        _check(lib().ogjk_graph_capture_scene(
            mode, C.byref(scene.scene), None if pairs is None else pairs.data_ptr(), n_pairs,
            out["simplices"].data_ptr(), out["distances"].data_ptr(), out["normals"].data_ptr(), f,
            C.byref(h)))
        self._h = h

    def launch(self, stream=None) -> None:
        _check(lib().ogjk_graph_launch(self._h, self._scene._stream(stream)))

    def close(self) -> None:
        if self._h is not None and self._h.value:
            _check(lib().ogjk_graph_destroy(self._h))
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def simplices_from_device(t) -> np.ndarray:
    """(n, 27) int32 device tensor of gkSimplex records -> numpy records."""
    return np.ascontiguousarray(t.cpu().numpy()).view(SIMPLEX_DTYPE).reshape(-1)


# ------------------------------------------------------------------ broad phase

class Grid:
    """Uniform grid of the broad phase (ogjk_grid): grid_size^3 cells of edge
    cell_size from -boundary; max_per_cell 0 lists every object of a cell."""

    def __init__(self, cell_size: float, boundary: float, grid_size: int, max_per_cell: int = 0):
        self.cell_size, self.boundary = float(cell_size), float(boundary)
        self.grid_size, self.max_per_cell = int(grid_size), int(max_per_cell)

    @classmethod
    def for_spheres(cls, max_radius: float, boundary: float, max_grid: int = 30,
                    max_per_cell: int = 0) -> "Grid":
        """The reference's sizing (visualization/integrate_final_gjk.cu:813-819):
        cells at least one largest diameter wide, at most max_grid per axis,
        computed in FP32."""
        f = np.float32
        cell = max(f(2.0) * f(max_radius), f(2.0) * f(boundary) / f(max_grid))
        g = max(1, min(int(f(2.0) * f(boundary) / f(cell)), max_grid))
        return cls(float(cell), boundary, g, max_per_cell)

    def _c(self) -> _Grid:
        return _Grid(self.cell_size, self.boundary, self.grid_size, self.max_per_cell)

    def __repr__(self):
        return (f"Grid(cell_size={self.cell_size}, boundary={self.boundary}, "
                f"grid_size={self.grid_size}, max_per_cell={self.max_per_cell})")


def _torch_stream(stream):
    import torch
    if stream is None:
        return torch.cuda.current_stream().cuda_stream
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else stream


class BroadPhase:
    """Device broad phase over n spheres with a fixed grid: owns its workspace
    and a pair buffer of max_pairs (default 50 per object, the reference's
    allocation, visualization/integrate_final_gjk.cu:730)."""

    def __init__(self, n_objects: int, grid: Grid, max_pairs: int | None = None, device=None):
        import torch
        self.torch = torch
        self.n, self.grid = int(n_objects), grid
        self.max_pairs = int(50 * self.n if max_pairs is None else max_pairs)
        dev = device or torch.device("cuda", torch.cuda.current_device())
        b = C.c_size_t()
        _check(lib().ogjk_broad_phase_workspace(self.n, C.byref(grid._c()), C.byref(b)))
        self.workspace = torch.empty(max(int(b.value), 1), dtype=torch.uint8, device=dev)
        self.pairs = torch.empty((max(self.max_pairs, 1), 2), dtype=torch.int32, device=dev)

    def __call__(self, spheres, stream=None):
        """spheres: (n, 4) float32 device tensor.  Returns (pairs (m, 2) view
        of the pair buffer, found); m = min(found, max_pairs)."""
        if spheres.shape != (self.n, 4) or spheres.dtype != self.torch.float32:
            raise OgjkError(f"spheres must be float32 of shape ({self.n}, 4)")
        if not spheres.is_contiguous():
            raise OgjkError("spheres must be contiguous")
        m, found = C.c_int32(), C.c_int64()
        _check(lib().ogjk_broad_phase(
            spheres.data_ptr(), self.n, C.byref(self.grid._c()), self.pairs.data_ptr(),
            self.max_pairs, C.byref(m), C.byref(found), self.workspace.data_ptr(),
            self.workspace.numel(), _torch_stream(stream)))
        return self.pairs[:m.value], found.value


def broad_phase(spheres, grid: Grid, max_pairs: int | None = None, stream=None):
    """One-shot device broad phase; returns (pairs (m, 2) int32 tensor, found)."""
    bp = BroadPhase(spheres.shape[0], grid, max_pairs, device=spheres.device)
    pairs, found = bp(spheres, stream)
    return pairs.clone(), found


def expand_pairs(body_pairs, sub_off, sub_cnt, max_sub_pairs: int | None = None, stream=None):
    """Body pairs -> sub-mesh pairs on the device (ogjk_expand_pairs).
    body_pairs (m, 2) int32, sub_off / sub_cnt int32 per body, all device
    tensors.  max_sub_pairs defaults to the exact count (computed on the host).
    Returns (sub pairs (k, 2) int32 tensor, found)."""
    import torch
    n = int(body_pairs.shape[0])
    nb = int(sub_cnt.shape[0])
    if max_sub_pairs is None:
        if n:
            bpi = body_pairs.long()
            ok = (bpi >= 0).all(1) & (bpi < nb).all(1)
            cnt = sub_cnt.long().clamp(min=0)
            bpc = bpi.clamp(0, max(nb - 1, 0))
            max_sub_pairs = int((cnt[bpc[:, 0]] * cnt[bpc[:, 1]] * ok).sum().item())
        else:
            max_sub_pairs = 0
    b = C.c_size_t()
    _check(lib().ogjk_expand_pairs_workspace(n, C.byref(b)))
    ws = torch.empty(max(int(b.value), 1), dtype=torch.uint8, device=body_pairs.device)
    out = torch.empty((max(max_sub_pairs, 1), 2), dtype=torch.int32, device=body_pairs.device)
    m, found = C.c_int32(), C.c_int64()
    _check(lib().ogjk_expand_pairs(
        body_pairs.data_ptr(), n, sub_off.data_ptr(), sub_cnt.data_ptr(), nb, out.data_ptr(),
        max_sub_pairs, C.byref(m), C.byref(found), ws.data_ptr(), ws.numel(),
        _torch_stream(stream)))
    return out[:m.value], found.value


class SceneStep:
    """Broad phase then narrow phase over a resident scene whose bodies move
    by one pose each (the reference's sim_broad_phase + sim_step without the
    integrator, visualization/integrate_final_gjk.cu:916-1035).  Body b's
    bounding sphere is centred at its pose's translation with radius radii[b]
    (which must bound |x| over the body's local vertices).  Buffers for
    max_pairs pairs are allocated once; contacts=True writes 40-byte contact
    records (out["contacts"]) instead of the gkSimplex outputs."""

    def __init__(self, scene: "DeviceScene", radii, grid: Grid, max_pairs: int | None = None,
                 f64_record: bool = False, contacts: bool = False):
        import torch
        self.torch, self.scene, self.grid = torch, scene, grid
        nb = scene.n_bodies
        dev = scene.x.device
        self.spheres = torch.zeros((nb, 4), dtype=torch.float32, device=dev)
        self.spheres[:, 3] = torch.as_tensor(np.asarray(radii, np.float32), device=dev)
        self.bp = BroadPhase(nb, grid, max_pairs, device=dev)
        self.use_contacts = contacts
        self.out = (scene.contacts(self.bp.max_pairs, f64_record) if contacts
                    else scene.outputs(self.bp.max_pairs, f64_record))
        self.workspace = torch.empty(max(int(lib().ogjk_workspace_bytes(self.bp.max_pairs)), 1),
                                     dtype=torch.uint8, device=dev)

    def __call__(self, body_poses, mode: int = GJK_EPA, stream=None):
        """body_poses (n_bodies, 7) float32 device tensor.  Returns (pairs,
        found, outputs): outputs hold the first len(pairs) results."""
        self.spheres[:, :3].copy_(body_poses[:, 4:7])
        pairs, found = self.bp(self.spheres, stream)
        if len(pairs):
            run = (self.scene.collide_posed_bodies_contacts if self.use_contacts
                   else self.scene.collide_posed_bodies)
            run(mode, pairs, body_poses, out=self.out, workspace=self.workspace, stream=stream)
        return pairs, found, self.out


# ---------------------------------------------------------------- mesh ingestion

def build_hulls(points, off, cnt, face_cap=None, stream=None):
    """Convex hulls of many clouds on the device (ogjk_build_hulls).  points
    (N, 3) float32, off int64 and cnt int32 per cloud, all device tensors;
    face_cap: faces per cloud (device int64 tensor), default 4 n + 16.
    Returns (faces (F, 3) int32, face_off (n + 1) int64, nfaces int32): cloud
    m's faces are faces[face_off[m]:face_off[m] + nfaces[m]] (-1: capacity too
    small, -2: non-finite input, 0: no hull)."""
    import torch
    n = int(cnt.numel())
    dev = points.device
    cap = (4 * cnt.long() + 16) if face_cap is None else face_cap.long()
    face_off = torch.zeros(n + 1, dtype=torch.int64, device=dev)
    if n:
        face_off[1:] = torch.cumsum(cap, 0)
    total = int(face_off[-1].item()) if n else 0
    faces = torch.empty((max(total, 1), 3), dtype=torch.int32, device=dev)
    nfaces = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    ws = torch.empty(max(12 * total, 1), dtype=torch.uint8, device=dev)
    _check(lib().ogjk_build_hulls(points.data_ptr(), off.data_ptr(), cnt.data_ptr(), n,
                                  face_off.data_ptr(), total, faces.data_ptr(), nfaces.data_ptr(),
                                  ws.data_ptr(), ws.numel(), _torch_stream(stream)))
    return faces, face_off, nfaces[:n]


def build_hulls_fit(points, off, cnt, stream=None):
    """build_hulls at 4 n + 16 faces per cloud, then once more at 16 n + 64 for
    the clouds that overflowed.  The reference's algorithm can grow its face
    list without bound on dense, nearly coplanar surface points (its 1e-7
    visibility threshold); a cloud still over 16 n + 64 keeps nfaces = -1, as
    the oracle reports it at that capacity, and its build time stays bounded."""
    import torch
    faces, face_off, nfaces = build_hulls(points, off, cnt, 4 * cnt.long() + 16, stream)
    over = nfaces == -1
    if bool(over.any()):
        cap = torch.where(over, 16 * cnt.long() + 64, 4 * cnt.long() + 16)
        faces, face_off, nfaces = build_hulls(points, off, cnt, cap, stream)
    return faces, face_off, nfaces


def hull_moments(points, off, faces, face_off, nfaces, stream=None):
    """Per cloud: FP64 moments (n, 7), centre of mass (n, 3) float32 and
    inertia factors (n, 3) float32 (ogjk_hull_moments)."""
    import torch
    n = int(nfaces.numel())
    dev = points.device
    mo = torch.empty((max(n, 1), 7), dtype=torch.float64, device=dev)
    com = torch.empty((max(n, 1), 3), dtype=torch.float32, device=dev)
    k = torch.empty((max(n, 1), 3), dtype=torch.float32, device=dev)
    _check(lib().ogjk_hull_moments(points.data_ptr(), off.data_ptr(), faces.data_ptr(),
                                   face_off.data_ptr(), nfaces.data_ptr(), n, mo.data_ptr(),
                                   com.data_ptr(), k.data_ptr(), _torch_stream(stream)))
    return mo[:n], com[:n], k[:n]


def center_points(points, cloud, com, stream=None) -> None:
    """points[i] -= com[cloud[i]] in place (ogjk_center_points)."""
    _check(lib().ogjk_center_points(points.data_ptr(), cloud.data_ptr(), int(points.shape[0]),
                                    com.data_ptr(), _torch_stream(stream)))


def ingest_hulls(points, off, cnt, stream=None):
    """The reference's ingestion of hull shapes for a batch of clouds
    (visualization/main_opengl.cpp:114-121, 313): centre each cloud on its
    hull's centre of mass, then the hull of the centred points (the faces it
    renders) and that hull's inertia factors.  Returns a dict of device
    tensors: points (centred copy), faces, face_off, nfaces, com, inertia_k,
    volume."""
    import torch
    faces, face_off, nfaces = build_hulls_fit(points, off, cnt, stream)
    _, com, _ = hull_moments(points, off, faces, face_off, nfaces, stream)
    centred = points.clone()
    n = int(cnt.numel())
    cloud = torch.repeat_interleave(torch.arange(n, dtype=torch.int32, device=points.device),
                                    cnt.long())
    center_points(centred, cloud, com, stream)
    faces2, face_off2, nfaces2 = build_hulls_fit(centred, off, cnt, stream)
    mo2, _, k2 = hull_moments(centred, off, faces2, face_off2, nfaces2, stream)
    return {"points": centred, "faces": faces2, "face_off": face_off2, "nfaces": nfaces2,
            "com": com, "inertia_k": k2, "volume": mo2[:, 0]}
