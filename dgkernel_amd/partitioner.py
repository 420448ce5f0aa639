"""Mesh partitioning on the element graph.

Mirrors the reference's ``mesh/partitioner`` package
(mesh/partitioner/mesh_partitioner.go): ``PartitionConfig`` and
``default_partition_config`` (:14-31), ``MeshPartitioner`` with its
replaceable compute and communication cost models (:33-85), the weighted
element graph (``build_graph``, buildMetisGraph :145-184), ``partition``
(:87-142), the partition analysis (:186-283), and the boundary-face and
element queries (:296-335). The METIS call is replaced by libdgk's own
multilevel k-way partitioner (``dgk_partition_kway``, host code in the same
library; no device is used). A partition is turned into the partition-major
element order the solver runs on with ``mesh.order_by_partition``.
"""
from __future__ import annotations

import ctypes as C
import logging
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .mesh import TetMesh, connect

log = logging.getLogger(__name__)

OBJECTIVES = {"cut": 0, "vol": 1}

# Relative cost of one element per geometry (mesh_partitioner.go:51-62).
BASE_COST = {"tet": 1, "hex": 8, "prism": 6, "pyramid": 5}


@dataclass
class PartitionConfig:
    """mesh_partitioner.go:14-20."""
    num_partitions: int
    imbalance_factor: float = 1.05
    use_edge_weights: bool = True
    use_vertex_weights: bool = True
    objective: str = "vol"


def default_partition_config(nparts: int) -> PartitionConfig:
    """mesh_partitioner.go:22-31: 5 % imbalance, both weights, volume objective."""
    return PartitionConfig(nparts, 1.05, True, True, "vol")


@dataclass
class PartitionStats:
    """mesh_partitioner.go:286-293."""
    id: int
    num_elements: int = 0
    compute_load: int = 0
    element_types: dict = field(default_factory=dict)
    num_neighbors: dict = field(default_factory=dict)  # neighbour part -> shared faces


def _compute_cost(elem_type: str, num_vertices: int) -> int:
    return BASE_COST.get(elem_type, 0)


def _comm_cost(face_vertices: int, is_boundary: bool) -> int:
    """mesh_partitioner.go:65-82: nothing across a physical boundary, else the
    face's vertex count."""
    return 0 if is_boundary else int(face_vertices)


class MeshPartitioner:
    """Partitions a tetrahedral mesh's elements (``mesh.EToP``)."""

    def __init__(self, mesh: TetMesh, config: PartitionConfig, seed: int = 0):
        self.mesh, self.config, self.seed = mesh, config, seed
        self.compute_cost_model = _compute_cost
        self.comm_cost_model = _comm_cost
        self.analysis: dict | None = None
        self._conn = None

    # -- graph -------------------------------------------------------------

    def _connectivity(self):
        if self._conn is None:
            K = self.mesh.K
            if K == 0:
                e = np.zeros((0, 4), dtype=np.int64)
                self._conn = (e, e, np.zeros((0, 4), dtype=np.int64))
            else:
                EToE, EToF = connect(self.mesh)
                own = np.repeat(np.arange(K, dtype=np.int64), 4).reshape(K, 4)
                key = np.minimum(4 * own + np.arange(4), 4 * EToE + EToF)
                _, fid = np.unique(key.ravel(), return_inverse=True)
                self._conn = (EToE, EToF, fid.reshape(K, 4).astype(np.int64))
        return self._conn

    def face_ids(self) -> np.ndarray:
        """[K,4] id of each element face in the mesh's face list (a face shared
        by two elements has one id)."""
        return self._connectivity()[2]

    def build_graph(self):
        """buildMetisGraph (mesh_partitioner.go:145-184): CSR adjacency over
        face neighbours (no self or boundary entries), vertex weights from the
        compute cost model and edge weights from the communication cost model;
        a weight array is None when its use is switched off."""
        K = self.mesh.K
        EToE, _, _ = self._connectivity()
        own = np.repeat(np.arange(K, dtype=np.int64), 4).reshape(K, 4)
        keep = EToE != own
        xadj = np.zeros(K + 1, dtype=np.int32)
        np.cumsum(keep.sum(axis=1), out=xadj[1:])
        adjncy = EToE[keep].astype(np.int32)
        vwgt = adjwgt = None
        if self.config.use_vertex_weights:
            vwgt = np.full(K, self.compute_cost_model("tet", 4), dtype=np.int32)
        if self.config.use_edge_weights:
            adjwgt = np.full(adjncy.size, self.comm_cost_model(3, False), dtype=np.int32)
        return xadj, adjncy, vwgt, adjwgt

    # -- partitioning --------------------------------------------------------

    def partition(self) -> int:
        """Partition (mesh_partitioner.go:87-142): sets ``mesh.EToP`` and
        ``self.analysis``; returns the objective value (edge cut or volume)."""
        cfg = self.config
        if cfg.objective not in OBJECTIVES:
            raise ValueError(f"objective must be 'cut' or 'vol', got {cfg.objective!r}")
        if cfg.num_partitions < 1:
            raise ValueError("num_partitions must be >= 1")
        log.info("Partitioning mesh with %d elements into %d parts", self.mesh.K, cfg.num_partitions)
        xadj, adjncy, vwgt, adjwgt = self.build_graph()
        part, objval = partition_graph(xadj, adjncy, vwgt, adjwgt, cfg.num_partitions,
                                       cfg.imbalance_factor, cfg.objective, self.seed)
        self.mesh.EToP = part.astype(np.int64)
        self.analysis = self.analyze(objval)
        return objval

    def analyze(self, objval: int) -> dict:
        """analyzePartition (mesh_partitioner.go:186-283): cut faces,
        communication volume (summed face costs), load imbalance, per-partition
        statistics and the faces between each pair of partitions."""
        etop = np.asarray(self.mesh.EToP, dtype=np.int64)
        k = int(self.config.num_partitions)
        K = self.mesh.K
        EToE, _, fid = self._connectivity()
        cost = self.compute_cost_model("tet", 4)
        counts = np.bincount(etop, minlength=k)[:k] if K else np.zeros(k, dtype=np.int64)
        stats = [PartitionStats(p, int(counts[p]), int(counts[p]) * cost,
                                {"tet": int(counts[p])} if counts[p] else {}) for p in range(k)]
        own = np.repeat(np.arange(K, dtype=np.int64), 4).reshape(K, 4)
        cut = (EToE > own) & (etop[EToE] != etop[own]) if K else np.zeros((0, 4), bool)
        e, f = np.nonzero(cut)
        p1, p2 = etop[e], etop[EToE[e, f]]
        lo, hi = np.minimum(p1, p2), np.maximum(p1, p2)
        interfaces: dict = {}
        for a, b, face in zip(lo.tolist(), hi.tolist(), fid[e, f].tolist()):
            interfaces.setdefault((a, b), []).append(face)
        for (a, b), faces in interfaces.items():
            stats[a].num_neighbors[b] = stats[a].num_neighbors.get(b, 0) + len(faces)
            stats[b].num_neighbors[a] = stats[b].num_neighbors.get(a, 0) + len(faces)
        loads = np.array([s.compute_load for s in stats], dtype=np.float64)
        avg = float(loads.mean()) if k else 0.0
        out = {
            "objective_value": int(objval),
            "cut_edges": int(e.size),
            "comm_volume": int(e.size) * self.comm_cost_model(3, False),
            "imbalance": float(loads.max() / avg - 1.0) if avg > 0 else 0.0,
            "load_min": int(loads.min()) if k else 0,
            "load_max": int(loads.max()) if k else 0,
            "load_avg": avg,
            "partitions": stats,
            "interfaces": interfaces,
        }
        log.info("Partition analysis: objective %d, cut edges %d, communication volume %d, "
                 "load imbalance %.2f%%, load range [%d, %d], avg %.1f", out["objective_value"],
                 out["cut_edges"], out["comm_volume"], 100 * out["imbalance"], out["load_min"],
                 out["load_max"], avg)
        return out

    def partition_boundary_faces(self) -> dict:
        """GetPartitionBoundaryFaces (mesh_partitioner.go:296-323): for each
        partition, the ids of its element faces on a physical or partition
        boundary."""
        etop = np.asarray(self.mesh.EToP, dtype=np.int64)
        K = self.mesh.K
        EToE, _, fid = self._connectivity()
        own = np.repeat(np.arange(K, dtype=np.int64), 4).reshape(K, 4)
        bnd = (EToE == own) | (etop[EToE] != etop[own])
        out: dict = {}
        for e, f in zip(*np.nonzero(bnd)):
            out.setdefault(int(etop[e]), []).append(int(fid[e, f]))
        return out

    def partition_elements(self, part: int) -> np.ndarray:
        """GetPartitionElements (mesh_partitioner.go:325-334)."""
        return np.nonzero(np.asarray(self.mesh.EToP) == part)[0]

    def export_partitioned_mesh(self, filename: str) -> None:
        """ExportPartitionedMesh (mesh_partitioner.go:337-341) is not
        implemented in the reference either; ``mesh.write_gambit_neutral``
        writes a partitioned mesh (groups ``pN``)."""
        raise NotImplementedError("export_partitioned_mesh: use mesh.write_gambit_neutral")


def partition_graph(xadj, adjncy, vwgt, adjwgt, nparts: int, ubfactor: float = 1.05,
                    objective: str = "cut", seed: int = 0):
    """dgk_partition_kway on a CSR graph; returns (part int32 [nv], objective value)."""
    xadj = np.ascontiguousarray(xadj, dtype=np.int64)
    adjncy = np.ascontiguousarray(adjncy, dtype=np.int64)
    nv = xadj.size - 1
    if nv < 0:
        raise ValueError("xadj must hold at least one entry")
    vw = None if vwgt is None else np.ascontiguousarray(vwgt, dtype=np.int64)
    ew = None if adjwgt is None else np.ascontiguousarray(adjwgt, dtype=np.int64)
    if vw is not None and vw.size != nv:
        raise ValueError("vwgt must have one weight per vertex")
    if ew is not None and ew.size != adjncy.size:
        raise ValueError("adjwgt must have one weight per adjacency entry")
    if adjncy.size < (xadj[-1] if nv >= 0 else 0):
        raise ValueError("adjncy is shorter than xadj[-1]")
    part = np.zeros(max(nv, 1), dtype=np.int32)
    obj = C.c_int64()
    p64 = C.POINTER(C.c_int64)
    def ptr(a):
        return None if a is None else a.ctypes.data_as(p64)
    _lib.call("dgk_partition_kway", nv, ptr(xadj), ptr(adjncy) if adjncy.size else None, ptr(vw),
              ptr(ew), int(nparts), float(ubfactor), OBJECTIVES[objective], int(seed) & (2 ** 64 - 1),
              part.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(obj))
    return part[:nv], int(obj.value)


def kway_partition(mesh: TetMesh, nparts: int, objective: str = "cut", ubfactor: float = 1.05,
                   seed: int = 0) -> np.ndarray:
    """EToP of a k-way partition of ``mesh`` (unit element weights, face
    weights 3), the graph-based counterpart of ``mesh.rcb_partition``."""
    mp = MeshPartitioner(TetMesh(mesh.VX, mesh.VY, mesh.VZ, mesh.EToV, mesh.EToVc,
                                 period=mesh.period, info=mesh.info),
                         PartitionConfig(nparts, ubfactor, True, True, objective), seed)
    mp.partition()
    return mp.mesh.EToP
