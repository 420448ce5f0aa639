"""Host-side mesh setup for the tetrahedral DG hot path.

* Structured cube meshes: every hex is split into the reference's six
  positively oriented tetrahedra around the main diagonal
  (utils/mesh_test_helpers.go:346-393, ``createCubeMesh``), optionally periodic
  and with seeded vertex jitter.
* The Gambit neutral reader for the reference's ``cube-partitioned.neu``
  fixture (mesh/readers/gambit_reader.go:14-338; partitions come from groups
  named ``pN``).
* Face connectivity: EToE/EToF by sorted face keys (the tiConnect3D method,
  element/library/gonudg/ticonnect3d.go:24-122) and VmapP. The reference
  matches face nodes by coordinates (buildmaps3d.go:46-153); here the match is
  done once on the reference element, by barycentric coordinates on the shared
  face, giving an orientation code per face that the device kernels use
  directly. The resulting VmapP is compared bit-exactly with the coordinate
  match in the tests.
* FaceConnector pick/place lists (utils/face_connector.go:43-187), vectorised
  but in the reference's exact append order.
* Seeded randomness uses splitmix64 (Steele, Lea & Flood, 2014) as a counter
  based generator so that host data is identical wherever it is generated.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from itertools import permutations

import numpy as np

from .nudg import RefTet

# Local vertices of the four faces, NUDG numbering (ticonnect3d.go:40-45;
# consistent with Fmask: t=-1, s=-1, r+s+t=-1, r=-1).
FACE_VERTS = np.array([[0, 1, 2], [0, 1, 3], [1, 2, 3], [0, 2, 3]], dtype=np.int64)
PERMS = list(permutations(range(3)))  # 6 orderings of a face's three vertices

# Hex corners (utils/mesh_test_helpers.go:82-91) and the six tets
# (utils/mesh_test_helpers.go:361-372), all with det[vb-va, vc-va, vd-va] > 0.
_CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                     [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.int64)
_KUHN6 = np.array([[0, 1, 2, 6], [0, 2, 3, 6], [0, 4, 5, 6],
                   [0, 7, 4, 6], [0, 5, 1, 6], [0, 3, 7, 6]], dtype=np.int64)

_M64 = (1 << 64) - 1


def splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser applied to a uint64 array (wrapping arithmetic)."""
    z = (x + np.uint64(0x9E3779B97F4A7C15)).astype(np.uint64)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, n: int, lo: float = 0.0, hi: float = 1.0,
            start: int = 0) -> np.ndarray:
    """n doubles in [lo, hi) from counter (seed, stream, index)."""
    key = np.uint64(splitmix64(np.array([(seed * 0x100000001B3 + stream) & _M64],
                                        dtype=np.uint64))[0])
    with np.errstate(over="ignore"):
        idx = np.arange(start, start + n, dtype=np.uint64) * np.uint64(0xD1B54A32D192ED03)
        z = splitmix64(idx ^ key)
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


@dataclass
class TetMesh:
    """A tetrahedral mesh: vertex coordinates and element-to-vertex lists.

    ``EToV`` indexes ``VX/VY/VZ`` (coordinates used for the affine map).
    ``EToVc`` is the numbering used to find shared faces; for a periodic mesh
    opposite boundary vertices share an id there.
    """

    VX: np.ndarray
    VY: np.ndarray
    VZ: np.ndarray
    EToV: np.ndarray  # [K,4] int64
    EToVc: np.ndarray  # [K,4] int64
    EToP: np.ndarray | None = None  # [K] partition of each element
    period: tuple = (0.0, 0.0, 0.0)
    info: dict = field(default_factory=dict)

    @property
    def K(self) -> int:
        return int(self.EToV.shape[0])


def cube_mesh(nx: int, ny: int, nz: int, periodic: bool = False, jitter: float = 0.0,
              seed: int = 0, kmax: int | None = None, L: float = 1.0) -> TetMesh:
    """Structured [0,L]^3 mesh of nx*ny*nz hexes, 6 tets each.

    Element order is x slowest, z fastest, 6 tets per hex
    (k = ((ix*ny + iy)*nz + iz)*6 + t), so contiguous element ranges are
    x-slabs. ``kmax`` keeps only the first kmax tets in that order.
    ``jitter`` moves every vertex by U(-jitter*h, jitter*h) per coordinate
    (seeded), except along the normal of a boundary plane it lies on; for a
    periodic mesh the jitter is periodic (opposite vertices move together).
    """
    hx, hy, hz = L / nx, L / ny, L / nz
    i = np.arange(nx + 1)
    j = np.arange(ny + 1)
    k = np.arange(nz + 1)
    I, J, Kk = np.meshgrid(i, j, k, indexing="ij")
    I, J, Kk = I.ravel(), J.ravel(), Kk.ravel()
    vid = lambda a, b, c: (a * (ny + 1) + b) * (nz + 1) + c  # noqa: E731
    VX, VY, VZ = I * hx, J * hy, Kk * hz
    VX, VY, VZ = VX.astype(np.float64), VY.astype(np.float64), VZ.astype(np.float64)
    if jitter > 0.0:
        if periodic:
            key = (I % nx) * ny * nz + (J % ny) * nz + (Kk % nz)
        else:
            key = vid(I, J, Kk)
        nkey = int(key.max()) + 1
        jx = uniform(seed, 101, nkey, -jitter * hx, jitter * hx)[key]
        jy = uniform(seed, 102, nkey, -jitter * hy, jitter * hy)[key]
        jz = uniform(seed, 103, nkey, -jitter * hz, jitter * hz)[key]
        if not periodic:
            jx = np.where((I == 0) | (I == nx), 0.0, jx)
            jy = np.where((J == 0) | (J == ny), 0.0, jy)
            jz = np.where((Kk == 0) | (Kk == nz), 0.0, jz)
        VX, VY, VZ = VX + jx, VY + jy, VZ + jz
    hi, hj, hk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    hi, hj, hk = hi.ravel(), hj.ravel(), hk.ravel()
    nh = hi.size
    if kmax is not None:
        nh = min(nh, -(-kmax // 6))
        hi, hj, hk = hi[:nh], hj[:nh], hk[:nh]
    c = _CORNERS
    corner = np.stack([vid(hi + c[q, 0], hj + c[q, 1], hk + c[q, 2]) for q in range(8)], axis=1)
    EToV = corner[:, _KUHN6].reshape(-1, 4)
    if periodic:
        pid = lambda a, b, cc: ((a % nx) * ny + (b % ny)) * nz + (cc % nz)  # noqa: E731
        cornerc = np.stack([pid(hi + c[q, 0], hj + c[q, 1], hk + c[q, 2]) for q in range(8)], axis=1)
        EToVc = cornerc[:, _KUHN6].reshape(-1, 4)
    else:
        EToVc = EToV
    if kmax is not None:
        EToV, EToVc = EToV[:kmax], EToVc[:kmax]
    per = (L, L, L) if periodic else (0.0, 0.0, 0.0)
    return TetMesh(VX, VY, VZ, EToV.astype(np.int64), EToVc.astype(np.int64), period=per,
                   info={"nx": nx, "ny": ny, "nz": nz, "jitter": jitter, "seed": seed})


def read_gambit_neutral(path: str) -> TetMesh:
    """Gambit .neu reader for tetrahedral meshes; mesh/readers/gambit_reader.go:14-338.

    Vertices and elements are numbered in file order; an element group named
    ``p<N>`` assigns partition N to its elements, others leave -1
    (gambit_reader.go:140-218).
    """
    with open(path) as fh:
        lines = fh.read().splitlines()
    n = len(lines)
    pos = 0
    numnp = nelem = ngrps = 0
    while pos < n:
        if "NUMNP" in lines[pos] and "NELEM" in lines[pos]:
            v = lines[pos + 1].split()
            numnp, nelem, ngrps = int(v[0]), int(v[1]), int(v[2])
            pos += 2
            break
        pos += 1
    verts = np.zeros((numnp, 3))
    node_map: dict[int, int] = {}
    elems = []
    elem_map: dict[int, int] = {}
    etop = None
    while pos < n:
        line = lines[pos].strip()
        if "NODAL COORDINATES" in line:
            for q in range(numnp):
                f = lines[pos + 1 + q].split()
                node_map[int(f[0])] = q
                verts[q] = [float(f[1]), float(f[2]), float(f[3])]
            pos += 1 + numnp
            continue
        if "ELEMENTS/CELLS" in line:
            for q in range(nelem):
                f = lines[pos + 1 + q].split()
                if int(f[1]) != 6:
                    raise ValueError("only tetrahedra (Gambit type 6) are supported")
                elem_map[int(f[0])] = q
                elems.append([node_map[int(x)] for x in f[3:3 + int(f[2])]])
            pos += 1 + nelem
            continue
        if "ELEMENT GROUP" in line:
            if ngrps > 1 and etop is None:
                etop = np.full(nelem, -1, dtype=np.int64)
            f = lines[pos + 1].split()
            count = int(f[f.index("ELEMENTS:") + 1])
            name = lines[pos + 2].strip()
            part = -1
            if name.startswith("p") and name[1:].isdigit():
                part = int(name[1:])
            pos += 4  # group line, name, flags
            got = 0
            while got < count and pos < n:
                line2 = lines[pos].strip()
                if line2 == "ENDOFSECTION":
                    break
                for tok in line2.split():
                    eid = int(tok)
                    if eid > 0 and eid in elem_map and part >= 0 and etop is not None:
                        etop[elem_map[eid]] = part
                    got += 1
                pos += 1
            continue
        pos += 1
    EToV = np.array(elems, dtype=np.int64)
    return TetMesh(verts[:, 0].copy(), verts[:, 1].copy(), verts[:, 2].copy(), EToV, EToV,
                   EToP=etop, info={"file": path})


def write_gambit_neutral(mesh: TetMesh, path: str, title: str = "dgkernel_amd tetrahedral mesh") -> None:
    """Writes ``mesh`` as a Gambit neutral file that read_gambit_neutral (and the
    reference's mesh/readers/gambit_reader.go:14-338) reads back to the same
    vertices, elements and partitions.

    Vertices and elements are numbered 1.. in array order; coordinates carry 17
    significant digits, so they round-trip exactly. As in Gmsh's output, a
    material group holds every element, and each partition follows as an
    element group named ``p<N>``. A periodic identification (``EToVc``) has no
    place in this format: use save_mesh for periodic meshes.
    """
    if mesh.EToVc is not mesh.EToV and not np.array_equal(mesh.EToVc, mesh.EToV):
        raise ValueError("write_gambit_neutral: periodic meshes need save_mesh")
    K, nv = mesh.K, int(mesh.VX.size)
    parts = [] if mesh.EToP is None else [int(q) for q in np.unique(mesh.EToP) if q >= 0]
    out = ["        CONTROL INFO 2.4.6", "** GAMBIT NEUTRAL FILE", title,
           "PROGRAM:            dgkernel_amd     VERSION:  1.0", "",
           "     NUMNP     NELEM     NGRPS    NBSETS     NDFCD     NDFVL",
           f"{nv:10d}{K:10d}{len(parts) + 1:10d}{0:10d}{3:10d}{3:10d}", "ENDOFSECTION",
           "   NODAL COORDINATES 2.4.6"]
    for q in range(nv):
        out.append(f"{q + 1:10d} {mesh.VX[q]:24.16e} {mesh.VY[q]:24.16e} {mesh.VZ[q]:24.16e}")
    out += ["ENDOFSECTION", "      ELEMENTS/CELLS 2.4.6"]
    for k in range(K):
        v = mesh.EToV[k] + 1
        out.append(f"{k + 1:8d} {6:2d} {4:2d} {v[0]:8d}{v[1]:8d}{v[2]:8d}{v[3]:8d}")
    out.append("ENDOFSECTION")
    groups = [("Material group 1", np.arange(1, K + 1))]
    groups += [("p" + str(part), np.nonzero(mesh.EToP == part)[0] + 1) for part in parts]
    for gi, (name, ids) in enumerate(groups):
        out += ["       ELEMENT GROUP 2.4.6",
                f"GROUP: {gi + 1:10d} ELEMENTS: {ids.size:10d} MATERIAL: {0:10d} NFLAGS: {1:10d}",
                f"{name:>32s}", "       0"]
        for a in range(0, ids.size, 10):
            out.append("".join(f"{int(e):8d}" for e in ids[a:a + 10]))
        out.append("ENDOFSECTION")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def save_mesh(mesh: TetMesh, path: str) -> None:
    """Stores a mesh, periodic identification and partitions included, as a
    NumPy .npz (no pickled objects), so that separate runs -- CPU oracle and
    GPU -- read identical inputs. ``info`` must be JSON-serialisable."""
    import json
    np.savez(path, VX=mesh.VX, VY=mesh.VY, VZ=mesh.VZ, EToV=mesh.EToV, EToVc=mesh.EToVc,
             EToP=np.zeros(0, np.int64) if mesh.EToP is None else mesh.EToP,
             has_etop=np.array(mesh.EToP is not None), period=np.array(mesh.period, np.float64),
             info=np.array(json.dumps(mesh.info)))


def load_mesh(path: str) -> TetMesh:
    """Reads a mesh written by save_mesh (numpy.load with pickles refused)."""
    import json
    with np.load(path, allow_pickle=False) as z:
        etop = z["EToP"].astype(np.int64) if bool(z["has_etop"]) else None
        return TetMesh(z["VX"].copy(), z["VY"].copy(), z["VZ"].copy(), z["EToV"].astype(np.int64),
                       z["EToVc"].astype(np.int64), EToP=etop,
                       period=tuple(float(x) for x in z["period"]), info=json.loads(str(z["info"])))


def connect(mesh: TetMesh):
    """EToE, EToF [K,4] by matching sorted face vertex keys (ticonnect3d.go:24-122).

    Faces that appear once keep EToE = self, EToF = self.
    """
    K = mesh.K
    fv = np.sort(mesh.EToVc[:, FACE_VERTS], axis=2).reshape(-1, 3)  # [4K,3]
    nn = int(mesh.EToVc.max()) + 1
    key = (fv[:, 0] * nn + fv[:, 1]) * nn + fv[:, 2]
    order = np.argsort(key, kind="stable")
    ks = key[order]
    EToE = np.repeat(np.arange(K, dtype=np.int64), 4).reshape(K, 4).copy()
    EToF = np.tile(np.arange(4, dtype=np.int64), (K, 1))
    same = np.nonzero(ks[:-1] == ks[1:])[0]
    a, b = order[same], order[same + 1]
    EToE.ravel()[a] = b // 4
    EToF.ravel()[a] = b % 4
    EToE.ravel()[b] = a // 4
    EToF.ravel()[b] = a % 4
    return EToE, EToF


def face_node_tables(ref: RefTet) -> np.ndarray:
    """NODEP[f, code, fp] = volume node of the neighbour matching face node fp.

    code = 6*f2 + sigma: the neighbour's face f2, and sigma the ordering such
    that own face vertex slot q sits at neighbour face slot PERMS[sigma][q].
    Face nodes are matched by their barycentric coordinates on the face
    (exact lattice match after rounding), which is the coordinate match of
    buildmaps3d.go:118-150 carried out on the reference element.
    """
    Nfp = ref.Nfp
    lam = np.stack([-(1.0 + ref.r + ref.s + ref.t) / 2.0, (1.0 + ref.r) / 2.0,
                    (1.0 + ref.s) / 2.0, (1.0 + ref.t) / 2.0], axis=1)  # [Np,4]
    table = np.full((4, 24, Nfp), -1, dtype=np.int32)
    for f in range(4):
        bf = np.round(lam[ref.Fmask[f]][:, FACE_VERTS[f]], 9)  # [Nfp,3]
        for f2 in range(4):
            b2 = np.round(lam[ref.Fmask[f2]][:, FACE_VERTS[f2]], 9)
            for sg, perm in enumerate(PERMS):
                # own slot q corresponds to neighbour slot perm[q]
                want = np.zeros_like(bf)
                want[:, list(perm)] = bf
                for fp in range(Nfp):
                    hit = np.nonzero(np.all(np.abs(b2 - want[fp]) < 1e-7, axis=1))[0]
                    if hit.size != 1:
                        raise ValueError("face node sets are not symmetric")
                    table[f, 6 * f2 + sg, fp] = ref.Fmask[f2][hit[0]]
    return table


def face_codes(mesh: TetMesh, EToE: np.ndarray, EToF: np.ndarray) -> np.ndarray:
    """Orientation code 6*f2 + sigma for every (k, f); boundary faces map to themselves."""
    K = mesh.K
    own = mesh.EToVc[:, FACE_VERTS]  # [K,4,3] global ids in own slot order
    nbv = mesh.EToVc[EToE.ravel()].reshape(K, 4, 4)  # neighbour's four vertices
    nbface = np.take_along_axis(nbv, FACE_VERTS[EToF], axis=2)  # [K,4,3] neighbour slot order
    code = np.full((K, 4), -1, dtype=np.int64)
    for sg, perm in enumerate(PERMS):
        ok = np.all(own == nbface[..., list(perm)], axis=2)
        code = np.where(ok & (code < 0), 6 * EToF + sg, code)
    if (code < 0).any():
        raise ValueError("non-conforming face found while building orientation codes")
    return code.astype(np.int32)


def vmap_pm(ref: RefTet, EToE: np.ndarray, codes: np.ndarray):
    """VmapM, VmapP [K*4*Nfp] from the orientation codes (buildmaps3d.go:5-43 layout)."""
    K = EToE.shape[0]
    Np, Nfp = ref.Np, ref.Nfp
    table = face_node_tables(ref)
    base = (np.arange(K, dtype=np.int64) * Np)[:, None, None]
    vmapM = (base + ref.Fmask[None, :, :]).reshape(-1)
    f = np.arange(4)[None, :]
    nodes = table[np.broadcast_to(f, codes.shape), codes]  # [K,4,Nfp]
    vmapP = (EToE[:, :, None].astype(np.int64) * Np + nodes).reshape(-1)
    return vmapM, vmapP


@dataclass
class FaceConnector:
    """Pick/place index lists per partition pair; utils/face_connector.go:8-40."""

    num_partitions: int
    K: int
    Nfaces: int
    Nfp: int
    Np: int
    elems_per_partition: np.ndarray
    local_to_global: list
    global_to_local: np.ndarray
    pick: dict  # (src, dst) -> int64 array of local source nodes
    place: dict  # (dst, src) -> int64 array of P-buffer slots

    def get_pick_indices(self, src: int, dst: int):
        """GetPickIndices; face_connector.go:190-196 (None when out of range)."""
        if not (0 <= src < self.num_partitions and 0 <= dst < self.num_partitions):
            return None
        return self.pick[(src, dst)]

    def get_place_indices(self, dst: int, src: int):
        """GetPlaceIndices; face_connector.go:199-205."""
        if not (0 <= src < self.num_partitions and 0 <= dst < self.num_partitions):
            return None
        return self.place[(dst, src)]

    def verify(self) -> None:
        """Verify; face_connector.go:208-253. Raises ValueError on the first violation."""
        NP = self.num_partitions
        for p in range(NP):
            mx = int(self.elems_per_partition[p]) * self.Np
            for q in range(NP):
                idx = self.pick[(p, q)]
                if idx.size and (idx.min() < 0 or idx.max() >= mx):
                    raise ValueError(f"invalid pick index for partition {p} (max {mx - 1})")
        for p in range(NP):
            for q in range(NP):
                if self.pick[(p, q)].size != self.place[(q, p)].size:
                    raise ValueError(f"length mismatch: pick[{p}][{q}] vs place[{q}][{p}]")
        total = sum(v.size for v in self.pick.values())
        want = int(self.elems_per_partition.sum()) * self.Nfaces * self.Nfp
        if total != want:
            raise ValueError(f"conservation error: total picks {total} != total face points {want}")


def new_face_connector(K: int, Nfaces: int, Nfp: int, Np: int, VmapP, EToP) -> FaceConnector:
    """NewFaceConnector; utils/face_connector.go:43-187, same append order.

    For each destination partition p the face points of its elements are
    visited in (localElem, face, fp) order; each goes to the pick list of its
    source partition and the place list of p, so a stable sort of p's face
    points by source partition yields every pick[src][p] / place[p][src].
    """
    if K <= 0 or Nfaces <= 0 or Nfp <= 0 or Np <= 0:
        raise ValueError(f"invalid dimensions: K={K}, Nfaces={Nfaces}, Nfp={Nfp}, Np={Np}")
    VmapP = np.asarray(VmapP, dtype=np.int64)
    EToP = np.asarray(EToP, dtype=np.int64)
    if VmapP.size != K * Nfaces * Nfp:
        raise ValueError(f"VmapP length {VmapP.size} does not match expected {K * Nfaces * Nfp}")
    if EToP.size != K:
        raise ValueError(f"EToP length {EToP.size} does not match K={K}")
    NP = int(EToP.max()) + 1
    epp = np.bincount(EToP, minlength=NP)
    order = np.argsort(EToP, kind="stable")  # global elements grouped by partition
    l2g = np.split(order, np.cumsum(epp)[:-1])
    g2l = np.empty(K, dtype=np.int64)
    for p in range(NP):
        g2l[l2g[p]] = np.arange(l2g[p].size)
    NF = Nfaces * Nfp
    pick, place = {}, {}
    for p in range(NP):
        ge = l2g[p]
        gfp = (ge[:, None] * NF + np.arange(NF)[None, :]).reshape(-1)
        src = VmapP[gfp]
        se = src // Np
        sp = EToP[se]
        lsn = g2l[se] * Np + src % Np
        pos = np.arange(ge.size * NF, dtype=np.int64)
        srt = np.argsort(sp, kind="stable")
        cnt = np.bincount(sp, minlength=NP)
        bounds = np.concatenate(([0], np.cumsum(cnt)))
        for q in range(NP):
            sel = srt[bounds[q]:bounds[q + 1]]
            pick[(q, p)] = lsn[sel]
            place[(p, q)] = pos[sel]
    return FaceConnector(NP, K, Nfaces, Nfp, Np, epp, l2g, g2l, pick, place)


def slab_partition(K: int, nparts: int, imbalance: float = 0.0, seed: int = 0) -> np.ndarray:
    """Element counts of ``nparts`` contiguous partitions covering K elements.

    Weights w_p ~ U(1-imbalance, 1+imbalance) (seeded) are normalised to sum K
    with floor; the remainder goes one each to the first partitions.
    """
    if imbalance > 0:
        w = uniform(seed, 201, nparts, 1.0 - imbalance, 1.0 + imbalance)
    else:
        w = np.ones(nparts)
    kp = np.floor(w / w.sum() * K).astype(np.int64)
    rem = K - int(kp.sum())
    kp[:rem] += 1
    return kp


def _spread3(v: np.ndarray) -> np.ndarray:
    """Interleave the low 21 bits of v with two zero bits (Morton encoding)."""
    v = v.astype(np.uint64) & np.uint64(0x1FFFFF)
    v = (v | (v << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
    v = (v | (v << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
    v = (v | (v << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
    v = (v | (v << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
    v = (v | (v << np.uint64(2))) & np.uint64(0x1249249249249249)
    return v


def morton_within_partitions(mesh: TetMesh, Kp: np.ndarray) -> TetMesh:
    """Renumber elements inside each contiguous partition along a Morton
    (Z-order) curve of their centroids, keeping every partition's element
    set. Face neighbours then sit close in memory, so the neighbour gathers
    of the stage kernel hit in cache. Returns a new TetMesh (same vertices)."""
    K = mesh.K
    cx = mesh.VX[mesh.EToV].mean(axis=1)
    cy = mesh.VY[mesh.EToV].mean(axis=1)
    cz = mesh.VZ[mesh.EToV].mean(axis=1)
    scale = float(1 << 20)
    lo = np.array([cx.min(), cy.min(), cz.min()])
    span = np.array([np.ptp(cx), np.ptp(cy), np.ptp(cz)]) + 1e-30
    qx = ((cx - lo[0]) / span[0] * (scale - 1)).astype(np.uint64)
    qy = ((cy - lo[1]) / span[1] * (scale - 1)).astype(np.uint64)
    qz = ((cz - lo[2]) / span[2] * (scale - 1)).astype(np.uint64)
    key = _spread3(qx) | (_spread3(qy) << np.uint64(1)) | (_spread3(qz) << np.uint64(2))
    part = etop_from_counts(np.asarray(Kp))
    order = np.lexsort((np.arange(K), key, part))
    etop = None if mesh.EToP is None else mesh.EToP[order]
    return TetMesh(mesh.VX, mesh.VY, mesh.VZ, mesh.EToV[order], mesh.EToVc[order], EToP=etop,
                   period=mesh.period, info=dict(mesh.info, order="morton-within-partitions"))


def element_centroids(mesh: TetMesh) -> np.ndarray:
    """[K, 3] centroid of each element (affine-map coordinates)."""
    return np.stack([mesh.VX[mesh.EToV].mean(axis=1), mesh.VY[mesh.EToV].mean(axis=1),
                     mesh.VZ[mesh.EToV].mean(axis=1)], axis=1)


def rcb_partition(mesh: TetMesh, counts) -> np.ndarray:
    """Recursive coordinate bisection of the element centroids into
    ``len(counts)`` partitions of exactly ``counts[p]`` elements; returns EToP.

    A geometric stand-in for the reference's METIS partitioner
    (mesh/partitioner/), which is not available here: each split cuts the
    longest extent of the current element set at the element count the left
    half of the partitions asks for, so partitions come out compact and the
    number of faces between partitions -- the exchange volume -- stays near the
    surface-to-volume minimum instead of growing like the slabs' cross-section.
    """
    counts = np.asarray(counts, dtype=np.int64)
    if counts.ndim != 1 or counts.size == 0 or (counts < 0).any() or int(counts.sum()) != mesh.K:
        raise ValueError("rcb_partition: counts must be non-negative and sum to K")
    cen = element_centroids(mesh)
    etop = np.empty(mesh.K, dtype=np.int64)

    def split(idx: np.ndarray, parts: np.ndarray) -> None:
        if parts.size == 1:
            etop[idx] = parts[0]
            return
        half = parts.size // 2
        nl = int(counts[parts[:half]].sum())
        c = cen[idx]
        axis = int(np.argmax(np.ptp(c, axis=0))) if idx.size else 0
        order = np.argsort(c[:, axis], kind="stable")
        split(idx[order[:nl]], parts[:half])
        split(idx[order[nl:]], parts[half:])

    split(np.arange(mesh.K), np.arange(counts.size))
    return etop


def order_by_partition(mesh: TetMesh, etop: np.ndarray) -> tuple[TetMesh, np.ndarray]:
    """Renumbers elements partition-major (stable within each partition), the
    order the partitioned state arrays and the stage's tile lists use; returns
    the renumbered mesh (EToP set) and the element count of each partition."""
    etop = np.asarray(etop, dtype=np.int64)
    order = np.argsort(etop, kind="stable")
    kp = np.bincount(etop, minlength=int(etop.max()) + 1 if etop.size else 0).astype(np.int64)
    return (TetMesh(mesh.VX, mesh.VY, mesh.VZ, mesh.EToV[order], mesh.EToVc[order], EToP=etop[order],
                    period=mesh.period, info=dict(mesh.info, order="partition-major")), kp)


def etop_from_counts(kp: np.ndarray) -> np.ndarray:
    return np.repeat(np.arange(kp.size, dtype=np.int64), kp)
