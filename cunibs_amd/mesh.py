"""Head meshes: the ``HeadMesh`` container and a Gmsh 2.2 reader (mesh.py of the reference,
src/cunibs/mesh.py:136-386, restated).

``load_mesh`` reads Gmsh 2.2 files in binary (file-type 1, little-endian) and ASCII (file-type 0),
optionally gzip-compressed. It keeps tetrahedra (type 4) and triangles (type 2), skips the other
element types Gmsh writes (lines, quadrangles, hexahedra, points), rejects tetrahedra whose tissue
tag is unknown (dropping them would silently delete a tissue), filters triangles to the known
surface tags, drops nodes no tetrahedron references and rebases indices to 0. The skin is the
triangles tagged 1005.
"""

from __future__ import annotations

import gzip
from pathlib import Path
from typing import NamedTuple

import numpy as np

SKIN_SURFACE_TAG = 1005
VOLUME_KEY_TO_LABEL: dict[int, str] = {
    1: "white_matter", 2: "gray_matter", 3: "csf", 4: "average_bone", 5: "scalp",
    6: "eye_balls", 7: "cortical_bone", 8: "cancellous_bone", 9: "blood", 10: "muscle",
}  # mesh.py:25-36
SURFACE_KEY_TO_LABEL: dict[int, str] = {1000 + k: v for k, v in VOLUME_KEY_TO_LABEL.items()}
SURFACE_KEY_TO_LABEL[1099] = "internal_air"  # mesh.py:37-49
_NODES_PER_TYPE = {1: 2, 2: 3, 3: 4, 4: 4, 5: 8, 15: 1}


class MeshArrays(NamedTuple):
    nodes_mm: np.ndarray   # (N, 3) float64
    tet_nodes: np.ndarray  # (M, 4) int32
    tet_tags: np.ndarray   # (M,) int32
    surf_tris: np.ndarray  # (S, 3) int32
    surf_tags: np.ndarray  # (S,) int32


class _Reader:
    """Cursor over the file's bytes; Gmsh 2.2 interleaves text lines and binary blocks."""

    def __init__(self, data: bytes, path: Path) -> None:
        self.data = data
        self.path = path
        self.pos = 0

    def fail(self, what: str):
        raise ValueError(f"{self.path}: {what}")

    def line(self) -> str:
        end = self.data.find(b"\n", self.pos)
        if end < 0:
            self.fail("truncated: a line has no end")
        out = self.data[self.pos:end].decode(errors="replace").strip()
        self.pos = end + 1
        return out

    def expect(self, marker: str) -> None:
        got = self.line()
        if got != marker:
            self.fail(f"expected {marker}, found {got!r}")

    def count(self, what: str) -> int:
        text = self.line()
        if not text.isdigit():
            self.fail(f"expected a {what} count, found {text!r}")
        return int(text)

    def skip_newline(self) -> None:
        """Gmsh itself ends a binary block with a newline; some writers do not."""
        if self.data[self.pos:self.pos + 1] == b"\n":
            self.pos += 1

    def take(self, n_bytes: int) -> memoryview:
        if n_bytes < 0 or self.pos + n_bytes > len(self.data):
            self.fail("truncated: a block runs past the end of the file")
        out = memoryview(self.data)[self.pos:self.pos + n_bytes]
        self.pos += n_bytes
        return out


def _read_bytes(path: Path) -> bytes:
    raw = path.read_bytes()
    return gzip.decompress(raw) if raw[:2] == b"\x1f\x8b" else raw


def parse_msh(mesh_file: str | Path) -> MeshArrays:
    """Parse a Gmsh 2.2 mesh (binary or ASCII, optionally gzipped) into 0-based arrays."""
    path = Path(mesh_file)
    r = _Reader(_read_bytes(path), path)
    r.expect("$MeshFormat")
    fields = r.line().split()
    if len(fields) != 3 or fields[0] not in ("2.2", "2.1", "2"):
        r.fail(f"expected a Gmsh 2.2 header, found {' '.join(fields)!r}")
    binary = fields[1] == "1"
    if binary:
        if fields[2] != "8":
            r.fail(f"expected 8-byte reals, found {fields[2]}")
        if np.frombuffer(r.take(4), "<i4")[0] != 1:
            r.fail("byte-order marker is not 1; big-endian meshes are not supported")
        r.line()
    r.expect("$EndMeshFormat")

    # Skip sections before $Nodes ($PhysicalNames and the like).
    while True:
        head = r.line()
        if head == "$Nodes":
            break
        if head.startswith("$") and not head.startswith("$End"):
            end = "$End" + head[1:]
            while r.line() != end:
                pass
        elif head and not head.startswith("$"):
            r.fail(f"unexpected line {head!r} before $Nodes")
    n_nodes = r.count("node")
    if binary:
        rec = np.frombuffer(r.take(28 * n_nodes), dtype=np.dtype([("id", "<i4"), ("xyz", "<f8", (3,))]))
        node_ids = rec["id"].astype(np.int64)
        xyz = np.array(rec["xyz"], dtype=np.float64)
        r.skip_newline()
        r.expect("$EndNodes")
    else:
        rows = [r.line().split() for _ in range(n_nodes)]
        node_ids = np.array([int(t[0]) for t in rows], dtype=np.int64)
        xyz = np.array([[float(v) for v in t[1:4]] for t in rows], dtype=np.float64)
        r.expect("$EndNodes")

    r.expect("$Elements")
    n_elem = r.count("element")
    tets, tet_tags, tris, tri_tags = [], [], [], []
    if binary:
        seen = 0
        while seen < n_elem:
            etype, cnt, ntags = (int(v) for v in np.frombuffer(r.take(12), "<i4"))
            nper = _NODES_PER_TYPE.get(etype)
            if nper is None:
                r.fail(f"unsupported Gmsh element type {etype}")
            if cnt <= 0:
                r.fail("an element block declares no elements")
            block = np.frombuffer(r.take(4 * cnt * (1 + ntags + nper)), "<i4").reshape(cnt, 1 + ntags + nper)
            if etype == 4:
                tets.append(block[:, 1 + ntags:])
                tet_tags.append(block[:, 1])
            elif etype == 2:
                tris.append(block[:, 1 + ntags:])
                tri_tags.append(block[:, 1])
            seen += cnt
        r.skip_newline()
    else:
        for _ in range(n_elem):
            t = [int(v) for v in r.line().split()]
            etype, ntags = t[1], t[2]
            nodes = t[3 + ntags:]
            if etype not in _NODES_PER_TYPE or len(nodes) != _NODES_PER_TYPE[etype]:
                r.fail(f"unsupported or malformed element {t[0]} of type {etype}")
            if etype == 4:
                tets.append(np.array([nodes])); tet_tags.append(np.array([t[3]]))
            elif etype == 2:
                tris.append(np.array([nodes])); tri_tags.append(np.array([t[3]]))
    r.expect("$EndElements")
    if not tets:
        r.fail("no tetrahedra")

    tet_ids = np.concatenate(tets).astype(np.int64)
    tet_tag = np.concatenate(tet_tags).astype(np.int32)
    tri_ids = np.concatenate(tris).astype(np.int64) if tris else np.empty((0, 3), np.int64)
    tri_tag = np.concatenate(tri_tags).astype(np.int32) if tris else np.empty(0, np.int32)

    known = np.isin(tet_tag, list(VOLUME_KEY_TO_LABEL))
    if not known.all():
        bad, cnt = np.unique(tet_tag[~known], return_counts=True)
        listing = ", ".join(f"{int(b)}: {int(c)}" for b, c in zip(bad, cnt))
        r.fail(f"{int(cnt.sum())} tetrahedra carry unknown volume tags ({listing})")
    keep_tri = np.isin(tri_tag, list(SURFACE_KEY_TO_LABEL))
    tri_ids, tri_tag = tri_ids[keep_tri], tri_tag[keep_tri]

    # Node ids -> 0-based rows of the table, then keep only nodes some tetrahedron uses.
    order = np.argsort(node_ids, kind="stable")
    sorted_ids = node_ids[order]

    def rows_of(ids: np.ndarray, what: str) -> np.ndarray:
        pos = np.searchsorted(sorted_ids, ids)
        ok = (pos < sorted_ids.size) & (sorted_ids[np.minimum(pos, sorted_ids.size - 1)] == ids)
        if not ok.all():
            r.fail(f"{what} references a node id absent from the node table")
        return order[pos]

    tet_rows = rows_of(tet_ids, "a tetrahedron")
    tri_rows = rows_of(tri_ids, "a surface triangle")
    used = np.zeros(n_nodes, dtype=bool)
    used[tet_rows.ravel()] = True
    new_index = np.full(n_nodes, -1, dtype=np.int64)
    new_index[used] = np.arange(int(used.sum()))
    tri_new = new_index[tri_rows]
    if (tri_new < 0).any():
        r.fail(f"{int((tri_new < 0).any(axis=1).sum())} surface triangles reference nodes no tetrahedron uses")
    nodes = np.ascontiguousarray(xyz[used])
    if not np.isfinite(nodes).all():
        r.fail(f"{int((~np.isfinite(nodes).all(axis=1)).sum())} referenced nodes have non-finite coordinates")
    return MeshArrays(nodes, new_index[tet_rows].astype(np.int32), tet_tag, tri_new.astype(np.int32), tri_tag)


class HeadMesh:
    """A tetrahedral head mesh: node coordinates in mm, zero-based connectivity, tissue tags and
    the skin triangles placements project onto (mesh.py:307-373)."""

    def __init__(self, nodes_mm: np.ndarray, tet_nodes: np.ndarray, tet_tags: np.ndarray,
                 skin_tris: np.ndarray) -> None:
        self.nodes_mm = np.ascontiguousarray(nodes_mm, dtype=np.float64).reshape(-1, 3)
        self.tet_nodes = np.ascontiguousarray(tet_nodes, dtype=np.int32).reshape(-1, 4)
        self.tet_tags = np.ascontiguousarray(tet_tags, dtype=np.int32).reshape(-1)
        self.skin_tris = np.ascontiguousarray(skin_tris, dtype=np.int32).reshape(-1, 3)
        if self.tet_tags.shape[0] != self.tet_nodes.shape[0]:
            raise ValueError("tet_tags must have one entry per tetrahedron")

    @property
    def n_nodes(self) -> int:
        return int(self.nodes_mm.shape[0])

    @property
    def tet_barycenters_mm(self) -> np.ndarray:
        return self.nodes_mm[self.tet_nodes].mean(axis=1)

    def __repr__(self) -> str:
        return f"HeadMesh(n_nodes={self.n_nodes}, n_tets={self.tet_nodes.shape[0]})"


def load_mesh(mesh_file: str | Path) -> HeadMesh:
    """Load a Gmsh 2.2 head mesh (mesh.py:376-386); the skin is the surface tagged 1005."""
    a = parse_msh(mesh_file)
    return HeadMesh(a.nodes_mm, a.tet_nodes, a.tet_tags, a.surf_tris[a.surf_tags == SKIN_SURFACE_TAG])


def write_msh(path: str | Path, mesh: HeadMesh, binary: bool = True) -> None:
    """Write a Gmsh 2.2 file (tetrahedra, then the skin as tag-1005 triangles)."""
    n, m, s = mesh.n_nodes, mesh.tet_nodes.shape[0], mesh.skin_tris.shape[0]
    out = bytearray()
    out += b"$MeshFormat\n" + (b"2.2 1 8\n" if binary else b"2.2 0 8\n")
    if binary:
        out += np.array([1], "<i4").tobytes() + b"\n"
    out += b"$EndMeshFormat\n$Nodes\n" + f"{n}\n".encode()
    ids = np.arange(1, n + 1)
    if binary:
        rec = np.empty(n, dtype=np.dtype([("id", "<i4"), ("xyz", "<f8", (3,))]))
        rec["id"], rec["xyz"] = ids, mesh.nodes_mm
        out += rec.tobytes() + b"\n"
    else:
        out += "".join(f"{i} {x!r} {y!r} {z!r}\n" for i, (x, y, z) in zip(ids, mesh.nodes_mm.tolist())).encode()
    out += b"$EndNodes\n$Elements\n" + f"{m + s}\n".encode()
    if binary:
        for etype, conn, tags, first in ((4, mesh.tet_nodes, mesh.tet_tags, 1),
                                         (2, mesh.skin_tris, np.full(s, SKIN_SURFACE_TAG), m + 1)):
            cnt = conn.shape[0]
            if cnt == 0:
                continue
            out += np.array([etype, cnt, 2], "<i4").tobytes()
            blk = np.column_stack([np.arange(first, first + cnt), tags, tags, conn + 1]).astype("<i4")
            out += blk.tobytes()
        out += b"\n"
    else:
        lines = [f"{e + 1} 4 2 {t} {t} " + " ".join(str(v + 1) for v in c)
                 for e, (c, t) in enumerate(zip(mesh.tet_nodes.tolist(), mesh.tet_tags.tolist()))]
        lines += [f"{m + j + 1} 2 2 {SKIN_SURFACE_TAG} {SKIN_SURFACE_TAG} " + " ".join(str(v + 1) for v in c)
                  for j, c in enumerate(mesh.skin_tris.tolist())]
        out += ("\n".join(lines) + "\n").encode()
    out += b"$EndElements\n"
    Path(path).write_bytes(bytes(out))
