"""Seeded synthetic inputs for the E-field path: shell heads, a figure-8 coil, placement sets.

Nothing here touches the GPU. Every generator is a pure function of its arguments and seed, so a
benchmark or a test that names the same arguments gets the same arrays on any machine.

Head
    Concentric shells at radii 70/75/78/86/92 mm tagged 1 WM / 2 GM / 3 CSF / 4 bone / 5 scalp.
    A cubic grid of spacing ``h`` is split into the six Kuhn tetrahedra sharing each cube's main
    diagonal (the same six as the reference's unit-cube fixture, ``tests/conftest.py:37-47``), and a
    tetrahedron is kept when its centroid lies inside the outer shell. Its tag is the innermost
    shell whose radius its centroid lies within. The perturbed heads move every shell by one
    shared radial offset ``delta(theta, phi)``, a seeded sum of real spherical harmonics of degree
    1..6 rescaled to a given RMS over the sphere; one shared offset keeps the shells from crossing.
    The outer boundary faces, wound outward, are the scalp surface (tag 1005).

Coil
    Two wings (centres x = -38 / +38 mm) of three layers (z = 0, -1.5, -3 mm in the coil frame,
    i.e. on the side away from the head) of 400 dipoles on a Vogel spiral filling a 35 mm disc.
    Moments are +z on the left wing and -z on the right, each of magnitude
    ``turns * pi * r_wing^2 / 400`` A m^2 per ampere, positions jittered by N(0, 0.1 mm).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

SHELL_RADII_MM: tuple[float, ...] = (70.0, 75.0, 78.0, 86.0, 92.0)
SHELL_TAGS: tuple[int, ...] = (1, 2, 3, 4, 5)
SKIN_SURFACE_TAG = 1005

# Corner index of a unit cube = x + 2y + 4z.
_CUBE_OFFSETS = np.array(
    [[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1], [1, 1, 1]],
    dtype=np.int64,
)
# The six tetrahedra around the 0-7 diagonal. A translate of this split is conforming with its
# neighbours, so a grid of them is a valid tetrahedral mesh.
_KUHN_TETS = np.array(
    [[0, 1, 3, 7], [0, 3, 2, 7], [0, 2, 6, 7], [0, 6, 4, 7], [0, 4, 5, 7], [0, 5, 1, 7]],
    dtype=np.int64,
)


@dataclass(frozen=True)
class SyntheticHead:
    nodes_mm: np.ndarray  # (n, 3) float64
    tet_nodes: np.ndarray  # (m, 4) int32
    tet_tags: np.ndarray  # (m,) int32
    skin_tris: np.ndarray  # (s, 3) int32, wound outward


def _real_sph_harm(l: int, m: int, cos_t: np.ndarray, phi: np.ndarray) -> np.ndarray:
    """Orthonormal real spherical harmonic Y_lm (unit L2 norm over the sphere)."""
    from scipy.special import lpmv

    am = abs(m)
    norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
    p = lpmv(am, l, cos_t)
    if m == 0:
        return norm * p
    if m > 0:
        return math.sqrt(2.0) * norm * p * np.cos(m * phi)
    return math.sqrt(2.0) * norm * p * np.sin(am * phi)


class ShellPerturbation:
    """delta(direction) = s * sum_{1<=l<=lmax} a_lm Y_lm, a_lm ~ N(0,1), RMS over the sphere fixed."""

    def __init__(self, rms_mm: float, seed: int, lmax: int = 6) -> None:
        rng = np.random.default_rng(seed)
        self.terms = [(l, m) for l in range(1, lmax + 1) for m in range(-l, l + 1)]
        coeffs = rng.standard_normal(len(self.terms))
        # With orthonormal Y_lm the mean square over the sphere is sum(a^2) / (4 pi).
        ms = float((coeffs**2).sum()) / (4.0 * math.pi)
        self.coeffs = coeffs * (rms_mm / math.sqrt(ms)) if rms_mm > 0.0 else coeffs * 0.0
        self.rms_mm = rms_mm

    def __call__(self, dirs: np.ndarray) -> np.ndarray:
        if self.rms_mm == 0.0:
            return np.zeros(dirs.shape[0])
        cos_t = np.clip(dirs[:, 2], -1.0, 1.0)
        phi = np.arctan2(dirs[:, 1], dirs[:, 0])
        out = np.zeros(dirs.shape[0])
        for (l, m), a in zip(self.terms, self.coeffs):
            out += a * _real_sph_harm(l, m, cos_t, phi)
        return out


def _boundary_faces(tets: np.ndarray, nodes: np.ndarray) -> np.ndarray:
    """Faces owned by exactly one tetrahedron, wound so (b-a)x(c-a) points away from the owner."""
    n = int(nodes.shape[0])
    if n >= (1 << 21):
        raise ValueError("face keys pack three node ids into 63 bits; mesh too large")
    faces = tets[:, [[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]]].reshape(-1, 3).astype(np.int64)
    opposite = tets[:, [3, 2, 1, 0]].reshape(-1).astype(np.int64)
    s = np.sort(faces, axis=1)
    keys = (s[:, 0] << 42) | (s[:, 1] << 21) | s[:, 2]
    order = np.argsort(keys, kind="stable")
    ks = keys[order]
    first = np.ones(ks.shape[0], dtype=bool)
    first[1:] = ks[1:] != ks[:-1]
    last = np.ones(ks.shape[0], dtype=bool)
    last[:-1] = ks[1:] != ks[:-1]
    single = order[first & last]
    single.sort()  # keep the tetrahedron order, so the surface is laid out like the volume
    f = faces[single]
    opp = opposite[single]
    a = nodes[f[:, 0]]
    normal = np.cross(nodes[f[:, 1]] - a, nodes[f[:, 2]] - a)
    inward = np.einsum("ij,ij->i", normal, nodes[opp] - a) > 0
    f[inward] = f[inward][:, [0, 2, 1]]
    return f.astype(np.int32)


def shell_head(
    spacing_mm: float,
    *,
    perturb_rms_mm: float = 0.0,
    seed: int = 0,
    radii_mm: tuple[float, ...] = SHELL_RADII_MM,
    tags: tuple[int, ...] = SHELL_TAGS,
) -> SyntheticHead:
    """Kuhn-split shell head; ``perturb_rms_mm = 0`` gives the concentric-sphere head."""
    h = float(spacing_mm)
    pert = ShellPerturbation(perturb_rms_mm, seed)
    outer = radii_mm[-1] + 4.0 * perturb_rms_mm + h
    cells = int(math.ceil(2.0 * outer / h))
    lo = -0.5 * cells * h
    side = cells + 1  # nodes per axis
    radii = np.asarray(radii_mm, dtype=np.float64)
    tag_of = np.asarray(tags, dtype=np.int32)
    # Centroid offset (in cells) of each of the six tetrahedra of a cube.
    cen_off = _CUBE_OFFSETS[_KUHN_TETS].mean(axis=1)  # (6, 3)
    corner_lin = _CUBE_OFFSETS[:, 0] + side * (_CUBE_OFFSETS[:, 1] + side * _CUBE_OFFSETS[:, 2])
    tet_corner_lin = corner_lin[_KUHN_TETS]  # (6, 4) linear offsets of each tet's corners

    ii, jj = np.meshgrid(np.arange(cells), np.arange(cells), indexing="ij")
    ii = ii.ravel().astype(np.int64)
    jj = jj.ravel().astype(np.int64)
    tet_chunks: list[np.ndarray] = []
    tag_chunks: list[np.ndarray] = []
    for k in range(cells):
        cz = lo + h * (k + cen_off[:, 2])  # (6,)
        cx = lo + h * (ii[:, None] + cen_off[None, :, 0])  # (C, 6)
        cy = lo + h * (jj[:, None] + cen_off[None, :, 1])
        r = np.sqrt(cx * cx + cy * cy + cz[None, :] ** 2)
        inside_box = r <= outer
        if not inside_box.any():
            continue
        ci, ti = np.nonzero(inside_box)
        rr = r[ci, ti]
        # |delta| stays within a few RMS, so only centroids near a shell need it evaluated.
        delta = np.zeros(rr.shape[0])
        near = rr > radii[0] - 6.0 * perturb_rms_mm - h
        if perturb_rms_mm > 0.0 and near.any():
            cxs, cys = cx[ci[near], ti[near]], cy[ci[near], ti[near]]
            dirs = np.stack([cxs, cys, cz[ti[near]]], axis=1) / rr[near][:, None]
            delta[near] = pert(dirs)
        shell = np.searchsorted(radii, rr - delta, side="left")  # first radius >= r - delta
        keep = shell < radii.shape[0]
        ci, ti, shell = ci[keep], ti[keep], shell[keep]
        base = ii[ci] + side * (jj[ci] + side * k)
        tet_chunks.append(base[:, None] + tet_corner_lin[ti])
        tag_chunks.append(tag_of[shell])
    tets_lin = np.concatenate(tet_chunks)
    tet_tags = np.concatenate(tag_chunks).astype(np.int32)

    used = np.zeros(side**3, dtype=bool)
    used[tets_lin.ravel()] = True
    node_lin = np.nonzero(used)[0]
    remap = np.full(side**3, -1, dtype=np.int64)
    remap[node_lin] = np.arange(node_lin.shape[0])
    tet_nodes = remap[tets_lin].astype(np.int32)
    iz, rem = np.divmod(node_lin, side * side)
    iy, ix = np.divmod(rem, side)
    nodes_mm = np.stack([lo + h * ix, lo + h * iy, lo + h * iz], axis=1).astype(np.float64)
    skin = _boundary_faces(tet_nodes, nodes_mm)
    return SyntheticHead(
        nodes_mm=np.ascontiguousarray(nodes_mm),
        tet_nodes=np.ascontiguousarray(tet_nodes),
        tet_tags=np.ascontiguousarray(tet_tags),
        skin_tris=np.ascontiguousarray(skin),
    )


@dataclass(frozen=True)
class SyntheticCoil:
    positions_m: np.ndarray  # (D, 3) float64
    moments: np.ndarray  # (D, 3) float64, A m^2 per ampere
    name: str


def figure8_coil(
    *,
    seed: int = 0,
    per_layer: int = 400,
    wing_radius_mm: float = 35.0,
    wing_offset_mm: float = 38.0,
    layer_z_mm: tuple[float, ...] = (0.0, -1.5, -3.0),
    jitter_mm: float = 0.1,
    turns: float = 10.0,
) -> SyntheticCoil:
    rng = np.random.default_rng(seed)
    golden = math.pi * (3.0 - math.sqrt(5.0))
    i = np.arange(per_layer, dtype=np.float64)
    rad = wing_radius_mm * np.sqrt((i + 0.5) / per_layer)
    ang = i * golden
    disc = np.stack([rad * np.cos(ang), rad * np.sin(ang)], axis=1)
    moment = turns * math.pi * (wing_radius_mm * 1e-3) ** 2 / per_layer
    pos, mom = [], []
    for sign, mz in ((-1.0, 1.0), (1.0, -1.0)):
        for z in layer_z_mm:
            p = np.zeros((per_layer, 3))
            p[:, 0] = disc[:, 0] + sign * wing_offset_mm
            p[:, 1] = disc[:, 1]
            p[:, 2] = z
            pos.append(p)
            m = np.zeros((per_layer, 3))
            m[:, 2] = mz * moment
            mom.append(m)
    pos_mm = np.concatenate(pos) + rng.normal(0.0, jitter_mm, size=(2 * len(layer_z_mm) * per_layer, 3))
    return SyntheticCoil(
        positions_m=np.ascontiguousarray(pos_mm * 1e-3),
        moments=np.ascontiguousarray(np.concatenate(mom)),
        name="synthetic_figure8",
    )


def _tangent_frame(d: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    ref = np.array([1.0, 0.0, 0.0]) if abs(d[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = np.cross(d, ref)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(d, e1)
    return e1, e2


def _site(direction: np.ndarray, handle_angle: float, distance_mm: float, radius_mm: float):
    d = direction / np.linalg.norm(direction)
    e1, e2 = _tangent_frame(d)
    center = radius_mm * d
    handle = center + 50.0 * (math.cos(handle_angle) * e1 + math.sin(handle_angle) * e2)
    return center, handle, float(distance_mm)


def cap_sites(
    n: int, *, seed: int = 0, max_polar_deg: float = 60.0, distance_mm: float = 4.0,
    radius_mm: float = SHELL_RADII_MM[-1], distance_range_mm: tuple[float, float] | None = None,
) -> list[tuple[np.ndarray, np.ndarray, float]]:
    """``n`` (center, handle, distance) triples uniform on the upper cap, handle angle U[0, 2pi);
    the distance is ``distance_mm`` or, with ``distance_range_mm``, uniform on that range."""
    rng = np.random.default_rng(seed)
    cmin = math.cos(math.radians(max_polar_deg))
    out = []
    for _ in range(n):
        ct = rng.uniform(cmin, 1.0)
        ph = rng.uniform(0.0, 2.0 * math.pi)
        a = rng.uniform(0.0, 2.0 * math.pi)
        d = distance_mm if distance_range_mm is None else rng.uniform(*distance_range_mm)
        st = math.sqrt(max(0.0, 1.0 - ct * ct))
        out.append(_site(np.array([st * math.cos(ph), st * math.sin(ph), ct]), a, d, radius_mm))
    return out


def grid_sites(
    n_theta: int = 16, n_phi: int = 16, *, max_polar_deg: float = 60.0, distance_mm: float = 4.0,
    radius_mm: float = SHELL_RADII_MM[-1],
) -> list[tuple[np.ndarray, np.ndarray, float]]:
    """theta_i = max_polar * (i + 0.5) / n_theta, phi_j = 2 pi j / n_phi; handle along +phi."""
    out = []
    for i in range(n_theta):
        th = math.radians(max_polar_deg) * (i + 0.5) / n_theta
        for j in range(n_phi):
            ph = 2.0 * math.pi * j / n_phi
            d = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
            out.append(_site(d, 0.5 * math.pi + ph, distance_mm, radius_mm))
    return out


# The single placement of the 4 mm sphere configuration.
SPHERE_SITE = (np.array([0.0, 20.0, 92.0]), np.array([0.0, 70.0, 92.0]), 4.0)
