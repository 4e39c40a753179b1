"""The learned warm start: a small MLP trained on the GPU (the library's own kernels, cn_mlp_*)
maps a placement's pose and the tissue conductivities to the Galerkin coefficients of its
solution in the reduced basis of solver-generated solutions; the solve then starts from the
basis expansion of the prediction (cn_warm_set_coefficients + warm_start = 2) and still iterates
to the same tolerance, so the plain method remains the fallback and the arbiter of accuracy.

No reference counterpart: cuNIBS has no learned component.
"""

from __future__ import annotations

import ctypes
from collections.abc import Sequence

import numpy as np

from cunibs_amd import _lib
from cunibs_amd._lib import check, lib, ptr

N_POSE = 6
N_SIGMA = 10
N_INPUTS = N_POSE + N_SIGMA  # target direction (x, y), handle direction (3), distance + 10 log sigma


def pose_features(site, sigma: Sequence[float] | None = None) -> np.ndarray:
    """The pose as continuous coordinates: the target direction's (x, y) (its z follows on the upper
    cap), the handle direction projected on the tangent plane at the target as a global unit vector
    (3), the coil distance; then log sigma of the ten tissues (zeros when the table is not varied).
    (theta, phi) and an in-plane handle angle would carry the same information, but angles wrap at
    +-pi and the tangent frame turns fast near the pole, and the network fits a discontinuous input
    badly; the pose takes six inputs instead of four, sixteen in all."""
    center, handle, dist = (np.asarray(site[0], np.float64), np.asarray(site[1], np.float64), float(site[2]))
    r = center / np.linalg.norm(center)
    hd = handle - center
    hd = hd - (hd @ r) * r
    hd /= np.linalg.norm(hd)
    logs = np.zeros(N_SIGMA) if sigma is None else np.log(np.asarray(sigma, np.float64))
    return np.concatenate([[r[0], r[1], hd[0], hd[1], hd[2], dist], logs])


class WarmStartNetwork:
    """MLP N_INPUTS -> hidden -> hidden -> n_out, GELU (cn_mlp_*)."""

    def __init__(self, n_out: int, hidden: int = 512, seed: int = 0, n_in: int = N_INPUTS) -> None:
        h = ctypes.c_void_p()
        check(lib().cn_mlp_create(n_in, hidden, n_out, seed, ctypes.byref(h)), "cn_mlp_create")
        self._h = h
        self.n_in, self.hidden, self.n_out = n_in, hidden, n_out
        self.mean = np.zeros(n_in, np.float32)
        self.scale = np.ones(n_in, np.float32)
        self.out_scale = np.ones(n_out, np.float32)

    def free(self) -> None:
        if getattr(self, "_h", None):
            lib().cn_mlp_destroy(self._h)
            self._h = None

    def __del__(self) -> None:
        try:
            self.free()
        except Exception:
            pass

    @property
    def n_params(self) -> int:
        n = ctypes.c_int64()
        check(lib().cn_mlp_n_params(self._h, ctypes.byref(n)), "cn_mlp_n_params")
        return int(n.value)

    def params(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        check(lib().cn_mlp_get_params(self._h, ptr(out)), "cn_mlp_get_params")
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        if p.shape != (self.n_params,):
            raise ValueError("parameter vector has the wrong length")
        check(lib().cn_mlp_set_params(self._h, ptr(p)), "cn_mlp_set_params")

    def gradients(self, X: np.ndarray, T: np.ndarray) -> tuple[float, np.ndarray]:
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.empty(self.n_params, np.float32)
        loss = ctypes.c_double()
        check(lib().cn_mlp_gradients(self._h, ptr(X), ptr(T), X.shape[0], ptr(g), ctypes.byref(loss)),
              "cn_mlp_gradients")
        return float(loss.value), g

    def _inputs(self, X: np.ndarray) -> np.ndarray:
        return np.ascontiguousarray((np.asarray(X, np.float32) - self.mean) / self.scale, np.float32)

    def fit(self, X: np.ndarray, T: np.ndarray, epochs: int = 200, batch: int = 256, lr: float = 1e-3,
            seed: int = 0, target_scale: str = "energy") -> float:
        """Standardize inputs on the training set, scale the targets, then train with Adam.
        target_scale "energy" divides every target by one number (their RMS), so the squared loss
        is the energy-norm error of the expansion in A-orthonormal modes over the mean solution
        energy; "per_output" standardizes each coefficient, weighting the trailing modes, which
        carry little of the solution, as much as the leading ones."""
        if target_scale not in ("energy", "per_output"):
            raise ValueError("target_scale must be 'energy' or 'per_output'")
        X = np.asarray(X, np.float32)
        T = np.asarray(T, np.float32)
        self.mean = X.mean(axis=0)
        sd = X.std(axis=0)
        self.scale = np.where(sd > 0, sd, 1.0).astype(np.float32)
        if target_scale == "per_output":
            osd = T.std(axis=0)
            self.out_scale = np.where(osd > 0, osd, 1.0).astype(np.float32)
        else:
            rms = float(np.sqrt(np.mean(np.square(T, dtype=np.float64))))
            self.out_scale = np.full(T.shape[1], rms if rms > 0 else 1.0, np.float32)
        xs = self._inputs(X)
        ts = np.ascontiguousarray(T / self.out_scale, np.float32)
        loss = ctypes.c_double()
        check(lib().cn_mlp_train(self._h, ptr(xs), ptr(ts), xs.shape[0], epochs, batch, lr, seed,
                                 ctypes.byref(loss)), "cn_mlp_train")
        return float(loss.value)

    def predict_raw(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        y = np.empty((X.shape[0], self.n_out), np.float32)
        check(lib().cn_mlp_predict(self._h, ptr(X), X.shape[0], ptr(y)), "cn_mlp_predict")
        return y

    def predict(self, X: np.ndarray) -> np.ndarray:
        return self.predict_raw(self._inputs(X)) * self.out_scale


def train_warm_start(ctx, dip_pos_m, dip_moment, sites, didt: float = 1e6, *, modes: int = 256,
                     hidden: int = 512, epochs: int = 300, seed: int = 0, block: int = 16):
    """Solve the training placements (collecting their solutions), build the modes, and fit the
    network to each training solution's Galerkin coefficients. Returns the network."""
    from cunibs_amd import fem
    ctx.warm_reset(len(sites))
    for i in range(0, len(sites), block):
        fem.solve_placements_block(ctx, dip_pos_m, dip_moment, sites[i:i + block], didt, metrics=False,
                                   magnitude=False, collect=True)
    r = ctx.warm_finalize(max_modes=modes)
    targets = ctx.warm_targets()[:, :r]
    X = np.stack([pose_features(s) for s in sites])
    net = WarmStartNetwork(n_out=(r + 15) // 16 * 16, hidden=hidden, seed=seed)
    T = np.zeros((len(sites), net.n_out), np.float32)
    T[:, :r] = targets
    net.fit(X, T, epochs=epochs, seed=seed)
    net.modes = r
    return net


def solve_with_network(ctx, net: WarmStartNetwork, dip_pos_m, dip_moment, sites, didt: float = 1e6, **kw):
    """A block warm-started from the network's predicted coefficients."""
    from cunibs_amd import fem
    X = np.stack([pose_features(s) for s in sites])
    c = net.predict(X)[:, :net.modes].astype(np.float64)
    ctx.warm_set_coefficients(c)
    return fem.solve_placements_block(ctx, dip_pos_m, dip_moment, sites, didt, warm_start=2, **kw)
