"""Seeded synthetic VELO-pixel events in the reference's per-event layout.

Geometry: 48 pixel planes at z_k = -300 + 22 k mm (k = 0..47, int32 mm),
active area |x|, |y| <= 42 mm with a hole at r < 5.1 mm; a hit's z is the
plane z as float32.  Straight tracks from vertices with x, y ~ N(0, 0.05 mm)
and z ~ N(0, 50 mm); 80 % of tracks with eta ~ U[2, 5], 20 % with
eta ~ U[-4, -2], phi ~ U[0, 2 pi).  A crossed plane inside the active area
records a hit with efficiency 0.98 and N(0, 0.012 mm) smear in x and y.
Uniform noise hits add a fraction of the signal hit count.  Hits get unique
random uint32 IDs and are sorted by x within each sensor; a tied x within a
sensor is redrawn, so no sort-order ambiguity remains.

Workloads (seeds are part of the definition):

* ``single``   1 event, 150 tracks, seed 1 (about 2,500 hits)
* ``plain51``  51 events, Poisson(150) tracks, seeds 1001..1051
* ``batch1000`` 1,000 events, Poisson(7.6) vertices x Poisson(20) tracks,
  1 % noise, event i seeded 2000*10**4 + i
* ``stress1000`` 1,000 events, Poisson(15) vertices x Poisson(25) tracks,
  5 % noise, event i seeded 3000*10**4 + i
* ``serve256`` the serving pool: 256 events drawn like ``plain51``
  (Poisson(150) tracks), seeds 1001..1256; a serving run at fraction r of the
  batched peak draws arrival gaps ~ Exp and pool indices from the generator
  seeded 4000 + round(100 r) (:func:`serve_arrivals`)
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .event import Event

N_PLANES = 48
PLANE_Z = (-300 + 22 * np.arange(N_PLANES)).astype(np.int32)
HALF_SIZE = 42.0
HOLE_R = 5.1
EFFICIENCY = 0.98
SMEAR = 0.012
VTX_XY_SIGMA = 0.05
VTX_Z_SIGMA = 50.0


@dataclass
class Truth:
    """Monte-Carlo truth of a generated event (hit order of the packed event)."""

    particle: np.ndarray  # int32[N], -1 for noise
    n_particles: int


def _unique_ids(rng, n):
    ids = rng.integers(0, 2**32, n, dtype=np.uint64)
    while True:
        u, first = np.unique(ids, return_index=True)
        if u.size == n:
            return ids.astype(np.uint32)
        dup = np.ones(n, bool)
        dup[first] = False
        ids[dup] = rng.integers(0, 2**32, int(dup.sum()), dtype=np.uint64)


def generate_event(seed: int, *, n_tracks=None, vertices_mean=None, tracks_per_vertex_mean=None,
                   tracks_mean=None, noise_fraction=0.01, return_truth=False):
    """One event.  Exactly one of ``n_tracks`` (fixed count, one vertex),
    ``tracks_mean`` (Poisson count, one vertex) or
    ``vertices_mean``/``tracks_per_vertex_mean`` (Poisson pileup) is given."""
    rng = np.random.default_rng(seed)
    if n_tracks is not None:
        per_vertex = [int(n_tracks)]
    elif tracks_mean is not None:
        per_vertex = [int(rng.poisson(tracks_mean))]
    else:
        nv = int(rng.poisson(vertices_mean))
        per_vertex = [int(c) for c in rng.poisson(tracks_per_vertex_mean, nv)]
    n_trk = int(sum(per_vertex))

    vx = rng.normal(0.0, VTX_XY_SIGMA, len(per_vertex))
    vy = rng.normal(0.0, VTX_XY_SIGMA, len(per_vertex))
    vz = rng.normal(0.0, VTX_Z_SIGMA, len(per_vertex))
    vid = np.repeat(np.arange(len(per_vertex)), per_vertex)

    forward = rng.random(n_trk) < 0.8
    eta = np.where(forward, rng.uniform(2.0, 5.0, n_trk), rng.uniform(-4.0, -2.0, n_trk))
    phi = rng.uniform(0.0, 2 * np.pi, n_trk)
    theta = 2.0 * np.arctan(np.exp(-eta))
    tan_t = np.tan(theta)
    tx = tan_t * np.cos(phi)
    ty = tan_t * np.sin(phi)
    direction = np.sign(np.cos(theta))

    zk = PLANE_Z.astype(np.float64)
    dz = zk[None, :] - vz[vid][:, None]  # [trk, plane]
    x = vx[vid][:, None] + tx[:, None] * dz
    y = vy[vid][:, None] + ty[:, None] * dz
    r = np.hypot(x, y)
    crossed = (dz * direction[:, None]) > 0
    inside = (np.abs(x) <= HALF_SIZE) & (np.abs(y) <= HALF_SIZE) & (r >= HOLE_R)
    seen = crossed & inside & (rng.random(dz.shape) < EFFICIENCY)
    ti, pk = np.nonzero(seen)
    hx_true = x[ti, pk]
    hy_true = y[ti, pk]
    hx = hx_true + rng.normal(0.0, SMEAR, ti.size)
    hy = hy_true + rng.normal(0.0, SMEAR, ti.size)
    plane = pk.astype(np.int64)
    particle = ti.astype(np.int32)

    n_noise = int(round(noise_fraction * ti.size))
    if n_noise:
        nplane = rng.integers(0, N_PLANES, n_noise)
        nx = np.empty(n_noise)
        ny = np.empty(n_noise)
        todo = np.arange(n_noise)
        while todo.size:
            cx = rng.uniform(-HALF_SIZE, HALF_SIZE, todo.size)
            cy = rng.uniform(-HALF_SIZE, HALF_SIZE, todo.size)
            ok = np.hypot(cx, cy) >= HOLE_R
            nx[todo[ok]] = cx[ok]
            ny[todo[ok]] = cy[ok]
            todo = todo[~ok]
        hx_true = np.concatenate([hx_true, nx])
        hy_true = np.concatenate([hy_true, ny])
        hx = np.concatenate([hx, nx])
        hy = np.concatenate([hy, ny])
        plane = np.concatenate([plane, nplane])
        particle = np.concatenate([particle, np.full(n_noise, -1, np.int32)])

    X = hx.astype(np.float32)
    Y = hy.astype(np.float32)
    # Redraw tied x within a sensor until every x of a sensor is distinct.
    while True:
        key = np.lexsort((X, plane))
        sp, sx = plane[key], X[key]
        tie = np.zeros(X.size, bool)
        same = (sp[1:] == sp[:-1]) & (sx[1:] == sx[:-1])
        tie[key[1:][same]] = True
        if not tie.any():
            break
        k = int(tie.sum())
        X[tie] = (hx_true[tie] + rng.normal(0.0, SMEAR, k)).astype(np.float32)

    order = np.lexsort((X, plane))
    X, Y, plane, particle = X[order], Y[order], plane[order], particle[order]
    N = X.size
    hitNums = np.bincount(plane, minlength=N_PLANES).astype(np.int32)
    hitStarts = np.concatenate([[0], np.cumsum(hitNums)[:-1]]).astype(np.int32)
    Z = PLANE_Z[plane].astype(np.float32)
    ids = _unique_ids(rng, N)
    ev = Event(PLANE_Z.copy(), hitStarts, hitNums, ids, X, Y, Z)
    if return_truth:
        return ev, Truth(particle.astype(np.int32), n_trk)
    return ev


# --------------------------------------------------------------------------
# Named workloads
# --------------------------------------------------------------------------
def single_event(return_truth=False):
    """1 event, 150 tracks, seed 1."""
    return generate_event(1, n_tracks=150, noise_fraction=0.01, return_truth=return_truth)


def plain_event(i: int, return_truth=False):
    """Event i of the plain workload (0..50) and of the serving pool (0..255),
    seed 1001 + i."""
    return generate_event(1001 + i, tracks_mean=150, noise_fraction=0.01, return_truth=return_truth)


def batch_event(i: int, return_truth=False):
    """Event i of the 1,000-event batched workload (Poisson(7.6) x Poisson(20))."""
    return generate_event(2000 * 10**4 + i, vertices_mean=7.6, tracks_per_vertex_mean=20,
                          noise_fraction=0.01, return_truth=return_truth)


def stress_event(i: int, return_truth=False):
    """Event i of the 1,000-event stress workload (Poisson(15) x Poisson(25), 5 % noise)."""
    return generate_event(3000 * 10**4 + i, vertices_mean=15, tracks_per_vertex_mean=25,
                          noise_fraction=0.05, return_truth=return_truth)


def serve_arrivals(fraction: float, rate_per_s: float, n: int, pool_size: int):
    """Open-loop Poisson arrivals for the serving benchmark: (pick, arrival_ns),
    the first arrival at 0."""
    rng = np.random.default_rng(4000 + int(round(100 * fraction)))
    gaps = rng.exponential(1e9 / rate_per_s, n)
    arrival = np.cumsum(gaps)
    arrival -= arrival[0]
    pick = rng.integers(0, pool_size, n).astype(np.int32)
    return pick, arrival.astype(np.int64)


WORKLOADS = {
    "single": (1, lambda i: single_event()),
    "plain51": (51, plain_event),
    "batch1000": (1000, batch_event),
    "stress1000": (1000, stress_event),
    "serve256": (256, plain_event),
}


def workload_blobs(name: str, n: int | None = None) -> list[bytes]:
    count, fn = WORKLOADS[name]
    if n is None:
        n = count
    return [fn(i).pack() for i in range(n)]
