"""Seeded synthetic point clouds for the five workloads of the neighbour-search path.

All generators use numpy's PCG64 with fixed seeds and return float32 (n, 3) arrays.
  plane_sphere_cloud  100,000 points: planes z=0.2, x=0.8, y=0.5 over [0,1]^2, spheres r=0.15 at
                      (0.3,0.3,0.6) and (0.7,0.6,0.4), area-weighted surface sampling, N(0, 0.002) noise,
                      2% uniform outliers in [0,1]^3, shuffled (seed 42)
  lidar_frame         64 rings in [-25 deg, +3 deg] x 18,750 azimuths, sensor at z=1.7 m, ground z=0 plus
                      40 axis-aligned boxes (edges 0.5-5 m, centres within +-80 m), 80 m max range, range
                      noise 0.02 m; no-hit rays dropped (seed 1001)
  indoor_scan         20,000,000 points: 10x8x3 m room (floor, ceiling, 4 walls), 60 boxes, 20 cylinders,
                      area-weighted sampling, N(0, 0.003) noise (seed 2002)
  batch_clouds        256 clouds, sizes log-uniform in [2,000, 500,000], each a plane_sphere_cloud scaled
                      to a random box of edge 1-20 m (seeds 3000+i)
  lidar_stream        frames with 16/32/64/128 rings (seeds 5000+j), Poisson arrivals (seed 4004)
"""
from __future__ import annotations

import numpy as np


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


# ------------------------------------------------------------------------------------------------------
# plane + sphere + outliers (100k points)
# ------------------------------------------------------------------------------------------------------
def plane_sphere_cloud(n=100_000, seed=42, noise=0.002, outlier_frac=0.02, scale=1.0):
    rng = _rng(seed)
    n_out = int(round(n * outlier_frac))
    n_surf = n - n_out
    r = 0.15
    areas = np.array([1.0, 1.0, 1.0, 4 * np.pi * r * r, 4 * np.pi * r * r])
    counts = rng.multinomial(n_surf, areas / areas.sum())
    parts = []
    # planes z=0.2, x=0.8, y=0.5 clipped to [0,1]^2
    u = rng.random((counts[0], 2)); parts.append(np.c_[u[:, 0], u[:, 1], np.full(counts[0], 0.2)])
    u = rng.random((counts[1], 2)); parts.append(np.c_[np.full(counts[1], 0.8), u[:, 0], u[:, 1]])
    u = rng.random((counts[2], 2)); parts.append(np.c_[u[:, 0], np.full(counts[2], 0.5), u[:, 1]])
    for c, centre in zip(counts[3:], [(0.3, 0.3, 0.6), (0.7, 0.6, 0.4)]):
        d = rng.normal(size=(c, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        parts.append(np.asarray(centre) + r * d)
    surf = np.concatenate(parts)
    surf += rng.normal(0.0, noise, surf.shape)
    pts = np.concatenate([surf, rng.random((n_out, 3))])
    rng.shuffle(pts, axis=0)
    return (pts * scale).astype(np.float32)


# ------------------------------------------------------------------------------------------------------
# LiDAR-like frame
# ------------------------------------------------------------------------------------------------------
def _boxes(rng, nbox=40, extent=80.0):
    size = rng.uniform(0.5, 5.0, (nbox, 3))
    ctr = rng.uniform(-extent, extent, (nbox, 2))
    lo = np.c_[ctr - size[:, :2] / 2, np.zeros(nbox)]
    hi = np.c_[ctr + size[:, :2] / 2, size[:, 2]]
    return lo, hi


def lidar_frame(rings=64, azimuths=18_750, seed=1001, sensor_z=1.7, max_range=80.0, noise=0.02, nbox=40,
                chunk=2_000_000):
    rng = _rng(seed)
    lo, hi = _boxes(rng, nbox)
    elev = np.radians(np.linspace(-25.0, 3.0, rings))
    az = np.linspace(0.0, 2 * np.pi, azimuths, endpoint=False)
    E, A = np.meshgrid(elev, az, indexing="ij")
    d = np.stack([np.cos(E) * np.cos(A), np.cos(E) * np.sin(A), np.sin(E)], -1).reshape(-1, 3)
    o = np.array([0.0, 0.0, sensor_z])
    out = []
    for s in range(0, len(d), chunk):
        dd = d[s:s + chunk]
        t = np.full(len(dd), np.inf)
        down = dd[:, 2] < 0
        t[down] = -sensor_z / dd[down, 2]                      # ground z = 0
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = 1.0 / dd
            for b in range(len(lo)):                           # slab test per box
                t1 = (lo[b] - o) * inv
                t2 = (hi[b] - o) * inv
                tn = np.nanmax(np.minimum(t1, t2), axis=1)
                tf = np.nanmin(np.maximum(t1, t2), axis=1)
                hit = (tf >= np.maximum(tn, 0.0)) & (tn > 0)
                t = np.where(hit & (tn < t), tn, t)
        keep = t <= max_range
        tt = t[keep] + rng.normal(0.0, noise, keep.sum())
        out.append(o + dd[keep] * tt[:, None])
    return np.concatenate(out).astype(np.float32)


# ------------------------------------------------------------------------------------------------------
# dense indoor scan
# ------------------------------------------------------------------------------------------------------
def indoor_scan(n=20_000_000, seed=2002, noise=0.003):
    rng = _rng(seed)
    W, D, H = 10.0, 8.0, 3.0
    rects = []   # (origin, u, v) parallelograms

    def rect(o, u, v):
        rects.append((np.asarray(o, float), np.asarray(u, float), np.asarray(v, float)))

    rect((0, 0, 0), (W, 0, 0), (0, D, 0)); rect((0, 0, H), (W, 0, 0), (0, D, 0))
    rect((0, 0, 0), (W, 0, 0), (0, 0, H)); rect((0, D, 0), (W, 0, 0), (0, 0, H))
    rect((0, 0, 0), (0, D, 0), (0, 0, H)); rect((W, 0, 0), (0, D, 0), (0, 0, H))
    for _ in range(60):
        sx, sy = rng.uniform(0.3, 1.5, 2)
        h = rng.uniform(0.3, 2.0)
        x0, y0 = rng.uniform(0, W - sx), rng.uniform(0, D - sy)
        rect((x0, y0, h), (sx, 0, 0), (0, sy, 0))
        rect((x0, y0, 0), (sx, 0, 0), (0, 0, h)); rect((x0, y0 + sy, 0), (sx, 0, 0), (0, 0, h))
        rect((x0, y0, 0), (0, sy, 0), (0, 0, h)); rect((x0 + sx, y0, 0), (0, sy, 0), (0, 0, h))
    cyl = []
    for _ in range(20):
        r = rng.uniform(0.1, 0.4)
        h = rng.uniform(0.4, 2.0)
        cx, cy = rng.uniform(r, W - r), rng.uniform(r, D - r)
        cyl.append((cx, cy, r, h))
    a_rect = np.array([np.linalg.norm(np.cross(u, v)) for _, u, v in rects])
    a_side = np.array([2 * np.pi * r * h for _, _, r, h in cyl])
    a_top = np.array([np.pi * r * r for _, _, r, _ in cyl])
    areas = np.concatenate([a_rect, a_side, a_top])
    counts = rng.multinomial(n, areas / areas.sum())
    out = np.empty((n, 3), np.float32)
    k = 0
    for i, (o, u, v) in enumerate(rects):
        c = counts[i]
        st = rng.random((c, 2))
        out[k:k + c] = o + st[:, :1] * u + st[:, 1:] * v
        k += c
    nr = len(rects)
    for j, (cx, cy, r, h) in enumerate(cyl):
        c = counts[nr + j]
        th = rng.uniform(0, 2 * np.pi, c)
        out[k:k + c] = np.c_[cx + r * np.cos(th), cy + r * np.sin(th), rng.uniform(0, h, c)]
        k += c
        c = counts[nr + len(cyl) + j]
        rr = r * np.sqrt(rng.random(c))
        th = rng.uniform(0, 2 * np.pi, c)
        out[k:k + c] = np.c_[cx + rr * np.cos(th), cy + rr * np.sin(th), np.full(c, h)]
        k += c
    out += rng.normal(0.0, noise, out.shape).astype(np.float32)
    rng.shuffle(out, axis=0)
    return out


# ------------------------------------------------------------------------------------------------------
# batched / streaming workloads
# ------------------------------------------------------------------------------------------------------
def batch_sizes(count=256, lo=2_000, hi=500_000, seed=3000):
    rng = _rng(seed)
    return np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(np.int64)


def batch_cloud(i, n, seed=3000):
    """Cloud i of the batch (n points): a plane+sphere cloud scaled by s ~ U[1, 20]; returns (cloud, s)."""
    rng = _rng(seed + 1 + i)
    s = rng.uniform(1.0, 20.0)
    return plane_sphere_cloud(int(n), seed=seed + 1 + i, noise=0.002, scale=s), s


def batch_clouds(count=256, seed=3000, with_scales=False, which=None):
    """`count` plane+sphere clouds with log-uniform sizes in [2e3, 5e5], each scaled by s ~ U[1, 20] (all
    lengths, noise included); with_scales also returns the scales (the radius of cloud i is 0.02 s_i).
    which: indices of the clouds to generate (default all), in that order."""
    sizes = batch_sizes(count, seed=seed)
    out, scales = [], []
    for i in (range(count) if which is None else which):
        c, s = batch_cloud(i, sizes[i], seed)
        out.append(c)
        scales.append(s)
    return (out, np.array(scales)) if with_scales else out


def lidar_stream(frames=32, seed=4004):
    rng = _rng(seed)
    rings = rng.choice([16, 32, 64, 128], frames)
    return [(int(r), 5000 + j) for j, r in enumerate(rings)]
