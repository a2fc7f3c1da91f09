"""Synthetic inputs of the benchmark workloads (host side, generated once).

All randomness comes from a SplitMix64 stream so that the same arrays can be
rebuilt anywhere: uniform u = (x >> 40) * 2^-24, Gaussians by Box-Muller in
FP64 rounded to FP32.

Workloads (sizes are the benchmark's):
  dam_break(nx, ny, spacing)  square lattice dam break at ((i+1/2)s, (j+1/2)s),
                              jitter U(+-1e-3 s) per component, kernel support
                              radius 2s (the reference's params.support h = s,
                              cut-off and cell 2h), hex walls 3h thick around
                              the unit box
                              (sph2d_box/solids.cpp:139-156 without the random
                              obstacles), rest density 1000, zero velocity.
      12,800 particles:  80 x 160 at s = 0.005, dt = 2e-4
      1,048,576:         1024 x 1024 at s = 1/2560, dt = 1.5625e-5 (same CFL number)
  batch_dam_breaks(256)       256 dam breaks, N ~ log-U[2k, 200k], block aspect
                              (width / height) ~ U[0.25, 2], spacing so that
                              the block's longer side is 0.8, h = s, largest
                              dt = CFL 0.4 at the free-fall speed sqrt(2 g H)
                              of the block height H (the device CFL 0.4 over
                              the live max |v| takes over when it is lower)
  search_points(n)            half uniform in [0,1)^2, half from 8 Gaussian
                              clusters (sigma 0.02) with centres U[0.1, 0.9]^2,
                              samples outside [0,1)^2 redrawn.
"""

from __future__ import annotations

import dataclasses
import math

import numpy as np

MASK64 = (1 << 64) - 1


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n outputs of SplitMix64 started at seed (uint64)."""
    out = np.empty(n, np.uint64)
    # Vectorised: state_k = seed + (k+1) * gamma, then the finaliser.
    gamma = np.uint64(0x9E3779B97F4A7C15)
    k = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & MASK64) + k * gamma
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    out[:] = z
    return out


def uniform01(seed: int, n: int) -> np.ndarray:
    """u = (x >> 40) * 2^-24 in [0, 1), float64 (exact 24-bit values)."""
    x = splitmix64(seed, n)
    return (x >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)


def cubic_W(r: np.ndarray, h: float) -> np.ndarray:
    """sph_common/kernels.h:25-35 in float64."""
    q = r / h
    k = h * h * h * float(np.float32(3.14159274101257324218750))
    w = np.where(q <= 1.0, 1.0 + q * q * (q * 0.75 - 1.5),
                 np.where(q <= 2.0, (2.0 - q) ** 3 / 4.0, 0.0))
    return w / k


def lattice_mass(spacing: float, h: float, rho0: float) -> np.float32:
    """rho0 / (W(0) + sum W) over an interior square lattice: the mass that
    puts the unjittered lattice at rest density (an explicit override of the
    reference's hex calibration, config.cpp:16-102)."""
    m = int(math.ceil(2.0 * h / spacing)) + 1
    i, j = np.meshgrid(np.arange(-m, m + 1), np.arange(-m, m + 1), indexing="ij")
    r = spacing * np.sqrt((i * i + j * j).astype(np.float64))
    s = float(np.sum(cubic_W(r[r < 2.0 * h], h)))
    return np.float32(rho0 / s)


def lattice_positions(nx: int, ny: int, spacing: float, seed: int, jitter: float) -> np.ndarray:
    """Square lattice at ((i + 1/2) s, (j + 1/2) s), particle p = j * nx + i,
    jittered by (2u - 1) * jitter * s with u = uniform01(seed, 2 nx ny)[2p],
    [2p + 1] (evaluated in FP64, rounded to FP32; the C ABI's
    emsph_lattice_positions)."""
    import ctypes
    from ._capi import lib
    out = np.empty((nx * ny, 2), np.float32)
    lib().emsph_lattice_positions(nx, ny, float(spacing), int(seed) & MASK64, float(jitter),
                                  out.ctypes.data_as(ctypes.c_void_p))
    return out


@dataclasses.dataclass
class DamBreak:
    positions: np.ndarray
    velocities: np.ndarray
    solids: np.ndarray
    support: float
    spacing: float
    dt: float
    mass: np.float32
    rho0: float = 1000.0
    gravity: float = 9.81


def dam_break(nx: int, ny: int, spacing: float, dt: float, seed: int = 1,
              jitter: float = 1e-3, with_walls: bool = True) -> DamBreak:
    from . import sph2d_box as sb
    s = float(spacing)
    # The kernel vanishes at 2h (sph_common/kernels.h:25-35); a support radius
    # of two lattice spacings is h = s. (h = 2s makes the reference's IISPH
    # relaxed-Jacobi solve diverge on this wall-bounded lattice.)
    h = s
    pos = lattice_positions(nx, ny, s, seed, jitter)
    vel = np.zeros_like(pos)
    solids = np.zeros((0, 2), np.float32)
    if with_walls:
        params = sb.Parameters(support=np.float32(h), length=1.0)
        solids = sb.world_walls_initial_solid_state(params, random_obstacles=False).positions
    return DamBreak(pos, vel, solids, float(np.float32(h)), s, float(dt),
                    lattice_mass(s, float(np.float32(h)), 1000.0))


def dam_break_12800(seed: int = 1) -> DamBreak:
    return dam_break(80, 160, 0.005, 2e-4, seed)


def dam_break_1m(seed: int = 2) -> DamBreak:
    return dam_break(1024, 1024, 1.0 / 2560.0, 1.5625e-5, seed)


def batch_dam_breaks(count: int = 256, seed: int = 4, n_min: int = 2000, n_max: int = 200_000,
                     aspect_min: float = 0.25, aspect_max: float = 2.0,
                     cfl: float = 0.4, gravity: float = 9.81) -> list:
    """Independent dam-break scenes of uneven size for the batched form."""
    u = uniform01(seed, 2 * count).reshape(count, 2)
    out = []
    for i in range(count):
        n = math.exp(math.log(n_min) + u[i, 0] * (math.log(n_max) - math.log(n_min)))
        aspect = aspect_min + u[i, 1] * (aspect_max - aspect_min)
        nx = max(1, int(round(math.sqrt(n * aspect))))
        ny = max(1, int(round(n / nx)))
        s = 0.8 / max(nx, ny)
        h = float(np.float32(s))
        dt = cfl * h / math.sqrt(2.0 * gravity * ny * s)
        out.append(dam_break(nx, ny, s, float(np.float32(dt)), seed=1000 + seed * 10007 + i))
    return out


def pcisph_denom(dt: float, mass: float, rho0: float, spacing: float, h: float) -> np.float32:
    """PCISPH correction denominator 2 (dt m / rho0)^2 (|sum gradW|^2 +
    sum |gradW|^2) over an interior square lattice (the analogue of
    config.cpp:113-121 for the lattice mass override)."""
    m = int(math.ceil(2.0 * h / spacing)) + 1
    i, j = np.meshgrid(np.arange(-m, m + 1), np.arange(-m, m + 1), indexing="ij")
    dx = spacing * i.astype(np.float64).ravel()
    dy = spacing * j.astype(np.float64).ravel()
    r = np.hypot(dx, dy)
    keep = (r > 0) & (r < 2.0 * h)
    dx, dy, r = dx[keep], dy[keep], r[keep]
    q = r / h
    k = h * h * h * float(np.float32(3.14159274101257324218750))
    dw = np.where(q <= 1.0, -3.0 * r / (h * h) + 2.25 * r * r / (h ** 3),
                  -3.0 * (2.0 - q) ** 2 / (4.0 * h)) / k
    gx, gy = dx / r * dw, dy / r * dw
    beta = 2.0 * (dt * float(mass) / rho0) ** 2
    return np.float32(beta * (gx.sum() ** 2 + gy.sum() ** 2 + np.sum(gx * gx + gy * gy)))


def config_for(scene: DamBreak, method: str = "iisph", avg_threshold: float = 0.001):
    """Simulation_config for a dam-break workload (thresholds per the benchmark:
    IISPH avg 0.1%, max 5%, 30 iterations, omega 0.5)."""
    from . import sph2d_box as sb
    params = sb.Parameters(support=scene.support, length=1.0, gravity=scene.gravity,
                           target_density=scene.rho0, method=method)
    params.iisph.error_average_threshold = avg_threshold
    params.dfsph.divergence_error_average_threshold = avg_threshold
    cfg = sb.Simulation_config.from_parameters(params)
    cfg.mass_per_particle = float(scene.mass)
    cfg.seconds_per_sub_step = float(scene.dt)
    cfg.pressure_correction_denom = float(
        pcisph_denom(scene.dt, float(scene.mass), scene.rho0, scene.spacing, scene.support))
    return cfg


def search_points(n: int, seed: int = 3, clusters: int = 8, sigma: float = 0.02) -> np.ndarray:
    """n points: n/2 uniform in [0,1)^2, n/2 from Gaussian clusters; samples
    outside [0,1)^2 are redrawn so every 32-bit cell coordinate is in range."""
    half = n // 2
    u = uniform01(seed, 2 * half).reshape(half, 2)
    pts = [u]
    rest = n - half
    centres = 0.1 + 0.8 * uniform01(seed + 1, 2 * clusters).reshape(clusters, 2)
    per = [rest // clusters + (1 if c < rest % clusters else 0) for c in range(clusters)]
    stream = seed + 2
    for c in range(clusters):
        need = per[c]
        got = []
        while need > 0:
            m = max(2 * need, 1024)
            a = uniform01(stream, 2 * m).reshape(m, 2)
            stream += 1
            u1 = np.maximum(a[:, 0], 2.0 ** -25)
            rad = np.sqrt(-2.0 * np.log(u1))
            z = np.stack([rad * np.cos(2 * np.pi * a[:, 1]), rad * np.sin(2 * np.pi * a[:, 1])], 1)
            p = centres[c] + sigma * z
            p32 = p.astype(np.float32)
            ok = np.all((p32 >= 0.0) & (p32 < 1.0), axis=1)
            p32 = p32[ok][:need]
            got.append(p32)
            need -= len(p32)
        pts.append(np.concatenate(got).astype(np.float64))
    out = np.concatenate(pts).astype(np.float32)
    return np.ascontiguousarray(out)
