"""Deterministic synthetic particle sets for the MFM hydro pass.

Every random draw is counter-based SplitMix64, so a configuration is a pure
function of (config, seed): the same arrays come out in numpy on any host.
All boxes are periodic with side 1; positions lie in [0, 1).

Configurations (sizes as used by the benchmark and the parity tests):
  lattice_gaussian  jittered n^3 lattice, Gaussian velocities, u = 1
                    (the 32^3 CPU reference case at n = 32)
  blast             jittered n^3 lattice, v = 0, u = 1e-6, total energy 1
                    deposited kernel-weighted into the 64 central particles
                    (the 128^3 plain FP64 GPU case at n = 128)
  perfect_lattice   unjittered lattice with a uniform state (invariant tests)

The initial smoothing length is the radius that encloses DesNumNgb particles
at the mean number density, h0 = (3 N_des / (4 pi n))^(1/3).
"""
from __future__ import annotations

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
SEED_BASE = 0x475A4D00


def splitmix64(counter: np.ndarray) -> np.ndarray:
    """SplitMix64 output for the given 64-bit states (vectorised)."""
    z = counter.astype(np.uint64, copy=True)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, stream: int, count: int) -> np.ndarray:
    """`count` doubles in [0, 1) from stream `stream` of seed `seed`."""
    base = np.uint64((seed * 0x100000001B3 + stream * 0x1F3D5B79) & 0xFFFFFFFFFFFFFFFF)
    k = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = splitmix64(base + k * GOLDEN)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def gaussian(seed: int, stream: int, count: int) -> np.ndarray:
    """Standard normal draws by Box-Muller from two uniform streams."""
    u1 = uniform(seed, 2 * stream + 1000, count)
    u2 = uniform(seed, 2 * stream + 1001, count)
    u1 = np.where(u1 <= 0.0, 1e-300, u1)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def initial_hsml(n_total: int, volume: float = 1.0, des_num_ngb: float = 32.0) -> float:
    number_density = n_total / volume
    return (3.0 * des_num_ngb / (4.0 * np.pi * number_density)) ** (1.0 / 3.0)


def jittered_lattice(n: int, jitter: float, seed: int, stream: int = 0) -> np.ndarray:
    """n^3 lattice at (i + 1/2)/n, each coordinate jittered by U(-jitter, jitter)/n."""
    g = (np.arange(n, dtype=np.float64) + 0.5) / n
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    if jitter > 0:
        N = pos.shape[0]
        d = (uniform(seed, stream, 3 * N).reshape(N, 3) * 2.0 - 1.0) * (jitter / n)
        pos = pos + d
    pos = np.mod(pos, 1.0)
    pos[pos >= 1.0] -= 1.0  # mod can round up to exactly 1.0
    return np.ascontiguousarray(pos)


def _state(pos, vel, mass, u, hsml, name, **extra):
    N = pos.shape[0]
    out = dict(
        name=name,
        box=np.array([1.0, 1.0, 1.0]),
        pos=np.ascontiguousarray(pos, dtype=np.float64),
        vel=np.ascontiguousarray(vel, dtype=np.float64),
        mass=np.ascontiguousarray(mass, dtype=np.float64),
        u=np.ascontiguousarray(u, dtype=np.float64),
        hsml=np.ascontiguousarray(hsml, dtype=np.float64),
        cn=np.ones(N, dtype=np.float64),  # SphP.ConditionNumber at start-up (init.c:514)
    )
    out.update(extra)
    return out


def lattice_gaussian(n: int = 32, seed: int = 1, sigma_v: float = 0.1, u0: float = 1.0,
                     jitter: float = 0.1) -> dict:
    s = SEED_BASE + seed
    pos = jittered_lattice(n, jitter, s, 0)
    N = pos.shape[0]
    vel = sigma_v * gaussian(s, 1, 3 * N).reshape(N, 3)
    mass = np.full(N, 1.0 / N)
    u = np.full(N, u0)
    hsml = np.full(N, initial_hsml(N))
    return _state(pos, vel, mass, u, hsml, f"lattice_gaussian_{n}")


def blast(n: int = 128, seed: int = 2, e_blast: float = 1.0, u0: float = 1e-6, n_hot: int = 64,
          jitter: float = 0.1) -> dict:
    s = SEED_BASE + seed
    pos = jittered_lattice(n, jitter, s, 0)
    N = pos.shape[0]
    vel = np.zeros((N, 3))
    m = 1.0 / N
    mass = np.full(N, m)
    u = np.full(N, u0)
    d = pos - 0.5
    r = np.sqrt((d * d).sum(axis=1))
    order = np.argsort(r, kind="stable")
    hot = order[:n_hot]
    h_b = r[order[n_hot]]  # distance of the (n_hot+1)-th nearest particle
    q = r[hot] / h_b
    w = np.where(q < 0.5, 1.0 + 6.0 * (q - 1.0) * q * q, 2.0 * (1.0 - q) ** 3)
    w = np.maximum(w, 0.0)
    u[hot] += e_blast * w / (m * w.sum())
    hsml = np.full(N, initial_hsml(N))
    return _state(pos, vel, mass, u, hsml, f"blast_{n}", hot=hot)


def perfect_lattice(n: int = 16, rho: float = 1.0, pressure: float = 1.0, gamma: float = 5.0 / 3.0,
                    velocity=(0.0, 0.0, 0.0)) -> dict:
    pos = jittered_lattice(n, 0.0, 0)
    N = pos.shape[0]
    vel = np.tile(np.asarray(velocity, dtype=np.float64), (N, 1))
    mass = np.full(N, rho / N)
    u = np.full(N, pressure / ((gamma - 1.0) * rho))
    hsml = np.full(N, initial_hsml(N))
    return _state(pos, vel, mass, u, hsml, f"perfect_lattice_{n}")


def random_state(n: int = 12, seed: int = 7, sigma_v: float = 0.3, jitter: float = 0.25) -> dict:
    """Jittered lattice with random velocities and a random, smooth-ish energy
    field: exercises every branch of the force loop (used by the invariant tests)."""
    s = SEED_BASE + seed
    pos = jittered_lattice(n, jitter, s, 0)
    N = pos.shape[0]
    vel = sigma_v * gaussian(s, 1, 3 * N).reshape(N, 3)
    mass = np.full(N, 1.0 / N) * (0.75 + 0.5 * uniform(s, 5, N))
    u = np.exp(1.5 * (uniform(s, 6, N) - 0.5)) * (1.0 + 0.4 * np.sin(2 * np.pi * pos[:, 0]))
    hsml = np.full(N, initial_hsml(N))
    return _state(pos, vel, mass, u, hsml, f"random_state_{n}")


CONFIGS = {
    "lattice_gaussian": lattice_gaussian,
    "blast": blast,
    "perfect_lattice": perfect_lattice,
    "random_state": random_state,
}
