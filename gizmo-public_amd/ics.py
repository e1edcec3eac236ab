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
  shear_layers      two lattices of equal-mass particles, density ~2 inside
                    |y - 0.5| < 0.25 and 1 outside, vx = +-0.5, P = 2.5
                    (the 16.8M mixed-precision case at n_out = 224)
  batch_boxes       many independent jittered-lattice boxes of log-uniform
                    particle counts (the batched case: 512 boxes, N_b in
                    [16^3, 64^3])
  plummer_clusters  1000 Plummer spheres over a 20% uniform background, about
                    four decades of density (the clustered case at 2^23)

The initial smoothing length is the radius that encloses DesNumNgb particles
at the local number density, h0 = (3 N_des / (4 pi n))^(1/3): the mean density
of the box (or of the layer) for the lattice cases, the distances to the 32nd and
33rd nearest neighbours for the clustered case.
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


def _lattice_slab(nx: int, ny: int, nz: int, y0: float, ly: float) -> np.ndarray:
    """nx * ny * nz lattice filling x, z in [0, 1) and y in [y0, y0 + ly) (wrapped)."""
    gx = (np.arange(nx, dtype=np.float64) + 0.5) / nx
    gy = y0 + (np.arange(ny, dtype=np.float64) + 0.5) * (ly / ny)
    gz = (np.arange(nz, dtype=np.float64) + 0.5) / nz
    x, y, z = np.meshgrid(gx, gy, gz, indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    pos = np.mod(pos, 1.0)
    pos[pos >= 1.0] -= 1.0
    return pos


def shear_layers(n_out: int = 224, n_in: int | None = None, pressure: float = 2.5, gamma: float = 5.0 / 3.0,
                 v_shear: float = 0.5, v_pert: float = 0.01) -> dict:
    """Dense band |y - 0.5| < 0.25 on an n_in x n_in/2 x n_in lattice, the rest
    on n_out x n_out/2 x n_out, equal masses with rho_out = 1 (so rho_in =
    (n_in / n_out)^3 ~ 2), P = 2.5, vx = +0.5 inside and -0.5 outside,
    vy = 0.01 sin(4 pi x).  n_out = 224, n_in = 282 gives 16,832,596 particles."""
    if n_in is None:
        n_in = 2 * int(round(n_out * 2.0 ** (1.0 / 3.0) / 2.0))
    inner = _lattice_slab(n_in, n_in // 2, n_in, 0.25, 0.5)
    outer = _lattice_slab(n_out, n_out // 2, n_out, 0.75, 0.5)
    pos = np.ascontiguousarray(np.concatenate([inner, outer]))
    n1, n2 = inner.shape[0], outer.shape[0]
    N = n1 + n2
    m = 0.5 / n2  # rho_out = n2 m / 0.5 = 1
    rho = np.concatenate([np.full(n1, n1 * m / 0.5), np.full(n2, 1.0)])
    vel = np.zeros((N, 3))
    vel[:n1, 0] = v_shear
    vel[n1:, 0] = -v_shear
    vel[:, 1] = v_pert * np.sin(4.0 * np.pi * pos[:, 0])
    u = pressure / ((gamma - 1.0) * rho)
    hsml = np.concatenate([np.full(n1, initial_hsml(n1, 0.5)), np.full(n2, initial_hsml(n2, 0.5))])
    return _state(pos, vel, np.full(N, m), u, hsml, f"shear_layers_{n_out}")


def batch_boxes(n_boxes: int = 512, seed: int = 4, n_min: int = 4096, n_max: int = 262144,
                jitter: float = 0.1) -> dict:
    """Independent unit boxes, box b a jittered n_b^3 lattice with
    n_b = round(exp(U(ln n_min, ln n_max))^(1/3)), m = 1 / N_b, Gaussian
    velocities with sigma ~ U(0.1, 1.0) and one u ~ log-uniform in [0.1, 10]
    per box.  Rows of box b: [box_offsets[b], box_offsets[b + 1])."""
    s = SEED_BASE + seed
    ub = uniform(s, 0, n_boxes)
    n_side = np.rint(np.exp(np.log(n_min) + ub * (np.log(n_max) - np.log(n_min))) ** (1.0 / 3.0)).astype(np.int64)
    sig = 0.1 + 0.9 * uniform(s, 1, n_boxes)
    ubox = np.exp(np.log(0.1) + uniform(s, 2, n_boxes) * (np.log(10.0) - np.log(0.1)))
    pos, vel, mass, u, hsml = [], [], [], [], []
    offsets = [0]
    for b in range(n_boxes):
        n = int(n_side[b])
        p = jittered_lattice(n, jitter, s + 1000 + b, 0)
        Nb = p.shape[0]
        pos.append(p)
        vel.append(sig[b] * gaussian(s + 1000 + b, 1, 3 * Nb).reshape(Nb, 3))
        mass.append(np.full(Nb, 1.0 / Nb))
        u.append(np.full(Nb, ubox[b]))
        hsml.append(np.full(Nb, initial_hsml(Nb)))
        offsets.append(offsets[-1] + Nb)
    out = _state(np.concatenate(pos), np.concatenate(vel), np.concatenate(mass), np.concatenate(u),
                 np.concatenate(hsml), f"batch_boxes_{n_boxes}")
    out["box_offsets"] = np.asarray(offsets, dtype=np.int64)
    out["box_sizes"] = np.ones((n_boxes, 3))
    return out


def sod_tube(n_left: int = 64, n_cross: int = 8, rho_left: float = 1.0, p_left: float = 1.0,
             rho_right: float = 0.125, p_right: float = 0.1, gamma: float = 5.0 / 3.0) -> dict:
    """Sod shock tube in a periodic box of sides (1, s, s): x in [0, 1/2) holds
    the left state on an n_left x n_cross x n_cross lattice of spacing
    d = 1/(2 n_left), x in [1/2, 1) the right state on a lattice of spacing 2d
    (density 1/8 with equal masses, so rho_right must be rho_left / 8); s =
    n_cross d.  At rest, u = p / ((gamma - 1) rho).  The periodic box holds two
    Riemann problems, at x = 1/2 and at x = 0, which do not meet before
    t ~ 0.15 (gamma 5/3)."""
    if abs(rho_right * 8.0 - rho_left) > 1e-12 * rho_left:
        raise ValueError("equal-mass lattices give a density ratio of exactly 8")
    if n_left % 2 or n_cross % 2:
        raise ValueError("n_left and n_cross must be even")
    d = 0.5 / n_left
    side = n_cross * d

    def lattice(nx, nc, x0, step):
        i, j, k = np.meshgrid(np.arange(nx), np.arange(nc), np.arange(nc), indexing="ij")
        return np.stack([x0 + (i.ravel() + 0.5) * step, (j.ravel() + 0.5) * step, (k.ravel() + 0.5) * step], axis=1)

    left = lattice(n_left, n_cross, 0.0, d)
    right = lattice(n_left // 2, n_cross // 2, 0.5, 2 * d)
    pos = np.concatenate([left, right])
    n = pos.shape[0]
    m = rho_left * d ** 3
    u = np.concatenate([np.full(len(left), p_left / ((gamma - 1) * rho_left)),
                        np.full(len(right), p_right / ((gamma - 1) * rho_right))])
    h = np.concatenate([np.full(len(left), (3 * 32 * d ** 3 / (4 * np.pi)) ** (1 / 3)),
                        np.full(len(right), (3 * 32 * 8 * d ** 3 / (4 * np.pi)) ** (1 / 3))])
    out = _state(pos, np.zeros((n, 3)), np.full(n, m), u, h, f"sod_tube_{n_left}")
    out["box"] = np.array([1.0, side, side])
    out["gamma"] = gamma
    return out


def box_slice(ic: dict, b: int) -> dict:
    """Box b of a batch_boxes set as a single-box particle set."""
    o = ic["box_offsets"]
    sl = slice(int(o[b]), int(o[b + 1]))
    out = {k: v[sl] for k, v in ic.items()
           if isinstance(v, np.ndarray) and v.shape[:1] == ic["pos"].shape[:1]}
    out["box"] = np.asarray(ic["box_sizes"][b], dtype=np.float64)
    out["name"] = f"{ic.get('name', 'batch')}[{b}]"
    return out


def local_hsml(pos: np.ndarray, des_num_ngb: int = 32, workers: int = -1) -> np.ndarray:
    """Midpoint of the distances to the des_num_ngb-th and the next nearest
    other particle (periodic unit box): about des_num_ngb neighbours, with no
    particle exactly at distance h.  (At r == h the pair test r^2 < h^2 can
    pass while r / h rounds to 1, so the kernel weights of both sides vanish
    and the volume-weighted face of compute_finitevol_faces.h:21 divides 0/0.)"""
    from scipy.spatial import cKDTree
    tree = cKDTree(pos, boxsize=1.0)
    d, _ = tree.query(pos, k=des_num_ngb + 2, workers=workers)
    return np.ascontiguousarray(0.5 * (d[:, des_num_ngb] + d[:, des_num_ngb + 1]))


def plummer_clusters(n_total: int = 8388608, n_spheres: int = 1000, seed: int = 5, f_background: float = 0.2,
                     slope: float = -1.8, m_range=(1.0, 1000.0), a_range=(1e-3, 3e-2), sigma_v: float = 0.2,
                     u0: float = 1.0) -> dict:
    """f_background of the particles uniform in the box, the rest in Plummer
    spheres with centres U(0,1)^3, masses dN/dM ~ M^slope on m_range
    (particle counts proportional to mass), scale radii log-uniform in
    a_range, radii truncated at 10 a; equal particle masses 1/N, Gaussian
    velocities, u = u0.  hsml starts between the 32nd and 33rd neighbour distances."""
    s = SEED_BASE + seed
    n_bg = int(round(f_background * n_total))
    n_cl = n_total - n_bg
    # sphere masses by inverse CDF of the power law
    q = 1.0 + slope
    lo, hi = m_range[0] ** q, m_range[1] ** q
    M = (lo + uniform(s, 0, n_spheres) * (hi - lo)) ** (1.0 / q)
    counts = np.floor(n_cl * M / M.sum()).astype(np.int64)
    rest = n_cl - counts.sum()
    order = np.argsort(-(n_cl * M / M.sum() - counts), kind="stable")
    counts[order[:rest]] += 1
    centres = uniform(s, 1, 3 * n_spheres).reshape(n_spheres, 3)
    a = np.exp(np.log(a_range[0]) + uniform(s, 2, n_spheres) * (np.log(a_range[1]) - np.log(a_range[0])))
    sph = np.repeat(np.arange(n_spheres), counts)
    x_max = 1000.0 / 101.0 ** 1.5  # enclosed-mass fraction at r = 10 a
    X = uniform(s, 3, n_cl) * x_max
    X = np.where(X <= 0.0, 1e-300, X)
    r = a[sph] / np.sqrt(X ** (-2.0 / 3.0) - 1.0)
    cos_t = 2.0 * uniform(s, 4, n_cl) - 1.0
    phi = 2.0 * np.pi * uniform(s, 5, n_cl)
    sin_t = np.sqrt(np.maximum(0.0, 1.0 - cos_t * cos_t))
    d = np.stack([r * sin_t * np.cos(phi), r * sin_t * np.sin(phi), r * cos_t], axis=1)
    pos = np.concatenate([uniform(s, 6, 3 * n_bg).reshape(n_bg, 3), centres[sph] + d])
    pos = np.mod(pos, 1.0)
    pos[pos >= 1.0] -= 1.0
    pos = np.ascontiguousarray(pos)
    N = pos.shape[0]
    vel = sigma_v * gaussian(s, 7, 3 * N).reshape(N, 3)
    hsml = local_hsml(pos)
    return _state(pos, vel, np.full(N, 1.0 / N), np.full(N, u0), hsml, f"plummer_clusters_{N}")


def random_subset(n: int, fraction: float = 0.1, seed: int = 6) -> np.ndarray:
    """Sorted random subset of round(fraction * n) indices (the active set)."""
    k = int(round(fraction * n))
    keys = uniform(SEED_BASE + seed, 0, n)
    return np.sort(np.argsort(keys, kind="stable")[:k]).astype(np.int32)


CONFIGS = {
    "lattice_gaussian": lattice_gaussian,
    "blast": blast,
    "perfect_lattice": perfect_lattice,
    "random_state": random_state,
    "shear_layers": shear_layers,
    "batch_boxes": batch_boxes,
    "plummer_clusters": plummer_clusters,
}
