"""Seeded synthetic inputs for the field engine (site coordinates, packed
Cartesian multipoles, exclusion lists, damping factors).

Multipoles are packed q, x, y, z, xx, xy, xz, yy, yz, zz (peter-reinholdt/fmm
README.md:38, src/fmm.F90:80-89).  Exclusion rows are 1-based, zero-padded,
and always list the site itself first: neither the reference direct sum
(src/fmm.F90:798-800) nor its near field (src/fmm.F90:704-736) skips j == i,
so a site that does not exclude itself gets a non-finite field.
All arrays are returned Fortran-ordered (column-major), as the reference takes
them.
"""
import numpy as np

WATER_DENSITY = 0.0334          # molecules / A^3
OH_BOND = 0.96                  # A
HOH_ANGLE = np.deg2rad(104.5)
Q_O, Q_H = -0.834, 0.417
ALPHA_O, ALPHA_H = 1.0, 0.5     # A^3
THOLE_A = 0.39


def ncomp(k):
    return (k + 1) * (k + 2) * (k + 3) // 6


def traceless(mult):
    """Remove the trace of the quadrupole block (columns 4, 7, 9) in place."""
    if mult.shape[1] >= 10:
        tr = (mult[:, 4] + mult[:, 7] + mult[:, 9]) / 3.0
        mult[:, 4] -= tr
        mult[:, 7] -= tr
        mult[:, 9] -= tr
    return mult


def self_exclusions(n):
    return np.asfortranarray(np.arange(1, n + 1, dtype=np.int32).reshape(n, 1))


def random_multipoles(rng, n, dim, q_sigma=0.5, mu_sigma=0.1, theta_sigma=0.05):
    """Charges N(0,q_sigma) shifted to sum zero, dipoles N(0,mu_sigma),
    quadrupoles N(0,theta_sigma) made traceless (higher orders N(0,theta_sigma))."""
    m = np.zeros((n, dim))
    m[:, 0] = rng.normal(0.0, q_sigma, n)
    m[:, 0] -= m[:, 0].mean()
    if dim > 1:
        m[:, 1:4] = rng.normal(0.0, mu_sigma, (n, 3))
    if dim > 4:
        m[:, 4:] = rng.normal(0.0, theta_sigma, (n, dim - 4))
    return np.asfortranarray(traceless(m))


def uniform_box(n, side, dim, seed):
    rng = np.random.default_rng(seed)
    x = np.asfortranarray(rng.uniform(0.0, side, (n, 3)))
    m = random_multipoles(rng, n, dim)
    return x, m


def _haar_rotations(rng, k):
    q = rng.normal(size=(k, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((k, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def water_box(n_mol, seed, jitter=0.25):
    """Rigid 3-site waters (O, H, H) on a jittered cubic grid at liquid
    density with Haar-random orientations.  Returns coordinates (3 n_mol x 3),
    charges (3 n_mol), exclusions (3 n_mol x 3: self + the two partners) and
    Thole damping factors sqrt(0.39) / alpha^(1/6)."""
    rng = np.random.default_rng(seed)
    a = (1.0 / WATER_DENSITY) ** (1.0 / 3.0)
    side = int(np.ceil(n_mol ** (1.0 / 3.0)))
    g = np.stack(np.meshgrid(np.arange(side), np.arange(side), np.arange(side), indexing="ij"), -1).reshape(-1, 3)
    centres = g[:n_mol] * a + rng.uniform(-jitter, jitter, (n_mol, 3))
    half = HOH_ANGLE / 2.0
    body = np.array([[0.0, 0.0, 0.0],
                     [OH_BOND * np.sin(half), 0.0, OH_BOND * np.cos(half)],
                     [-OH_BOND * np.sin(half), 0.0, OH_BOND * np.cos(half)]])
    R = _haar_rotations(rng, n_mol)
    sites = centres[:, None, :] + np.einsum("kij,sj->ksi", R, body)
    n = 3 * n_mol
    x = np.asfortranarray(sites.reshape(n, 3))
    q = np.tile([Q_O, Q_H, Q_H], n_mol)
    base = 3 * np.arange(n_mol)
    excl = np.zeros((n, 3), dtype=np.int32)
    for s in range(3):
        others = [t for t in range(3) if t != s]
        excl[base + s, 0] = base + s + 1
        excl[base + s, 1] = base + others[0] + 1
        excl[base + s, 2] = base + others[1] + 1
    alpha = np.tile([ALPHA_O, ALPHA_H, ALPHA_H], n_mol)
    df = np.sqrt(THOLE_A) / alpha ** (1.0 / 6.0)
    return x, q, np.asfortranarray(excl), df, rng


def gaussian_clusters(n_clusters, per_cluster, side, sigma, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, side, (n_clusters, 3))
    x = c[:, None, :] + rng.normal(0.0, sigma, (n_clusters, per_cluster, 3))
    return np.asfortranarray(x.reshape(-1, 3)), rng


def config(idx, scale=1.0):
    """The five benchmark workloads.  `scale` < 1 shrinks the site count
    (for tests) keeping every other parameter.  Returns a dict of the call
    arguments of field_fmm plus 'name'; idx 4 (many systems) returns the
    per-system arrays under 'systems'."""
    seed = 1000 + idx
    if idx == 0:
        n = max(int(4096 * scale), 8)
        x, m = uniform_box(n, 40.0 * scale ** (1.0 / 3.0), 10, seed)
        return dict(name="uniform-4096-quadrupole", coordinates=x, multipoles=m,
                    exclusions=self_exclusions(n), theta=0.5, ncrit=64, expansion_order=5,
                    field_order=1, damp_type="", damping_factors=np.zeros(0), sets=1)
    if idx in (1, 2):
        n_mol = max(int(66667 * scale), 2)
        x, q, excl, df, rng = water_box(n_mol, 1001)
        n = x.shape[0]
        if idx == 1:
            m = np.zeros((n, 4))
            m[:, 0] = q
            m[:, 1:4] = rng.normal(0.0, 0.05, (n, 3))
            return dict(name="water-200k-thole-dipole", coordinates=x, multipoles=np.asfortranarray(m),
                        exclusions=excl, theta=0.5, ncrit=64, expansion_order=6, field_order=2,
                        damp_type="THOLE", damping_factors=df, sets=1)
        rng = np.random.default_rng(seed)
        sets = 32
        m = rng.normal(0.0, 0.1, (n, 10, sets))
        for k in range(sets):
            traceless(m[:, :, k])
        return dict(name="water-200k-amoeba-quadrupole-32sets", coordinates=x,
                    multipoles=np.asfortranarray(m), exclusions=excl, theta=0.5, ncrit=64,
                    expansion_order=6, field_order=1, damp_type="AMOEBA", damping_factors=df, sets=sets)
    if idx == 3:
        per = 512
        ncl = max(int(2048 * scale), 1)
        x, rng = gaussian_clusters(ncl, per, 300.0 * scale ** (1.0 / 3.0), 3.0, seed)
        n = x.shape[0]
        m = random_multipoles(rng, n, 10)
        df = rng.uniform(0.8, 1.2, n)
        return dict(name="clusters-1M-erf-quadrupole", coordinates=x, multipoles=m,
                    exclusions=self_exclusions(n), theta=0.6, ncrit=64, expansion_order=8,
                    field_order=1, damp_type="ERF", damping_factors=df, sets=1)
    if idx == 4:
        # An uneven batch: 64 water boxes built as the default water box, sizes log-uniform between 10k and
        # 500k sites (times scale), THOLE, 8 multipole sets each (the water charges plus per-set dipoles
        # N(0, 0.05)), field order 1, learned M2L orders within a cap of 10 at a 1e-6 tolerance.  theta is not
        # given for this workload: at the default water box's 0.5 even order 10 leaves ~3e-5 of far-field
        # truncation error, and at 0.3 the largest THOLE systems still miss 1e-6 at order 10 (spot errors 1.0e-6 at 211k and 2.2e-6 at 494k
        # sites: damping shrinks the near field the error is relative to), so a 1e-6 tolerance could only fall
        # back there; 0.25 meets it on every system with learned orders (mean order 7-8).
        sizes = uneven_sizes(scale)
        systems = [uneven_system(k, size) for k, size in enumerate(sizes)]
        return dict(name="water-uneven-64-systems-8sets-thole", systems=systems, theta=0.25, ncrit=64,
                    expansion_order=10, field_order=1, damp_type="THOLE", sets=UNEVEN_SETS, order_tolerance=1e-6)
    raise ValueError(f"unknown config {idx}")


UNEVEN_SYSTEMS, UNEVEN_SETS = 64, 8


def uneven_sizes(scale=1.0):
    """Site counts (before rounding to whole molecules) of the uneven batch's systems."""
    rng = np.random.default_rng(1004)
    return np.exp(rng.uniform(np.log(10_000.0), np.log(500_000.0), UNEVEN_SYSTEMS)) * scale


def uneven_system(k, size):
    """System k of the uneven batch (a water box of about `size` sites, 8 dipole sets)."""
    x, q, excl, df, r2 = water_box(max(int(size / 3), 2), 1004 * 100 + k)
    n = x.shape[0]
    m = np.zeros((n, 4, UNEVEN_SETS), order="F")
    for u in range(UNEVEN_SETS):
        m[:, 0, u] = q
        m[:, 1:4, u] = r2.normal(0.0, 0.05, (n, 3))
    return dict(coordinates=x, multipoles=m, exclusions=excl, damping_factors=df)
