"""Synthetic initial conditions for the benchmark and test workloads.

All generators use numpy's PCG64 with the workload's seed, so every machine
regenerates identical particles.  Positions are FP64.
"""
import numpy as np


def uniform_ball(n, seed=0, radius=0.5, center=(0.5, 0.5, 0.5)):
    """n points uniform in a ball (r = R U^(1/3), isotropic direction)."""
    rng = np.random.default_rng(seed)
    u = rng.random(n)
    d = rng.standard_normal((3, n))
    d /= np.linalg.norm(d, axis=0)
    r = radius * np.cbrt(u)
    pos = np.asarray(center, np.float64)[:, None] + r * d
    return pos[0].copy(), pos[1].copy(), pos[2].copy()


def lattice(n_side):
    g = (np.arange(n_side, dtype=np.float64) + 0.5) / n_side
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    return X.ravel(), Y.ravel(), Z.ravel()


def zeldovich_field(n_side, seed, rms, spectral_index=-2.0, a0=0.02):
    """Lattice positions q (3, N), Zel'dovich displacements psi (3, N) of a
    Gaussian field with P(k) ~ k^n (psi_k = i k delta_k / k^2) scaled to the
    given RMS, and the growing-mode velocity factor v = vfac * psi in the
    drift convention dx/da = Cx v with Cx = sqrt(a) (Einstein-de Sitter),
    vfac = a0^-1.5."""
    rng = np.random.default_rng(seed)
    n = n_side
    w = rng.standard_normal((n, n, n))
    W = np.fft.rfftn(w)
    k1 = np.fft.fftfreq(n, d=1.0 / n) * 2 * np.pi
    kz = np.fft.rfftfreq(n, d=1.0 / n) * 2 * np.pi
    KX, KY, KZ = np.meshgrid(k1, k1, kz, indexing="ij")
    k2 = KX * KX + KY * KY + KZ * KZ
    k2[0, 0, 0] = 1.0
    amp = np.power(k2, spectral_index / 4.0)  # sqrt(P(k)) = k^(n/2)
    amp[0, 0, 0] = 0.0
    D = W * amp
    psi = []
    for K in (KX, KY, KZ):
        psi.append(np.fft.irfftn(1j * K * D / k2, s=(n, n, n), axes=(0, 1, 2)).ravel())
    psi = np.stack(psi)
    psi *= rms / np.sqrt(np.mean(np.sum(psi * psi, axis=0)))
    q = np.stack(lattice(n))
    return q, psi, 1.0 / (a0 * np.sqrt(a0))


def zeldovich(n_side, seed, rms, spectral_index=-2.0, a0=0.02):
    """Perturbed unit-cube lattice (zeldovich_field): returns x, y, z, vx, vy,
    vz, positions periodically wrapped into [0,1)."""
    q, psi, vfac = zeldovich_field(n_side, seed, rms, spectral_index, a0)
    x, y, z = (np.mod(q[d] + psi[d], 1.0) for d in range(3))
    return x, y, z, psi[0] * vfac, psi[1] * vfac, psi[2] * vfac


def plummer_clusters(x, y, z, vx, vy, vz, seed, n_spheres=64, scale=0.005, fraction=0.2,
                     rmax_factor=30.0):
    """Moves `fraction` of the particles into n_spheres Plummer spheres
    (r = b / sqrt(U^(-2/3) - 1), truncated at rmax_factor * b) with uniform
    centres; moved particles keep zero velocity.  Periodically wrapped."""
    rng = np.random.default_rng(seed + 1000)
    n = len(x)
    k = int(round(fraction * n))
    sel = rng.choice(n, size=k, replace=False)
    cen = rng.random((n_spheres, 3))
    which = rng.integers(0, n_spheres, k)
    umax = (1.0 + 1.0 / rmax_factor ** 2) ** -1.5
    u = rng.random(k) * umax
    r = scale / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    d = rng.standard_normal((3, k))
    d /= np.linalg.norm(d, axis=0)
    x, y, z, vx, vy, vz = [a.copy() for a in (x, y, z, vx, vy, vz)]
    x[sel] = np.mod(cen[which, 0] + r * d[0], 1.0)
    y[sel] = np.mod(cen[which, 1] + r * d[1], 1.0)
    z[sel] = np.mod(cen[which, 2] + r * d[2], 1.0)
    vx[sel] = 0.0
    vy[sel] = 0.0
    vz[sel] = 0.0
    return x, y, z, vx, vy, vz


def perturbed_lattice(n_side, seed, sigma):
    """Lattice with Gaussian jitter (sigma in box units), wrapped."""
    rng = np.random.default_rng(seed)
    x, y, z = lattice(n_side)
    n = len(x)
    return (np.mod(x + sigma * rng.standard_normal(n), 1.0),
            np.mod(y + sigma * rng.standard_normal(n), 1.0),
            np.mod(z + sigma * rng.standard_normal(n), 1.0))


def clustered(n, seed, n_clusters=4, scale=0.01, fraction=0.5):
    """Uniform background plus Plummer clusters (small tests)."""
    rng = np.random.default_rng(seed)
    x, y, z = rng.random((3, n))
    z0 = np.zeros(n)
    x, y, z, _, _, _ = plummer_clusters(x, y, z, z0, z0, z0, seed, n_spheres=n_clusters,
                                        scale=scale, fraction=fraction)
    return x, y, z


def write_vtk(path, x, y, z, vx, vy, vz, ids=None, unit_velocity=1.0):
    """Writes particles in the reference's IC layout (initialization.cpp:34-83):
    4 header lines, POINTS, then v_x, v_y, v_z and ID blocks, velocities
    multiplied by UnitVelocity_in_cm_per_s (the reader divides by it).  Values
    are written with 17 significant digits so they read back exactly."""
    n = len(x)
    ids = np.arange(n) if ids is None else ids
    with open(path, "w") as f:
        f.write("# vtk DataFile Version 3.0\nsynthetic initial conditions\nASCII\nDATASET POLYDATA\n")
        f.write("POINTS %d double\n" % n)
        np.savetxt(f, np.stack([x, y, z], 1), fmt="%.17g")
        f.write("\nPOINT_DATA %d\n" % n)
        for name, v in (("v_x", vx), ("v_y", vy), ("v_z", vz)):
            if name != "v_x":
                f.write("\n")
            f.write("SCALARS %s double\nLOOKUP_TABLE default\n" % name)
            np.savetxt(f, np.asarray(v) * unit_velocity, fmt="%.17g")
        f.write("\nSCALARS ID int\nLOOKUP_TABLE default\n")
        np.savetxt(f, ids, fmt="%d")


PARAM_TAGS = [
    "Omega", "OmegaLambda", "Redshift", "HubbleParam", "GravitionalConstant", "UnitLength_in_cm",
    "UnitMass_in_g", "UnitVelocity_in_cm_per_s", "ICFile", "ICFileType", "FileBase", "OutputFile",
    "BoundingBox_min_x", "BoundingBox_max_x", "BoundingBox_min_y", "BoundingBox_max_y",
    "BoundingBox_min_z", "BoundingBox_max_z", "Periodic", "DeltaA", "EndRedShift",
    "MassPerParticle", "MaxOctreeLevel", "CellSizeParameter", "NumberOfNeighbors",
    "DensityEstimator",
]


def write_param(path, **tags):
    """A .param file in the reference's Tag value %comment format."""
    with open(path, "w") as f:
        f.write("% written by apcloud.ic.write_param\n")
        for k, v in tags.items():
            f.write("%-28s %s\n" % (k, repr(v) if isinstance(v, float) else v))


MULTIBOX_EDGES = (16, 24, 32, 48, 64, 96)


def multibox_specs(nbox=48, seed=3, rms=0.01, edges=MULTIBOX_EDGES, index_range=(-2.5, -1.5)):
    """Independent periodic boxes for the multi-box form: lattice side drawn
    uniformly from `edges`, spectral index uniformly from `index_range`, and
    a per-box seed, all from one generator seeded with `seed`.  Returns a list
    of dicts (n_side, spectral_index, seed, rms); `multibox_box` realises one."""
    rng = np.random.default_rng(seed)
    sides = rng.choice(np.asarray(edges), size=nbox)
    index = rng.uniform(index_range[0], index_range[1], size=nbox)
    seeds = rng.integers(0, 2 ** 31 - 1, size=nbox)
    return [dict(n_side=int(sides[b]), spectral_index=float(index[b]), seed=int(seeds[b]),
                 rms=rms) for b in range(nbox)]


def multibox_box(spec):
    """x, y, z, vx, vy, vz of one multibox_specs entry (a Zel'dovich lattice)."""
    return zeldovich(spec["n_side"], spec["seed"], spec["rms"], spec["spectral_index"])
