"""Synthetic workloads of the benchmark configurations: planogram geometries, seeded ellipsoid
phantoms and Poisson data.  There are no datasets; every input is generated from a seed."""
import numpy as np

# (psi list, D, L, H, T) per geometry, and the volume (n, voxel mm) matched to it
GEOMETRIES = {
    # 32x32 crystals at 2 mm, 100 mm separation; 32^3 volume at 2 mm
    "panel32": dict(psis=(0.0,), D=100.0, L=64.0, H=64.0, T=2.0, n=32, voxel=2.0),
    # 128x128 crystals at 2 mm, 200 mm separation, two orientations; 128^3 volume at 2 mm
    "panel128": dict(psis=(0.0, 90.0), D=200.0, L=256.0, H=256.0, T=2.0, n=128, voxel=2.0),
    # 64x64 crystals, 200 mm separation, two orientations; 64^3 volume
    "panel64": dict(psis=(0.0, 90.0), D=200.0, L=128.0, H=128.0, T=2.0, n=64, voxel=2.0),
    "panel48": dict(psis=(0.0, 90.0), D=200.0, L=96.0, H=96.0, T=2.0, n=48, voxel=2.0),
    "panel96": dict(psis=(0.0, 90.0), D=200.0, L=192.0, H=192.0, T=2.0, n=96, voxel=2.0),
}


def setup_server(server, name):
    """Configures a PETLserver (or anything with add_planogram/set_volume) for a named geometry."""
    g = GEOMETRIES[name]
    for psi in g["psis"]:
        server.add_planogram(psi, g["D"], g["L"], g["H"], g["T"])
    server.set_volume(g["n"], g["n"], g["n"], g["voxel"])
    return g


def ellipsoid_phantom(n, voxel, n_ellipsoids, seed, background=0.1, shape=None):
    """Uniform axis-aligned ellipsoids over a constant background on an [nz][ny][nx] grid centred at the
    origin.  Centres uniform in the inner 60% of the field of view, semi-axes U[3,12] mm, values
    U[0.5,2]; later ellipsoids overwrite earlier ones."""
    nz, ny, nx = shape if shape is not None else (n, n, n)
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) - 0.5 * (nx - 1)) * voxel
    y = (np.arange(ny) - 0.5 * (ny - 1)) * voxel
    z = (np.arange(nz) - 0.5 * (nz - 1)) * voxel
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    f = np.full((nz, ny, nx), background, dtype=np.float32)
    ext = np.array([nx * voxel, ny * voxel, nz * voxel])
    for _ in range(n_ellipsoids):
        c = rng.uniform(-0.3, 0.3, size=3) * ext
        a = rng.uniform(3.0, 12.0, size=3)
        v = rng.uniform(0.5, 2.0)
        inside = ((X - c[0]) / a[0]) ** 2 + ((Y - c[1]) / a[1]) ** 2 + ((Z - c[2]) / a[2]) ** 2 <= 1.0
        f[inside] = v
    return f


def poisson_data(g_list, total_counts, seed):
    """Poisson counts with mean proportional to the projections, scaled to total_counts."""
    rng = np.random.default_rng(seed)
    tot = float(sum(np.sum(g, dtype=np.float64) for g in g_list))
    s = total_counts / tot
    return [rng.poisson(g.astype(np.float64) * s).astype(np.float32) for g in g_list]
