"""Synthetic density / activity phantoms (numpy, (NZ, NY, NX) float32, X fastest).

* test_phantom: the anisotropic heterogeneous phantom of the reference's GPU tests
  (tests/test_solver.py:143-157): 16x14x12 voxels of 1.1x1.6x2.3 mm, random density
  U(0.2, 1.8) inside an ellipsoid with an air cavity, seed 72.
* cylinder_phantom: 64^3 at 2 mm, water Z-cylinder of radius 50 mm, three hot spheres.
* thorax_phantom: 128x128x159 at 0.776x0.776x0.795 mm (68Ga workload): water elliptic body
  (45x35 mm semi-axes), two lungs at 0.30, a spine cylinder at 1.85, eight hot spheres,
  some placed on lung and bone boundaries.
* random_volume: one volume of the uneven batch (grid 96..160 x 96..160 x 80..200, voxels
  0.5-1.2 mm, 18F/11C/64Cu/68Ga, 1e6-1e8 counts, beta in [0, 0.2]): the cylinder generator
  with random ellipsoid lungs, bone inserts at 1.6-1.9 and 2-10 hot spheres.
* clinical_thorax: 344x344x127 at 2.04x2.04x2.03 mm: water body, lungs 0.26-0.32, ribs and
  spine at 1.85, a heart wall at 20x background and 10 lesions of radius 3-12 mm at 3-8:1.
Counts are applied by `poisson_counts` after the caller blurs the truth with the forward
projector (the reference's own --mode forward).
"""
import numpy as np

AIR = np.float32(0.0012)


def _grid(shape, spacing):
    """Voxel-centre coordinates in mm, origin at the volume centre: (z, y, x) arrays."""
    nz, ny, nx = shape
    dx, dy, dz = spacing
    z, y, x = np.meshgrid((np.arange(nz) - (nz - 1) / 2) * dz, (np.arange(ny) - (ny - 1) / 2) * dy,
                          (np.arange(nx) - (nx - 1) / 2) * dx, indexing="ij")
    return z, y, x


def test_phantom():
    """Returns (rho, x, y, spacing) of tests/test_solver.py:147-157 (seed 72)."""
    shape = (12, 14, 16)
    z, y, x = np.indices(shape)
    body = ((x - 7.5) / 6) ** 2 + ((y - 6.5) / 5) ** 2 + ((z - 5.5) / 5) ** 2 < 1
    cavity = (x - 7.5) ** 2 + (y - 6.5) ** 2 + (z - 5.5) ** 2 < 3
    rng = np.random.default_rng(72)
    rho = np.where(body & ~cavity, rng.uniform(0.2, 1.8, shape), 0.001).astype(np.float32)
    support = rho >= 0.05
    xa = (rng.uniform(0.1, 2, shape) * support).astype(np.float32)
    ya = (rng.uniform(0.1, 2, shape) * support).astype(np.float32)
    return rho, xa, ya, (1.1, 1.6, 2.3)


def cylinder_phantom(shape=(64, 64, 64), spacing=(2.0, 2.0, 2.0), seed=1001):
    """Water Z-cylinder r = 50 mm (air 0.0012 outside); activity 1 plus spheres of radius
    4, 6, 8 mm at 4:1. Returns (rho, truth)."""
    z, y, x = _grid(shape, spacing)
    cyl = x ** 2 + y ** 2 < 50.0 ** 2
    rho = np.where(cyl, np.float32(1.0), AIR).astype(np.float32)
    act = cyl.astype(np.float32)
    rng = np.random.default_rng(seed)
    for r in (4.0, 6.0, 8.0):
        ang = rng.uniform(0, 2 * np.pi)
        cx, cy, cz = 25.0 * np.cos(ang), 25.0 * np.sin(ang), rng.uniform(-20, 20)
        act[(x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 < r * r] = 4.0
    return rho, act


def thorax_phantom(shape=(159, 128, 128), spacing=(0.776, 0.776, 0.795), seed=1002):
    """68Ga workload phantom. Returns (rho, truth)."""
    z, y, x = _grid(shape, spacing)
    body = (x / 45.0) ** 2 + (y / 35.0) ** 2 < 1.0
    rho = np.where(body, np.float32(1.0), AIR).astype(np.float32)
    act = body.astype(np.float32)
    lungs = []
    for cx in (-22.0, 22.0):
        lung = ((x - cx) / 15.0) ** 2 + ((y + 2.0) / 20.0) ** 2 + (z / 50.0) ** 2 < 1.0
        lungs.append((cx, lung))
        rho[lung & body] = 0.30
        act[lung & body] = 0.3
    spine = (x ** 2 + (y - 24.0) ** 2 < 6.0 ** 2) & body
    rho[spine] = 1.85
    rng = np.random.default_rng(seed)
    # Eight hot spheres of radius 2-6 mm at 4-10:1; three sit on lung or bone boundaries.
    centres = [(-22.0 + 15.0, -2.0, 10.0), (22.0 - 15.0, -2.0, -12.0), (0.0, 24.0 - 6.0, 5.0)]
    while len(centres) < 8:
        cx, cy, cz = rng.uniform(-35, 35), rng.uniform(-25, 25), rng.uniform(-50, 50)
        if (cx / 40.0) ** 2 + (cy / 30.0) ** 2 < 1.0:
            centres.append((cx, cy, cz))
    for (cx, cy, cz), r, c in zip(centres, rng.uniform(2.0, 6.0, 8), rng.uniform(4.0, 10.0, 8)):
        sph = ((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 < r * r) & body
        act[sph] = np.float32(c) * np.where(rho[sph] < 0.5, np.float32(0.3), np.float32(1.0))
    return rho, act


def poisson_counts(blurred, total, seed):
    """Poisson-noised data with `total` expected counts: rng.poisson(b * total / sum(b))."""
    b = np.asarray(blurred, np.float64)
    lam = b * (total / b.sum())
    return np.random.default_rng(seed).poisson(lam).astype(np.float32)


def _ellipsoid(z, y, x, centre, semi):
    cz, cy, cx = centre
    az, ay, ax = semi
    return ((x - cx) / ax) ** 2 + ((y - cy) / ay) ** 2 + ((z - cz) / az) ** 2 < 1.0


def random_volume(rng):
    """One volume of the uneven batch. Returns a dict with shape (NZ, NY, NX), spacing
    (DX, DY, DZ), isotope, counts, beta, rho and truth."""
    nx, ny = (int(v) for v in rng.integers(96, 161, 2))
    nz = int(rng.integers(80, 201))
    sxy, sz = float(rng.uniform(0.5, 1.2)), float(rng.uniform(0.5, 1.2))
    shape, spacing = (nz, ny, nx), (sxy, sxy, sz)
    z, y, x = _grid(shape, spacing)
    half = 0.5 * min(nx, ny) * sxy
    radius = rng.uniform(0.7, 0.9) * half
    body = x ** 2 + y ** 2 < radius ** 2
    rho = np.where(body, np.float32(1.0), AIR).astype(np.float32)
    act = body.astype(np.float32)
    zext = 0.5 * nz * sz
    for side in (-1.0, 1.0):  # two random ellipsoid lungs
        semi = (rng.uniform(0.3, 0.6) * zext, rng.uniform(0.3, 0.5) * radius, rng.uniform(0.2, 0.3) * radius)
        centre = (rng.uniform(-0.2, 0.2) * zext, rng.uniform(-0.2, 0.1) * radius, side * rng.uniform(0.35, 0.5) * radius)
        lung = _ellipsoid(z, y, x, centre, semi) & body
        rho[lung] = np.float32(rng.uniform(0.25, 0.35))
        act[lung] = 0.3
    for _ in range(int(rng.integers(1, 4))):  # bone inserts: Z-cylinders at 1.6-1.9
        r = rng.uniform(0.05, 0.12) * radius
        cx, cy = rng.uniform(-0.5, 0.5, 2) * radius
        bone = ((x - cx) ** 2 + (y - cy) ** 2 < r * r) & body
        rho[bone] = np.float32(rng.uniform(1.6, 1.9))
    for _ in range(int(rng.integers(2, 11))):  # hot spheres of radius 1.5-6 mm at 2-8:1
        r = rng.uniform(1.5, 6.0)
        cz, cy, cx = rng.uniform(-0.8, 0.8) * zext, *(rng.uniform(-0.7, 0.7, 2) * radius)
        act[((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 < r * r) & body] = np.float32(rng.uniform(2.0, 8.0))
    return {"shape": shape, "spacing": spacing, "isotope": str(rng.choice(["18F", "11C", "64Cu", "68Ga"])),
            "counts": float(10 ** rng.uniform(6, 8)), "beta": float(rng.uniform(0.0, 0.2)), "rho": rho, "truth": act}


def uneven_batch(n=12, seed=1004):
    """The uneven batch: n random volumes from one seed."""
    rng = np.random.default_rng(seed)
    return [random_volume(rng) for _ in range(n)]


def clinical_thorax(shape=(127, 344, 344), spacing=(2.04, 2.04, 2.03), seed=1005):
    """Clinical-scale thorax. Returns (rho, truth)."""
    z, y, x = _grid(shape, spacing)
    rng = np.random.default_rng(seed)
    ax, ay = 165.0, 115.0  # body ellipse semi-axes [mm]; y grows towards the back
    body = (x / ax) ** 2 + (y / ay) ** 2 < 1.0
    rho = np.where(body, np.float32(1.0), AIR).astype(np.float32)
    act = body.astype(np.float32)
    for side in (-1.0, 1.0):
        lung = _ellipsoid(z, y, x, (10.0, -5.0, side * 72.0), (105.0, 78.0, 55.0)) & body
        rho[lung] = np.float32(rng.uniform(0.26, 0.32))
        act[lung] = 0.3
    # Ribs: bands of the elliptical shell 8-16 mm inside the skin, 10 mm wide every 25 mm in z,
    # over the sides and the back; the spine: a cylinder behind the mediastinum.
    shell = ((x / (ax - 8.0)) ** 2 + (y / (ay - 8.0)) ** 2 < 1.0) & ((x / (ax - 16.0)) ** 2 + (y / (ay - 16.0)) ** 2 >= 1.0)
    ribs = shell & (np.mod(z + 200.0, 25.0) < 10.0) & (y > -0.5 * ay)
    spine = x ** 2 + (y - 80.0) ** 2 < 16.0 ** 2
    rho[(ribs | spine) & body] = np.float32(1.85)
    # Heart: ellipsoid left of the midline, in front; its 9 mm wall at 20x background.
    centre, semi = (-20.0, -35.0, 28.0), (50.0, 45.0, 55.0)
    outer = _ellipsoid(z, y, x, centre, semi)
    inner = _ellipsoid(z, y, x, centre, tuple(s - 9.0 for s in semi))
    act[outer & ~inner & body] = 20.0
    act[inner & body] = 2.0  # blood pool
    placed = 0
    while placed < 10:  # lesions of radius 3-12 mm at 3-8:1, inside the body
        r = rng.uniform(3.0, 12.0)
        cx, cy, cz = rng.uniform(-ax, ax), rng.uniform(-ay, ay), rng.uniform(-110.0, 110.0)
        if (cx / (ax - r - 10)) ** 2 + (cy / (ay - r - 10)) ** 2 >= 1.0:
            continue
        act[((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 < r * r) & body] = np.float32(rng.uniform(3.0, 8.0))
        placed += 1
    return rho, act
