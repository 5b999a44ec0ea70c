"""Seeded synthetic initial conditions for the gravity path.

* ``plummer`` -- equal-mass Plummer sphere, G = M = a = 1, radii truncated at
  ``rmax`` scale radii, isotropic velocities from the Plummer distribution
  function.  The sampling method is the Aarseth-Henon-Wielen (1974) one that the
  reference's generator uses (inputgen/plummer.cpp:65-86): inverse-CDF radius,
  von Neumann rejection on g(q) = q^2 (1 - q^2)^(7/2) for the speed in units of
  the local escape speed.  Unlike that generator it keeps standard units
  (no Henon rescaling, no second cluster) and removes the centre-of-mass
  position and velocity.
* ``halo_mixture`` -- Plummer halos with log-normal masses and scale radii
  proportional to mass^(1/3), uniform centres in a cube, plus a uniform
  background fraction; equal particle masses.

All generators draw from ``numpy.random.default_rng(seed)`` so a seed fixes the
particle set bit for bit.
"""
import numpy as np


def _isotropic(rng, n):
    cz = rng.uniform(-1.0, 1.0, n)
    ph = rng.uniform(0.0, 2.0 * np.pi, n)
    sz = np.sqrt(1.0 - cz * cz)
    return np.stack([sz * np.cos(ph), sz * np.sin(ph), cz], axis=1)


def _plummer_unit(rng, n, rmax):
    """Positions/velocities of an n-particle Plummer sphere with G=M=a=1."""
    r = np.empty(0)
    while r.size < n:
        x1 = rng.uniform(0.0, 1.0, 2 * (n - r.size) + 16)
        x1 = x1[x1 > 0.0]
        rr = 1.0 / np.sqrt(x1 ** (-2.0 / 3.0) - 1.0)
        r = np.concatenate([r, rr[rr <= rmax]])
    r = r[:n]
    q = np.empty(0)
    while q.size < n:
        m = 4 * (n - q.size) + 16
        qq = rng.uniform(0.0, 1.0, m)
        yy = rng.uniform(0.0, 0.1, m)
        q = np.concatenate([q, qq[yy < qq * qq * (1.0 - qq * qq) ** 3.5]])
    q = q[:n]
    vesc = np.sqrt(2.0) * (1.0 + r * r) ** -0.25
    pos = _isotropic(rng, n) * r[:, None]
    vel = _isotropic(rng, n) * (q * vesc)[:, None]
    return pos, vel


def plummer(n, seed=12345, rmax=10.0):
    """Equal-mass Plummer sphere.  Returns (pos[n,3], vel[n,3], mass[n])."""
    rng = np.random.default_rng(seed)
    pos, vel = _plummer_unit(rng, n, rmax)
    pos -= pos.mean(axis=0)
    vel -= vel.mean(axis=0)
    mass = np.full(n, 1.0 / n)
    return pos, vel, mass


def halo_mixture(n, n_halos=64, seed=4242, box=100.0, background=0.2, sigma=1.0,
                 a0=1.0, rmax=10.0):
    """Plummer halos (log-normal masses, a_h = a0 (M_h/<M_h>)^(1/3)) with
    centres uniform in [0, box]^3 plus a uniform background fraction.
    Returns (pos, vel, mass) with equal particle masses 1/n."""
    rng = np.random.default_rng(seed)
    n_bg = int(round(background * n))
    n_halo_total = n - n_bg
    hm = rng.lognormal(0.0, sigma, n_halos)
    counts = np.floor(hm / hm.sum() * n_halo_total).astype(np.int64)
    counts[np.argmax(hm)] += n_halo_total - counts.sum()
    centres = rng.uniform(0.0, box, (n_halos, 3))
    scale = a0 * np.cbrt(hm / hm.mean())
    pos_parts, vel_parts = [], []
    for h in range(n_halos):
        if counts[h] == 0:
            continue
        p, v = _plummer_unit(rng, int(counts[h]), rmax)
        p -= p.mean(axis=0)
        v -= v.mean(axis=0)
        mh = counts[h] / n
        pos_parts.append(centres[h] + p * scale[h])
        vel_parts.append(v * np.sqrt(mh / scale[h]))
    pos_parts.append(rng.uniform(0.0, box, (n_bg, 3)))
    vel_parts.append(np.zeros((n_bg, 3)))
    pos = np.concatenate(pos_parts)
    vel = np.concatenate(vel_parts)
    order = rng.permutation(n)
    return pos[order].copy(), vel[order].copy(), np.full(n, 1.0 / n)


def read_tipsy(path):
    """Tipsy standard binary (big-endian) reader, as src/Reader.C:11-108 loads
    it: header {double time; int nbodies, ndim, nsph, ndark, nstar; pad}, then
    gas (48 B), dark (36 B) and star (44 B) float records.  Returns a dict with
    time, pos, vel, mass, soft (gas hsmooth / dark eps / star eps) and type
    (2 gas, 3 dark, 1 star, as Particle::Type)."""
    with open(path, "rb") as f:
        hdr = np.frombuffer(f.read(32), dtype=np.dtype(">f8,>i4,>i4,>i4,>i4,>i4,>i4"))[0]
        t, nbodies, _ndim, nsph, ndark, nstar = (hdr[i] for i in range(6))
        gas = np.frombuffer(f.read(48 * nsph), dtype=np.dtype(">f4")).reshape(nsph, 12)
        dark = np.frombuffer(f.read(36 * ndark), dtype=np.dtype(">f4")).reshape(ndark, 9)
        star = np.frombuffer(f.read(44 * nstar), dtype=np.dtype(">f4")).reshape(nstar, 11)
    if nsph + ndark + nstar != nbodies:
        raise ValueError("inconsistent tipsy header")
    mass = np.concatenate([gas[:, 0], dark[:, 0], star[:, 0]]).astype(np.float64)
    pos = np.concatenate([gas[:, 1:4], dark[:, 1:4], star[:, 1:4]]).astype(np.float64)
    vel = np.concatenate([gas[:, 4:7], dark[:, 4:7], star[:, 4:7]]).astype(np.float64)
    # gas: mass pos vel rho temp hsmooth metals phi; dark: mass pos vel eps phi;
    # star: mass pos vel metals tform eps phi (Reader.C sets soft = hsmooth,
    # eps, and 0 for stars).
    soft = np.concatenate([gas[:, 9], dark[:, 7], np.zeros(nstar, np.float32)]).astype(np.float64)
    ptype = np.concatenate([np.full(nsph, 2), np.full(ndark, 3), np.full(nstar, 1)]).astype(np.int8)
    return {"time": float(t), "pos": pos, "vel": vel, "mass": mass, "soft": soft, "type": ptype}


def write_acc(path, acc):
    """ASCII accelerations as Writer::do_write writes them (src/Writer.C:31-72):
    N, then the x block, y block and z block, one '%.14g' value per line."""
    acc = np.asarray(acc)
    with open(path, "w") as f:
        f.write("%d\n" % len(acc))
        for k in range(3):
            f.write("".join("%.14g\n" % v for v in acc[:, k]))


def read_acc(path):
    """Inverse of write_acc: returns an (N, 3) float64 array."""
    v = np.loadtxt(path, dtype=np.float64)
    n = int(v[0])
    return np.stack([v[1:1 + n], v[1 + n:1 + 2 * n], v[1 + 2 * n:1 + 3 * n]], axis=1)
