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
import os

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


# Tipsy record layouts (big-endian, as the utility's TipsyFile reads and
# writes them): mass, pos[3], vel[3] and then the per-type float fields.
# Positions and velocities are float, or double in a file written with
# double positions and velocities (TipsyWriter's use_double, src/Writer.C:110-116).
_TIPSY_HEADER = np.dtype(">f8,>i4,>i4,>i4,>i4,>i4,>i4")
_TIPSY_TAIL = {"gas": ("rho", "temp", "hsmooth", "metals", "phi"),
               "dark": ("eps", "phi"),
               "star": ("metals", "tform", "eps", "phi")}
# Particle::Type values (src/Particle.h:25-30)
TIPSY_GAS, TIPSY_DARK, TIPSY_STAR = 2, 3, 1


def _tipsy_dtype(kind, double):
    v = ">f8" if double else ">f4"
    fields = [("mass", ">f4"), ("pos", v, (3,)), ("vel", v, (3,))]
    fields += [(name, ">f4") for name in _TIPSY_TAIL[kind]]
    return np.dtype(fields)


def read_tipsy(path):
    """Tipsy standard binary (big-endian) reader, as src/Reader.C:11-108 loads
    it: header {double time; int nbodies, ndim, nsph, ndark, nstar; pad}, then
    the gas, dark and star records.  Float records (48 / 36 / 44 bytes) and
    records with double positions and velocities (72 / 60 / 68 bytes) are
    told apart by the file size.  Returns a dict with time, pos, vel, mass,
    soft (gas hsmooth, dark eps, 0 for stars as Reader.C sets them), type
    (Particle::Type: 2 gas, 3 dark, 1 star), order (the file index,
    Particle::order) and double (the record layout found)."""
    size = os.path.getsize(path)
    with open(path, "rb") as f:
        hdr = np.frombuffer(f.read(32), dtype=_TIPSY_HEADER)[0]
        t, nbodies, _ndim, nsph, ndark, nstar = (hdr[i] for i in range(6))
        if nsph + ndark + nstar != nbodies or min(nsph, ndark, nstar) < 0:
            raise ValueError("inconsistent tipsy header")
        double = None
        for d in (False, True):
            body = sum(c * _tipsy_dtype(k, d).itemsize
                       for c, k in ((nsph, "gas"), (ndark, "dark"), (nstar, "star")))
            if size == 32 + body:
                double = d
                break
        if double is None:
            raise ValueError("tipsy file size %d matches neither float nor double records" % size)
        gas = np.frombuffer(f.read(nsph * _tipsy_dtype("gas", double).itemsize),
                            dtype=_tipsy_dtype("gas", double))
        dark = np.frombuffer(f.read(ndark * _tipsy_dtype("dark", double).itemsize),
                             dtype=_tipsy_dtype("dark", double))
        star = np.frombuffer(f.read(nstar * _tipsy_dtype("star", double).itemsize),
                             dtype=_tipsy_dtype("star", double))
    recs = (gas, dark, star)
    mass = np.concatenate([r["mass"] for r in recs]).astype(np.float64)
    pos = np.concatenate([r["pos"].reshape(-1, 3) for r in recs]).astype(np.float64)
    vel = np.concatenate([r["vel"].reshape(-1, 3) for r in recs]).astype(np.float64)
    soft = np.concatenate([gas["hsmooth"], dark["eps"],
                           np.zeros(nstar, np.float32)]).astype(np.float64)
    ptype = np.concatenate([np.full(nsph, TIPSY_GAS), np.full(ndark, TIPSY_DARK),
                            np.full(nstar, TIPSY_STAR)]).astype(np.int8)
    return {"time": float(t), "pos": pos, "vel": vel, "mass": mass, "soft": soft,
            "type": ptype, "order": np.arange(nbodies, dtype=np.int64), "double": double}


def tipsy_snapshot_name(prefix, iteration):
    """TipsyWriter::do_write's file name (src/Writer.C:112)."""
    return "%s.%d.tipsy" % (prefix, iteration)


def write_tipsy(path, pos, vel, mass, ptype, time=0.0, order=None, use_double=True):
    """TipsyWriter::write / do_write (src/Writer.C:87-159): the particles
    sorted by Particle::order, one record per particle by its type with mass,
    position and velocity set and the other fields zero (the reference fills
    only those three), positions and velocities in double when use_double
    (the reference's choice when Real is double).  The header counts gas,
    dark and star particles; a Tipsy file holds the gas records first, then
    dark, then star, so the types must come in that order once sorted."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    vel = np.asarray(vel, dtype=np.float64).reshape(-1, 3)
    mass = np.asarray(mass, dtype=np.float64)
    ptype = np.asarray(ptype)
    n = len(mass)
    idx = np.arange(n) if order is None else np.argsort(np.asarray(order), kind="stable")
    ptype = ptype[idx]
    rank = np.select([ptype == TIPSY_GAS, ptype == TIPSY_DARK, ptype == TIPSY_STAR], [0, 1, 2], -1)
    if np.any(rank < 0):
        raise ValueError("particle type outside gas/dark/star")
    if np.any(np.diff(rank) < 0):
        raise ValueError("tipsy records must be gas, then dark, then star in particle order")
    counts = [int(np.count_nonzero(rank == r)) for r in range(3)]
    hdr = np.array([(float(time), n, 3, counts[0], counts[1], counts[2], 0)], dtype=_TIPSY_HEADER)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        start = 0
        for kind, c in zip(("gas", "dark", "star"), counts):
            sel = idx[start:start + c]
            start += c
            rec = np.zeros(c, dtype=_tipsy_dtype(kind, use_double))
            rec["mass"] = mass[sel]
            rec["pos"] = pos[sel]
            rec["vel"] = vel[sel]
            f.write(rec.tobytes())


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
