"""Synthetic molecules, bases and densities of the benchmark configurations.

* Water clusters: oxygens uniform in a cubic box with seeded rejection for
  O-O >= 2.7 A; each H 0.96 A from its O, H-O-H angle 104.5 deg, random
  orientation.  Coordinates are returned in bohr.
* Even-tempered uncontracted bases: alpha_k = alpha_0 * beta^k.  Shells with
  l >= 2 are spherical, l <= 1 Cartesian (GauXC's convention for parsed
  bases, tests/basis/parse_basis.cxx:220).
* Density P = 2 C C^T with C an nbf x width seeded normal matrix
  orthonormalised by QR.

Configurations (sizes from the recipes):
  c0  3 waters, 6 A box, O 4s3p1d / H 2s1p, beta 3 (O) / 4 (H)   -> 84 bf
      FineGrid (75 x 302), SVWN5
  c1  24 waters, 10 A box, O 5s4p2d1f / H 3s2p, beta 2.5          -> 1248 bf
      UltraFineGrid (99 x 590), PBE, screening 1e-10
  c4  150 waters, 17 A box, c1 basis                              -> 7800 bf
The c1 recipe names beta = 2.5 but no alpha_0; this module uses c0's
alpha_0 per angular momentum (s 0.2, p 0.3, d 0.8 for O; s 0.15, p 0.7 for
H) and 1.0 for the O f shell.
"""
import numpy as np

ANGSTROM = 1.8897261254578281  # bohr per angstrom

# (l, alpha_0, beta, count) per element
BASIS_C0 = {8: [(0, 0.2, 3.0, 4), (1, 0.3, 3.0, 3), (2, 0.8, 3.0, 1)],
            1: [(0, 0.15, 4.0, 2), (1, 0.7, 4.0, 1)]}
BASIS_C1 = {8: [(0, 0.2, 2.5, 5), (1, 0.3, 2.5, 4), (2, 0.8, 2.5, 2), (3, 1.0, 2.5, 1)],
            1: [(0, 0.15, 2.5, 3), (1, 0.7, 2.5, 2)]}


def water_cluster(n_water, box_angstrom, seed):
    """(Z, xyz[bohr]) of n_water waters placed as described above."""
    rng = np.random.default_rng(seed)
    oxy = []
    tries = 0
    while len(oxy) < n_water:
        tries += 1
        if tries > 200000:
            raise RuntimeError("could not place waters; box too small")
        p = rng.uniform(0.0, box_angstrom, 3)
        if all(np.linalg.norm(p - q) >= 2.7 for q in oxy):
            oxy.append(p)
    Z, xyz = [], []
    half = np.deg2rad(104.5) / 2
    for o in oxy:
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        u = q @ np.array([np.sin(half), 0.0, np.cos(half)])
        v = q @ np.array([-np.sin(half), 0.0, np.cos(half)])
        Z += [8, 1, 1]
        xyz += [o, o + 0.96 * u, o + 0.96 * v]
    return np.array(Z, dtype=np.int64), np.array(xyz) * ANGSTROM


def even_tempered_basis(Z, xyz, recipe):
    """Shell arrays (l, pure, alpha lists, coeff lists, origins), unnormalised."""
    l, pure, alpha, coeff, origin = [], [], [], [], []
    for z, r in zip(Z, xyz):
        for (am, a0, beta, count) in recipe[int(z)]:
            for k in range(count):
                l.append(am)
                pure.append(am >= 2)
                alpha.append([a0 * beta ** k])
                coeff.append([1.0])
                origin.append(r)
    return l, pure, alpha, coeff, origin


def nbf_of(l, pure):
    return int(sum(2 * a + 1 if p else (a + 1) * (a + 2) // 2 for a, p in zip(l, pure)))


def density(nbf, width, seed):
    rng = np.random.default_rng(seed + 1)
    C, _ = np.linalg.qr(rng.normal(size=(nbf, width)))
    return np.asfortranarray(2.0 * C @ C.T)


CONFIGS = {
    "c0": dict(n_water=3, box=6.0, recipe=BASIS_C0, grid="FineGrid", func="SVWN5", width=15,
               tol=1e-10),
    "c1": dict(n_water=24, box=10.0, recipe=BASIS_C1, grid="UltraFineGrid", func="PBE",
               width=120, tol=1e-10),
    "c4": dict(n_water=150, box=17.0, recipe=BASIS_C1, grid="UltraFineGrid", func="PBE",
               width=750, tol=1e-10),
}


def make_config(name, seed=0, n_water=None, grid=None):
    """Build (mol, basis, P, cfg) of a configuration through the product API."""
    from . import api
    cfg = dict(CONFIGS[name])
    if n_water is not None:
        cfg["n_water"] = n_water
    if grid is not None:
        cfg["grid"] = grid
    Z, xyz = water_cluster(cfg["n_water"], cfg["box"], seed)
    l, pure, alpha, coeff, origin = even_tempered_basis(Z, xyz, cfg["recipe"])
    mol = api.Molecule(Z, xyz)
    basis = api.BasisSet(l, pure, alpha, coeff, origin, normalize=True,
                         shell_tolerance=cfg["tol"])
    nbf = basis.nbf()
    width = cfg["width"] if n_water is None else max(1, min(nbf, 5 * cfg["n_water"]))
    P = density(nbf, width, seed)
    cfg.update(Z=Z, xyz=xyz, nbf=nbf)
    return mol, basis, P, cfg


def build_tasks(mol, basis, grid="UltraFineGrid", batch_size=512, ex="Host"):
    """Load balancer (batch screening on the host, or on the device with
    ex="Device"); its tasks are built on first use."""
    from . import api
    mg = api.MolGridFactory.create_default_molgrid(mol, "Unpruned", batch_size, "MuraKnowles",
                                                   grid)
    lb = api.LoadBalancerFactory(ex, "Default").get_instance(api.RuntimeEnvironment(), mol,
                                                              mg, basis)
    return lb
