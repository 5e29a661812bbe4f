"""Convert GauXC's HDF5 reference fixtures into small ``.npz`` golden files.

Runs only in the build container (it needs the read-only reference tree and
conda's ``libhdf5``); the ``.npz`` files it writes are committed and are all
the test-suite reads.  HDF5 is read through ``ctypes`` because the container
has no working ``h5py``.

Fixtures converted (reference paths relative to the GauXC tree):

* ``tests/ref_data/benzene_svwn5_cc-pvdz_ufg_ssf.hdf5`` and
  ``benzene_pbe0_cc-pvdz_ufg_ssf.hdf5`` -> ``benzene_{svwn5,pbe0}.npz``:
  MOLECULE, BASIS (stored already normalised, read back with
  ``normalize=false`` at ``src/external/hdf5_read.cxx:47-48``), DENSITY, VXC,
  EXC (checked by ``tests/xc_integrator.cxx:199-212``), TRIAL_DENSITY and FXC
  (``tests/2nd_derivative_test.cxx:121-126``), EXC_GRAD with weight
  derivatives and EXC_GRAD_HELLFEY without (``tests/xc_integrator.cxx:
  117-145,278-300``).
* ``tests/ref_data/benzene_svwn5_cc-pvdz_ufg_ssf_{robust,treutler}_prune.hdf5``
  -> ``benzene_svwn5_pruned.npz``: EXC, VXC, FXC and both gradients on the
  Robust and Treutler pruned grids (``tests/xc_integrator.cxx:410-420``,
  ``tests/2nd_derivative_test.cxx:198-208``); inputs are those of the
  unpruned SVWN5 file.
* ``tests/ref_data/benzene_weights_ssf.hdf5`` -> ``benzene_weights_ssf.npz``:
  15 tasks with raw (``tasks_unm``) points/weights and SSF-modified weights
  (``tasks_mod``), checked by ``tests/weights_host.hpp:22-57``.
* ``tests/ref_data/benzene_cc-pvdz_ufg_tasks_1mpi_rank0_pv1.hdf5`` ->
  ``benzene_tasks_ufg.npz``: the 664 load-balancer tasks checked by
  ``tests/load_balancer_test.cxx:40-82``.
* ``tests/ref_data/li_svwn5_sto3g_uks.bin`` (HDF5) -> ``li_svwn5_uks.npz``:
  the UKS check of ``tests/xc_integrator.cxx:218-237,448-453`` (scalar and
  z densities and potentials, EXC) and the UKS FXC check of
  ``tests/2nd_derivative_test.cxx:128-134`` (trial densities, FXC_s, FXC_z).
* ``tests/ref_data/cytosine_svwn5_cc-pvdz_ufg_ssf_robust_uks.hdf5`` ->
  ``cytosine_svwn5_uks.npz``: the UKS check of ``tests/xc_integrator.cxx:
  218-237,455-459`` on the Robust-pruned ultra-fine grid (EXC; scalar and z
  densities and potentials, stored as packed upper triangles since all four
  are symmetric) and the UKS gradient checks of :278-300 (EXC_GRAD with
  weight derivatives, EXC_GRAD_HELLFEY without).
* ``tests/ref_data/water_cc-pVDZ_collocation.hdf5`` ->
  ``water_collocation.npz``: gau2grid collocation + gradient golden values
  (``tests/collocation_host.hpp:111-180``).
* ``tests/basis/old/cc-pvdz.g94`` (H, C, O records) -> raw, unnormalised
  shells stored inside ``water_collocation.npz`` and
  ``benzene_tasks_ufg.npz``, and ``tests/basis/old/6-31g-star.g94`` (H, C)
  inside ``benzene_weights_ssf.npz``, so the product's own ``Shell``
  normalisation is exercised (``tests/basis/parse_basis.cxx:191-231``: l <= 1 cartesian,
  l >= 2 spherical).
"""
import ctypes
import os
import sys

import numpy as np

REF = "/root/reference/tests"
OUT = os.path.dirname(os.path.abspath(__file__))

_h5 = ctypes.CDLL("/opt/conda/lib/libhdf5.so", mode=ctypes.RTLD_GLOBAL)
_h5.H5open()
hid = ctypes.c_int64
for fn in ("H5Fopen", "H5Dopen2", "H5Dget_space", "H5Tcreate", "H5Tarray_create2",
           "H5Dget_type"):
    getattr(_h5, fn).restype = hid
_h5.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid]
_h5.H5Dopen2.argtypes = [hid, ctypes.c_char_p, hid]
_h5.H5Dget_space.argtypes = [hid]
_h5.H5Dget_type.argtypes = [hid]
_h5.H5Sget_simple_extent_ndims.argtypes = [hid]
_h5.H5Sget_simple_extent_dims.argtypes = [hid, ctypes.c_void_p, ctypes.c_void_p]
_h5.H5Dread.argtypes = [hid, hid, hid, hid, hid, ctypes.c_void_p]
_h5.H5Tcreate.argtypes = [ctypes.c_int, ctypes.c_size_t]
_h5.H5Tinsert.argtypes = [hid, ctypes.c_char_p, ctypes.c_size_t, hid]
_h5.H5Tarray_create2.argtypes = [hid, ctypes.c_uint, ctypes.c_void_p]
_h5.H5Lexists.argtypes = [hid, ctypes.c_char_p, hid]
for fn in ("H5Dclose", "H5Sclose", "H5Tclose", "H5Fclose"):
    getattr(_h5, fn).argtypes = [hid]

NATIVE_DOUBLE = hid.in_dll(_h5, "H5T_NATIVE_DOUBLE_g").value
NATIVE_INT32 = hid.in_dll(_h5, "H5T_NATIVE_INT_g").value
H5T_COMPOUND = 6
H5P_DEFAULT = 0
H5S_ALL = 0


class H5File:
    def __init__(self, path):
        self.fid = _h5.H5Fopen(path.encode(), 0, H5P_DEFAULT)
        if self.fid < 0:
            raise IOError(path)

    def exists(self, name):
        return _h5.H5Lexists(self.fid, name.encode(), H5P_DEFAULT) > 0

    def shape(self, name):
        did = _h5.H5Dopen2(self.fid, name.encode(), H5P_DEFAULT)
        sid = _h5.H5Dget_space(did)
        nd = _h5.H5Sget_simple_extent_ndims(sid)
        dims = (ctypes.c_uint64 * max(nd, 1))()
        _h5.H5Sget_simple_extent_dims(sid, dims, None)
        _h5.H5Sclose(sid)
        _h5.H5Dclose(did)
        return tuple(int(d) for d in dims[:nd])

    def read(self, name, dtype=np.float64):
        shape = self.shape(name)
        mtype = NATIVE_DOUBLE if np.dtype(dtype) == np.float64 else NATIVE_INT32
        arr = np.zeros(shape, dtype=dtype)
        if arr.size == 0:
            return arr
        did = _h5.H5Dopen2(self.fid, name.encode(), H5P_DEFAULT)
        if _h5.H5Dread(did, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, arr.ctypes.data) < 0:
            raise IOError(name)
        _h5.H5Dclose(did)
        return arr

    def read_compound(self, name, np_dtype, members):
        """members: list of (h5 member name, numpy field name)."""
        shape = self.shape(name)
        arr = np.zeros(shape, dtype=np_dtype)
        tid = _h5.H5Tcreate(H5T_COMPOUND, np_dtype.itemsize)
        extra = []
        for h5name, field in members:
            fdt, off = np_dtype.fields[field][:2]
            if fdt.subdtype is not None:
                base, sub = fdt.subdtype
                dims = (ctypes.c_uint64 * len(sub))(*sub)
                btype = NATIVE_DOUBLE if base == np.float64 else NATIVE_INT32
                mt = _h5.H5Tarray_create2(btype, len(sub), dims)
                extra.append(mt)
            else:
                mt = NATIVE_DOUBLE if fdt == np.float64 else NATIVE_INT32
            _h5.H5Tinsert(tid, h5name.encode(), off, mt)
        did = _h5.H5Dopen2(self.fid, name.encode(), H5P_DEFAULT)
        if _h5.H5Dread(did, tid, H5S_ALL, H5S_ALL, H5P_DEFAULT, arr.ctypes.data) < 0:
            raise IOError(name)
        _h5.H5Dclose(did)
        _h5.H5Tclose(tid)
        for mt in extra:
            _h5.H5Tclose(mt)
        return arr

    def close(self):
        _h5.H5Fclose(self.fid)


MOL_DT = np.dtype([("Z", np.int32), ("x", np.float64), ("y", np.float64), ("z", np.float64)],
                  align=True)
BASIS_DT = np.dtype([("nprim", np.int32), ("l", np.int32), ("pure", np.int32),
                     ("alpha", np.float64, (16,)), ("coeff", np.float64, (16,)),
                     ("origin", np.float64, (3,))], align=True)


def read_mol(f):
    m = f.read_compound("/MOLECULE", MOL_DT, [("Atomic Number", "Z"), ("X Coordinate", "x"),
                                             ("Y Coordinate", "y"), ("Z Coordinate", "z")])
    return m["Z"].astype(np.int32), np.stack([m["x"], m["y"], m["z"]], axis=1)


def read_basis(f):
    b = f.read_compound("/BASIS", BASIS_DT, [("NPRIM", "nprim"), ("L", "l"), ("PURE", "pure"),
                                            ("ALPHA", "alpha"), ("COEFF", "coeff"),
                                            ("ORIGIN", "origin")])
    return dict(sh_nprim=b["nprim"].astype(np.int32), sh_l=b["l"].astype(np.int32),
                sh_pure=b["pure"].astype(np.int32), sh_alpha=b["alpha"].copy(),
                sh_coeff=b["coeff"].copy(), sh_origin=b["origin"].copy())


def read_tasks(f, group, with_points):
    ntasks = int(f.read(f"/{group}/ntasks", np.int32)[0])
    out = dict(iParent=[], npts=[], nbe=[], dist_nearest=[], shell_off=[0], shell_list=[],
               weights=[], points=[])
    for t in range(ntasks):
        g = f"/{group}/task_{t}"
        out["iParent"].append(int(f.read(g + "/iParent", np.int32)[0]))
        out["npts"].append(int(f.read(g + "/npts", np.int32)[0]))
        out["nbe"].append(int(f.read(g + "/bfn_screening_nbe", np.int32)[0]))
        out["dist_nearest"].append(float(f.read(g + "/dist_nearest")[0]))
        sl = f.read(g + "/shell_list", np.int32)
        out["shell_list"].extend(sl.tolist())
        out["shell_off"].append(len(out["shell_list"]))
        out["weights"].append(f.read(g + "/weights"))
        if with_points:
            out["points"].append(f.read(g + "/points").reshape(-1, 3))
    res = dict(iParent=np.array(out["iParent"], np.int32), npts=np.array(out["npts"], np.int32),
               nbe=np.array(out["nbe"], np.int32),
               dist_nearest=np.array(out["dist_nearest"]),
               shell_off=np.array(out["shell_off"], np.int64),
               shell_list=np.array(out["shell_list"], np.int32),
               weights=np.concatenate(out["weights"]) if out["weights"] else np.zeros(0))
    if with_points:
        res["points"] = np.concatenate(out["points"])
    return res


AM = {"S": 0, "P": 1, "D": 2, "F": 3, "G": 4}
SYMB = {"H": 1, "C": 6, "N": 7, "O": 8}


def parse_g94(path, elements):
    """Raw (unnormalised) shells per element, as tests/basis/parse_basis.cxx:191-231 reads them."""
    recs, cur = [], []
    for line in open(path):
        s = line.strip()
        if not s or s.startswith("!"):
            continue
        if s == "****":
            if cur:
                recs.append(cur)
            cur = []
        else:
            cur.append(s)
    if cur:
        recs.append(cur)
    shells = {}
    for rec in recs:
        sym = rec[0].split()[0].upper()
        if sym not in elements:
            continue
        lst, i = [], 1
        while i < len(rec):
            tok = rec[i].split()
            typ, nprim = tok[0].upper(), int(tok[1])
            alpha, c1, c2 = [], [], []
            for k in range(nprim):
                p = rec[i + 1 + k].replace("D", "e").replace("d", "e").split()
                alpha.append(float(p[0]))
                c1.append(float(p[1]))
                if typ == "SP":
                    c2.append(float(p[2]))
            i += 1 + nprim
            if typ == "SP":
                lst.append((0, alpha, c1))
                lst.append((1, alpha, c2))
            else:
                lst.append((AM[typ], alpha, c1))
        shells[SYMB[sym]] = lst
    return shells


def raw_basis_for(Z, xyz, g94):
    nprim, l, pure, alpha, coeff, origin = [], [], [], [], [], []
    for z, r in zip(Z, xyz):
        for (am, a, c) in g94[int(z)]:
            nprim.append(len(a))
            l.append(am)
            pure.append(1 if am > 1 else 0)
            aa = np.zeros(16)
            cc = np.zeros(16)
            aa[:len(a)] = a
            cc[:len(c)] = c
            alpha.append(aa)
            coeff.append(cc)
            origin.append(np.asarray(r, dtype=np.float64))
    return dict(raw_nprim=np.array(nprim, np.int32), raw_l=np.array(l, np.int32),
                raw_pure=np.array(pure, np.int32), raw_alpha=np.array(alpha),
                raw_coeff=np.array(coeff), raw_origin=np.array(origin))


def main():
    g94 = parse_g94(os.path.join(REF, "basis/old/cc-pvdz.g94"), {"H", "C", "O"})

    for tag, fname in (("svwn5", "benzene_svwn5_cc-pvdz_ufg_ssf.hdf5"),
                       ("pbe0", "benzene_pbe0_cc-pvdz_ufg_ssf.hdf5")):
        f = H5File(os.path.join(REF, "ref_data", fname))
        Z, xyz = read_mol(f)
        d = dict(Z=Z, xyz=xyz, **read_basis(f))
        # HDF5 stores the Eigen (column-major) matrices in C order; both are
        # symmetric so the transpose is immaterial, keep it explicit anyway.
        d["density"] = f.read("/DENSITY").T.copy()
        d["vxc"] = f.read("/VXC").T.copy()
        d["exc"] = f.read("/EXC")
        # FXC contraction check (tests/2nd_derivative_test.cxx:73-92,121-126)
        # and the full nuclear gradient (/EXC_GRAD carries the attribute
        # includes_weight_derivatives = 1; tests/xc_integrator.cxx:117-137).
        d["trial_density"] = f.read("/TRIAL_DENSITY").T.copy()
        d["fxc"] = f.read("/FXC").T.copy()
        d["exc_grad"] = f.read("/EXC_GRAD")
        # Hellmann-Feynman gradient (tests/xc_integrator.cxx:138-145,291-300).
        d["exc_grad_hf"] = f.read("/EXC_GRAD_HELLFEY")
        f.close()
        np.savez_compressed(os.path.join(OUT, f"benzene_{tag}.npz"), **d)

    # Pruned grids (tests/xc_integrator.cxx:410-420, 2nd_derivative_test.cxx:
    # 198-208): the same molecule, basis, density and trial density as the
    # unpruned SVWN5 file (checked here), on the Robust and Treutler pruned
    # ultra-fine grids; EXC, VXC, FXC (packed upper triangles) and both
    # gradients.
    base = dict(np.load(os.path.join(OUT, "benzene_svwn5.npz")))
    d = {}
    for scheme in ("robust", "treutler"):
        f = H5File(os.path.join(REF, "ref_data", f"benzene_svwn5_cc-pvdz_ufg_ssf_{scheme}_prune.hdf5"))
        Z, xyz = read_mol(f)
        assert np.array_equal(Z, base["Z"]) and np.array_equal(xyz, base["xyz"])
        assert np.array_equal(f.read("/DENSITY").T, base["density"])
        assert np.array_equal(f.read("/TRIAL_DENSITY").T, base["trial_density"])
        iu = np.triu_indices(base["density"].shape[0])
        d[f"{scheme}_exc"] = f.read("/EXC")
        d[f"{scheme}_vxc_packed"] = f.read("/VXC").T[iu].copy()
        d[f"{scheme}_fxc_packed"] = f.read("/FXC").T[iu].copy()
        d[f"{scheme}_exc_grad"] = f.read("/EXC_GRAD")
        d[f"{scheme}_exc_grad_hf"] = f.read("/EXC_GRAD_HELLFEY")
        f.close()
    np.savez_compressed(os.path.join(OUT, "benzene_svwn5_pruned.npz"), **d)

    f = H5File(os.path.join(REF, "ref_data", "benzene_weights_ssf.hdf5"))
    Z, xyz = read_mol(f)
    unm = read_tasks(f, "tasks_unm", True)
    mod = read_tasks(f, "tasks_mod", False)
    f.close()
    # The fixture's tasks were made with 6-31G(d) at shell tolerance 1e-6
    # (tests/weights.cxx:40-44, standards.cxx:1411-1415): its raw shells are
    # stored so the load balancer's task point sets can be checked too.
    g631 = parse_g94(os.path.join(REF, "basis/old/6-31g-star.g94"), {"H", "C"})
    np.savez_compressed(os.path.join(OUT, "benzene_weights_ssf.npz"), Z=Z, xyz=xyz,
                        iParent=unm["iParent"], npts=unm["npts"],
                        dist_nearest=unm["dist_nearest"], points=unm["points"],
                        weights_unm=unm["weights"], weights_mod=mod["weights"],
                        iParent_mod=mod["iParent"], npts_mod=mod["npts"],
                        **raw_basis_for(Z, xyz, g631))

    f = H5File(os.path.join(REF, "ref_data", "benzene_cc-pvdz_ufg_tasks_1mpi_rank0_pv1.hdf5"))
    t = read_tasks(f, "tasks", False)
    f.close()
    fb = H5File(os.path.join(REF, "ref_data", "benzene_svwn5_cc-pvdz_ufg_ssf.hdf5"))
    Z, xyz = read_mol(fb)
    fb.close()
    np.savez_compressed(os.path.join(OUT, "benzene_tasks_ufg.npz"), Z=Z, xyz=xyz,
                        iParent=t["iParent"], npts=t["npts"], nbe=t["nbe"],
                        dist_nearest=t["dist_nearest"], shell_off=t["shell_off"],
                        shell_list=t["shell_list"], **raw_basis_for(Z, xyz, g94))

    # tests/standards.cxx:18-30 (make_water), bohr
    wZ = np.array([1, 8, 1], np.int32)
    wxyz = np.array([[0., 1.579252144093028, 2.174611055780858],
                     [0., 0., 0.],
                     [0., 1.579252144093028, -2.174611055780858]])
    f = H5File(os.path.join(REF, "ref_data", "water_cc-pVDZ_collocation.hdf5"))
    d = dict(Z=wZ, xyz=wxyz, **raw_basis_for(wZ, wxyz, g94))
    k = 0
    offs_p, offs_e, offs_m = [0], [0], [0]
    pts, mask, ev, dx, dy, dz = [], [], [], [], [], []
    while f.exists(f"/entry_{k}"):
        g = f"/entry_{k}"
        p = f.read(g + "/pts").reshape(-1, 3)
        pts.append(p)
        mask.append(f.read(g + "/mask", np.int32))
        ev.append(f.read(g + "/eval"))
        dx.append(f.read(g + "/deval_x"))
        dy.append(f.read(g + "/deval_y"))
        dz.append(f.read(g + "/deval_z"))
        offs_p.append(offs_p[-1] + len(p))
        offs_m.append(offs_m[-1] + len(mask[-1]))
        offs_e.append(offs_e[-1] + len(ev[-1]))
        k += 1
    f.close()
    d.update(pts=np.concatenate(pts), mask=np.concatenate(mask),
             eval=np.concatenate(ev), deval_x=np.concatenate(dx),
             deval_y=np.concatenate(dy), deval_z=np.concatenate(dz),
             pts_off=np.array(offs_p, np.int64), mask_off=np.array(offs_m, np.int64),
             eval_off=np.array(offs_e, np.int64))
    np.savez_compressed(os.path.join(OUT, "water_collocation.npz"), **d)

    # UKS (tests/xc_integrator.cxx:218-237, 448-453): Li / SVWN5 / STO-3G on
    # the unpruned ultra-fine grid; scalar (Pa + Pb) and z (Pa - Pb) densities
    # and potentials.
    f = H5File(os.path.join(REF, "ref_data", "li_svwn5_sto3g_uks.bin"))
    Z, xyz = read_mol(f)
    d = dict(Z=Z, xyz=xyz, **read_basis(f))
    for src, dst in (("/DENSITY_SCALAR", "density_s"), ("/DENSITY_Z", "density_z"),
                     ("/VXC_SCALAR", "vxc_s"), ("/VXC_Z", "vxc_z"),
                     ("/TRIAL_DENSITY_SCALAR", "trial_density_s"),
                     ("/TRIAL_DENSITY_Z", "trial_density_z"),
                     ("/FXC_SCALAR", "fxc_s"), ("/FXC_Z", "fxc_z")):
        d[dst] = f.read(src).T.copy()
    d["exc"] = f.read("/EXC")
    f.close()
    np.savez_compressed(os.path.join(OUT, "li_svwn5_uks.npz"), **d)

    # UKS with gradients (tests/xc_integrator.cxx:455-459, 278-300): doublet
    # cytosine / SVWN5 / cc-pVDZ, Robust pruning.  /EXC_GRAD carries
    # includes_weight_derivatives = 1 (read with h5dump -A).
    f = H5File(os.path.join(REF, "ref_data", "cytosine_svwn5_cc-pvdz_ufg_ssf_robust_uks.hdf5"))
    Z, xyz = read_mol(f)
    d = dict(Z=Z, xyz=xyz, **read_basis(f))
    for src, dst in (("/DENSITY_SCALAR", "density_s"), ("/DENSITY_Z", "density_z"),
                     ("/VXC_SCALAR", "vxc_s"), ("/VXC_Z", "vxc_z")):
        m = f.read(src).T
        assert np.abs(m - m.T).max() < 1e-13
        d[dst + "_packed"] = m[np.triu_indices(m.shape[0])].copy()
    d["nbf"] = np.array([m.shape[0]])
    d["exc"] = f.read("/EXC")
    d["exc_grad"] = f.read("/EXC_GRAD")
    d["exc_grad_hf"] = f.read("/EXC_GRAD_HELLFEY")
    f.close()
    np.savez_compressed(os.path.join(OUT, "cytosine_svwn5_uks.npz"), **d)
    print("fixtures written to", OUT)


if __name__ == "__main__":
    sys.exit(main())
