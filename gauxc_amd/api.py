"""Python mirror of GauXC's eval_exc_vxc API over the C API.

Usage mirrors GauXC's C++ test driver (tests/xc_integrator.cxx:163-199):

    mol   = Molecule(Z, xyz)
    basis = BasisSet.from_arrays(...)            # normalises like Shell<double>
    mg    = MolGridFactory.create_default_molgrid(mol, "Unpruned", 512,
                                                   "MuraKnowles", "UltraFineGrid")
    lb    = LoadBalancerFactory("Host", "Default").get_instance(rt, mol, mg, basis)
    MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    integ = XCIntegratorFactory("Device", "Replicated", "Default", "Default",
                                "Default").get_instance(Functional("PBE0"), lb)
    exc, vxc = integ.eval_exc_vxc(P)

Every call goes through the native libraries; nothing here computes.
"""
import ctypes

import numpy as np

from . import _native as N


def _enum_exec(ex):
    if isinstance(ex, int):
        return ex
    return {"host": N.EXEC_HOST, "device": N.EXEC_DEVICE}[ex.lower()]


class _Handle:
    _delete = None

    def __init__(self, h):
        self.h = h

    def __del__(self):
        try:
            if self.h is not None and self.h.ptr:
                st = N.Status()
                getattr(N.host_lib(), self._delete)(st.ptr, ctypes.byref(self.h))
        except Exception:
            pass


class Molecule(_Handle):
    _delete = "gauxc_molecule_delete"

    def __init__(self, Z, xyz):
        Z = np.asarray(Z, dtype=np.int64)
        xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
        atoms = (N.GauXCAtom * len(Z))()
        for i, (z, r) in enumerate(zip(Z, xyz)):
            atoms[i] = N.GauXCAtom(int(z), float(r[0]), float(r[1]), float(r[2]))
        st = N.Status()
        h = N.host_lib().gauxc_molecule_new_from_atoms(st.ptr, atoms, len(Z))
        st.raise_if()
        super().__init__(h)
        self.Z = Z
        self.xyz = xyz

    def natoms(self):
        return len(self.Z)


class BasisSet(_Handle):
    _delete = "gauxc_basisset_delete"

    def __init__(self, l, pure, alpha, coeff, origin, normalize=True, shell_tolerance=None):
        n = len(l)
        shells = (N.GauXCShell * n)()
        for i in range(n):
            a = np.asarray(alpha[i], dtype=np.float64)
            c = np.asarray(coeff[i], dtype=np.float64)
            s = shells[i]
            s.l = int(l[i])
            s.pure = bool(pure[i])
            s.nprim = len(a)
            for k in range(len(a)):
                s.exponents[k] = a[k]
                s.coefficients[k] = c[k]
            for k in range(3):
                s.origin[k] = float(origin[i][k])
            s.shell_tolerance = 0.0 if shell_tolerance is None else float(shell_tolerance)
        st = N.Status()
        h = N.host_lib().gauxc_basisset_new_from_shells(st.ptr, shells, n, bool(normalize))
        st.raise_if()
        super().__init__(h)
        self.nshells = n

    @classmethod
    def from_arrays(cls, nprim, l, pure, alpha, coeff, origin, normalize=True,
                    shell_tolerance=None):
        al = [np.asarray(alpha[i][:nprim[i]]) for i in range(len(l))]
        cl = [np.asarray(coeff[i][:nprim[i]]) for i in range(len(l))]
        return cls(l, pure, al, cl, origin, normalize, shell_tolerance)

    def nbf(self):
        st = N.Status()
        n = N.host_lib().gauxc_basisset_nbf(st.ptr, self.h)
        st.raise_if()
        return int(n)

    def shells(self):
        """Normalised shells as (struct array of GauXCShell, cutoff radii)."""
        arr = (N.GauXCShell * self.nshells)()
        rad = np.zeros(self.nshells)
        st = N.Status()
        N.host_lib().gauxc_basisset_get_shells(st.ptr, self.h, arr, rad.ctypes.data)
        st.raise_if()
        return arr, rad

    def xcq_shells(self):
        arr, rad = self.shells()
        out = (N.XcqShell * self.nshells)()
        for i, s in enumerate(arr):
            d = out[i]
            d.nprim, d.l, d.pure = s.nprim, s.l, int(s.pure)
            for k in range(s.nprim):
                d.alpha[k] = s.exponents[k]
                d.coeff[k] = s.coefficients[k]
            for k in range(3):
                d.origin[k] = s.origin[k]
            d.cutoff_radius = rad[i]
        return out


class MolGrid(_Handle):
    _delete = "gauxc_molgrid_delete"


class MolGridFactory:
    @staticmethod
    def create_default_molgrid(mol, pruning="Unpruned", batch_size=512, radial="MuraKnowles",
                               grid="UltraFineGrid"):
        if pruning not in N.PRUNING:
            raise N.GauXCError(f"Unknown pruning scheme {pruning!r}")
        rq = {"Becke": 0, "MuraKnowles": 1, "MurrayHandyLaming": 2, "TreutlerAhlrichs": 3}[radial]
        st = N.Status()
        h = N.host_lib().gauxc_molgrid_new_default(st.ptr, mol.h, N.PRUNING[pruning],
                                                   int(batch_size), rq, N.GRID[grid])
        st.raise_if()
        return MolGrid(h)


class RuntimeEnvironment(_Handle):
    """Host runtime environment (GauXC RuntimeEnvironment): device work runs
    on device 0 with the default fill fraction 0.9."""
    _delete = "gauxc_runtime_environment_delete"

    def __init__(self):
        st = N.Status()
        h = N.host_lib().gauxc_runtime_environment_new(st.ptr)
        st.raise_if()
        super().__init__(h)

    def __del__(self):
        try:
            if self.h is not None and (self.h.ptr or self.h.device_ptr):
                st = N.Status()
                N.host_lib().gauxc_runtime_environment_delete(st.ptr, ctypes.byref(self.h))
        except Exception:
            pass

    def comm_rank(self):
        st = N.Status()
        r = N.host_lib().gauxc_runtime_environment_comm_rank(st.ptr, self.h)
        st.raise_if()
        return r


class DeviceRuntimeEnvironment(RuntimeEnvironment):
    """GauXC DeviceRuntimeEnvironment: the integrator's working arena is
    ``fill_fraction`` of the free device memory, or the caller's device
    buffer ``mem`` (an integer device address, 256-byte aligned) of
    ``mem_size`` bytes."""

    def __init__(self, fill_fraction=0.9, mem=None, mem_size=0):
        st = N.Status()
        if mem is None:
            h = N.host_lib().gauxc_device_runtime_environment_new(st.ptr, float(fill_fraction))
        else:
            h = N.host_lib().gauxc_device_runtime_environment_new_mem(st.ptr, mem, mem_size)
        st.raise_if()
        _Handle.__init__(self, h)


class LoadBalancer(_Handle):
    _delete = "gauxc_load_balancer_delete"

    def task_counts(self):
        c = np.zeros(3, dtype=np.int64)
        st = N.Status()
        N.host_lib().gauxc_load_balancer_task_counts(st.ptr, self.h, c.ctypes.data)
        st.raise_if()
        return int(c[0]), int(c[1]), int(c[2])

    def get_tasks(self):
        """Tasks as flat numpy arrays (GauXC LoadBalancer::get_tasks())."""
        nt, npt, nsh = self.task_counts()
        t = dict(iparent=np.zeros(nt, np.int32), npts=np.zeros(nt, np.int32),
                 nshells=np.zeros(nt, np.int32), nbe=np.zeros(nt, np.int32),
                 dist_nearest=np.zeros(nt), points=np.zeros((npt, 3)), weights=np.zeros(npt),
                 shell_list=np.zeros(nsh, np.int32))
        st = N.Status()
        N.host_lib().gauxc_load_balancer_export_tasks(
            st.ptr, self.h, *[t[k].ctypes.data for k in
                              ("iparent", "npts", "nshells", "nbe", "dist_nearest", "points",
                               "weights", "shell_list")])
        st.raise_if()
        t["point_offset"] = np.concatenate([[0], np.cumsum(t["npts"])[:-1]]).astype(np.int64)
        t["shell_offset"] = np.concatenate([[0], np.cumsum(t["nshells"])[:-1]]).astype(np.int64)
        return t

    def weights_modified(self):
        st = N.Status()
        r = N.host_lib().gauxc_load_balancer_weights_modified(st.ptr, self.h)
        st.raise_if()
        return bool(r)


class LoadBalancerFactory(_Handle):
    _delete = "gauxc_load_balancer_factory_delete"

    def __init__(self, ex="Host", kernel="Default"):
        st = N.Status()
        h = N.host_lib().gauxc_load_balancer_factory_new(st.ptr, _enum_exec(ex), kernel.encode())
        st.raise_if()
        super().__init__(h)

    def get_instance(self, rt, mol, mg, basis):
        st = N.Status()
        h = N.host_lib().gauxc_load_balancer_factory_get_instance(st.ptr, self.h, rt.h, mol.h,
                                                                  mg.h, basis.h)
        st.raise_if()
        lb = LoadBalancer(h)
        lb._keep = (mol, mg, basis)
        return lb


class MolecularWeights(_Handle):
    _delete = "gauxc_molecular_weights_delete"

    def modify_weights(self, lb):
        st = N.Status()
        N.host_lib().gauxc_molecular_weights_modify_weights(st.ptr, self.h, lb.h)
        st.raise_if()


class MolecularWeightsFactory(_Handle):
    _delete = "gauxc_molecular_weights_factory_delete"

    def __init__(self, ex="Device", lwd="Default", weight_alg="SSF"):
        alg = {"NOTPARTITIONED": 0, "Becke": 1, "SSF": 2, "LKO": 3}[weight_alg]
        st = N.Status()
        h = N.host_lib().gauxc_molecular_weights_factory_new(
            st.ptr, _enum_exec(ex), lwd.encode(), N.GauXCMolecularWeightsSettings(alg, False))
        st.raise_if()
        super().__init__(h)

    def get_instance(self):
        st = N.Status()
        h = N.host_lib().gauxc_molecular_weights_factory_get_instance(st.ptr, self.h)
        st.raise_if()
        return MolecularWeights(h)


class Functional(_Handle):
    _delete = "gauxc_functional_delete"

    def __init__(self, name, polarized=False):
        st = N.Status()
        h = N.host_lib().gauxc_functional_from_string(st.ptr, name.encode(), bool(polarized))
        st.raise_if()
        super().__init__(h)
        self.name = name


class XCIntegrator(_Handle):
    _delete = "gauxc_integrator_delete"

    def _mat(self, P):
        P = np.asfortranarray(np.asarray(P, dtype=np.float64))
        if P.ndim != 2:
            raise N.GauXCError("P must be a matrix")
        return P

    def eval_exc_vxc(self, P, Pz=None):
        """RKS: (EXC, VXC).  UKS (Pz given; P is then P_s = P_a + P_b and
        Pz = P_a - P_b): (EXC, VXC_s, VXC_z), as GauXC's eval_exc_vxc(Ps, Pz)."""
        P = self._mat(P)
        m, n = P.shape
        exc = ctypes.c_double(0.)
        st = N.Status()
        if Pz is not None:
            Pz = self._mat(Pz)
            vs, vz = np.zeros((m, n), order="F"), np.zeros((m, n), order="F")
            N.host_lib().gauxc_integrator_eval_exc_vxc_uks(
                st.ptr, self.h, m, n, P.ctypes.data, m, Pz.ctypes.data, Pz.shape[0],
                ctypes.addressof(exc), vs.ctypes.data, m, vz.ctypes.data, m)
            st.raise_if()
            return exc.value, vs, vz
        vxc = np.zeros((m, n), order="F")
        N.host_lib().gauxc_integrator_eval_exc_vxc_rks(st.ptr, self.h, m, n, P.ctypes.data, m,
                                                       ctypes.addressof(exc), vxc.ctypes.data,
                                                       m)
        st.raise_if()
        return exc.value, vxc

    def eval_exc(self, P, Pz=None):
        P = self._mat(P)
        m, n = P.shape
        exc = ctypes.c_double(0.)
        st = N.Status()
        if Pz is not None:
            Pz = self._mat(Pz)
            N.host_lib().gauxc_integrator_eval_exc_uks(st.ptr, self.h, m, n, P.ctypes.data, m,
                                                       Pz.ctypes.data, Pz.shape[0],
                                                       ctypes.addressof(exc))
        else:
            N.host_lib().gauxc_integrator_eval_exc_rks(st.ptr, self.h, m, n, P.ctypes.data, m,
                                                       ctypes.addressof(exc))
        st.raise_if()
        return exc.value

    def eval_exc_grad(self, P, Pz=None, include_weight_derivatives=True):
        """EXC nuclear gradient, (natoms, 3), as GauXC's eval_exc_grad(P)
        (RKS) or eval_exc_grad(Ps, Pz) (UKS, when Pz is given), with the
        grid-weight derivatives unless include_weight_derivatives is False."""
        P = self._mat(P)
        m, n = P.shape
        g = np.zeros(3 * self.natoms)
        st = N.Status()
        L = N.host_lib()
        if Pz is None:
            fn = (L.gauxc_integrator_eval_exc_grad_rks if include_weight_derivatives
                  else L.gauxc_integrator_eval_exc_grad_hf_rks)
            fn(st.ptr, self.h, m, n, P.ctypes.data, m, g.ctypes.data)
        else:
            Pz = self._mat(Pz)
            if Pz.shape != P.shape:
                raise N.GauXCError("P_s and P_z differ in shape")
            fn = (L.gauxc_integrator_eval_exc_grad_uks if include_weight_derivatives
                  else L.gauxc_integrator_eval_exc_grad_hf_uks)
            fn(st.ptr, self.h, m, n, P.ctypes.data, m, Pz.ctypes.data, m, g.ctypes.data)
        st.raise_if()
        return g.reshape(-1, 3)

    def eval_fxc_contraction(self, P, tP, Pz=None, tPz=None):
        """RKS FXC contraction d VXC[P + e tP] / de, as GauXC's
        eval_fxc_contraction(P, tP); with Pz and tPz (P, tP then the total
        densities) the UKS form, returning (FXC_s, FXC_z)."""
        if Pz is not None or tPz is not None:
            mats = [self._mat(x) for x in (P, Pz, tP, tPz)]
            m, n = mats[0].shape
            fs, fz = np.zeros((m, n), order="F"), np.zeros((m, n), order="F")
            st = N.Status()
            args = []
            for x in mats:
                args += [x.ctypes.data, x.shape[0]]
            N.host_lib().gauxc_integrator_eval_fxc_contraction_uks(
                st.ptr, self.h, m, n, *args, fs.ctypes.data, m, fz.ctypes.data, m)
            st.raise_if()
            return fs, fz
        P, tP = self._mat(P), self._mat(tP)
        m, n = P.shape
        fxc = np.zeros((m, n), order="F")
        st = N.Status()
        N.host_lib().gauxc_integrator_eval_fxc_contraction_rks(
            st.ptr, self.h, m, n, P.ctypes.data, m, tP.ctypes.data, tP.shape[0],
            fxc.ctypes.data, m)
        st.raise_if()
        return fxc

    def integrate_den(self, P):
        P = self._mat(P)
        m, n = P.shape
        den = ctypes.c_double(0.)
        st = N.Status()
        N.host_lib().gauxc_integrator_integrate_den(st.ptr, self.h, m, n, P.ctypes.data, m,
                                                    ctypes.addressof(den))
        st.raise_if()
        return den.value

    def device_context(self):
        st = N.Status()
        p = N.host_lib().gauxc_integrator_device_context(st.ptr, self.h)
        st.raise_if()
        return p


class XCIntegratorFactory:
    def __init__(self, ex="Device", input_type="Replicated", integrator_kernel="Default",
                 lwd_kernel="Default", reduction_kernel="Default"):
        self.args = (_enum_exec(ex), input_type, integrator_kernel, lwd_kernel, reduction_kernel)

    def get_instance(self, func, lb):
        ex, it, ik, lk, rk = self.args
        st = N.Status()
        h = N.host_lib().gauxc_integrator_new(st.ptr, func.h, lb.h, ex, it.encode(), ik.encode(),
                                              lk.encode(), rk.encode())
        st.raise_if()
        integ = XCIntegrator(h)
        integ._keep = (func, lb)
        integ.natoms = lb._keep[0].natoms()
        return integ


class DeviceContext:
    """Direct handle on the device C ABI (include/gauxc_amd/device.h)."""

    def __init__(self, device=0, arena_bytes=0, fill_fraction=0.9):
        self.lib = N.device_lib()
        h = ctypes.c_void_p()
        N.xcq_check(self.lib.xcq_context_create(device, arena_bytes, fill_fraction,
                                                ctypes.byref(h)), "xcq_context_create")
        self.h = h
        self.nbf = None
        self.tasks = None

    def close(self):
        if self.h:
            self.lib.xcq_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_basis(self, xcq_shells):
        N.xcq_check(self.lib.xcq_upload_basis(self.h, len(xcq_shells), xcq_shells),
                    "xcq_upload_basis")

    def upload_tasks(self, t):
        nt = len(t["npts"])
        desc = N.pack_tasks(t)
        pts = np.ascontiguousarray(t["points"], dtype=np.float64)
        w = np.ascontiguousarray(t["weights"], dtype=np.float64)
        sl = np.ascontiguousarray(t["shell_list"], dtype=np.int32)
        N.xcq_check(self.lib.xcq_upload_tasks(self.h, nt, desc.ctypes.data, len(w), pts.ctypes.data,
                                              w.ctypes.data, len(sl), sl.ctypes.data),
                    "xcq_upload_tasks")
        self.tasks = t

    def upload_density(self, P):
        P = np.asfortranarray(P, dtype=np.float64)
        self.nbf = P.shape[0]
        N.xcq_check(self.lib.xcq_upload_density(self.h, P.shape[0], P.ctypes.data, P.shape[0]),
                    "xcq_upload_density")

    def ssf_weights(self, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        N.xcq_check(self.lib.xcq_ssf_weights_tasks(self.h, len(xyz), xyz.ctypes.data),
                    "xcq_ssf_weights_tasks")

    def weights(self):
        w = np.zeros(len(self.tasks["weights"]))
        N.xcq_check(self.lib.xcq_download_weights(self.h, w.ctypes.data), "xcq_download_weights")
        return w

    def exc_vxc(self, func="PBE", precision="fp64", want_vxc=True):
        N.xcq_check(self.lib.xcq_exc_vxc(self.h, N.XCQ_FUNC[func], N.XCQ_PREC[precision],
                                         int(bool(want_vxc))), "xcq_exc_vxc")

    def synchronize(self):
        N.xcq_check(self.lib.xcq_synchronize(self.h), "xcq_synchronize")

    def results(self, want_vxc=True):
        exc = ctypes.c_double(0.)
        nel = ctypes.c_double(0.)
        vxc = np.zeros((self.nbf, self.nbf), order="F") if want_vxc else None
        N.xcq_check(self.lib.xcq_download_results(
            self.h, ctypes.addressof(exc), ctypes.addressof(nel),
            vxc.ctypes.data if want_vxc else None, self.nbf), "xcq_download_results")
        return exc.value, nel.value, vxc

    def upload_density_z(self, Pz):
        """UKS spin density P_z = P_a - P_b (P_s = P_a + P_b goes through
        upload_density)."""
        Pz = np.asfortranarray(Pz, dtype=np.float64)
        N.xcq_check(self.lib.xcq_upload_density_z(self.h, Pz.shape[0], Pz.ctypes.data,
                                                  Pz.shape[0]), "xcq_upload_density_z")

    def exc_vxc_uks(self, func="PBE", want_vxc=True):
        N.xcq_check(self.lib.xcq_exc_vxc_uks(self.h, N.XCQ_FUNC[func], int(bool(want_vxc))),
                    "xcq_exc_vxc_uks")

    def results_uks(self, want_vxc=True):
        """(EXC, NEL, VXC_s, VXC_z) of the last UKS evaluation."""
        exc = ctypes.c_double(0.)
        nel = ctypes.c_double(0.)
        vs = np.zeros((self.nbf, self.nbf), order="F") if want_vxc else None
        vz = np.zeros((self.nbf, self.nbf), order="F") if want_vxc else None
        N.xcq_check(self.lib.xcq_download_results_uks(
            self.h, ctypes.addressof(exc), ctypes.addressof(nel),
            vs.ctypes.data if want_vxc else None, vz.ctypes.data if want_vxc else None,
            self.nbf), "xcq_download_results_uks")
        return exc.value, nel.value, vs, vz

    def exc_grad(self, func, xyz, weight_derivs=True, uks=False):
        """EXC gradient of the uploaded density (RKS; UKS with P_z uploaded
        and uks=True); xyz (natoms, 3) in the order of the tasks' parent
        indices."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        self._grad_natoms = len(xyz)
        fn = self.lib.xcq_exc_grad_uks if uks else self.lib.xcq_exc_grad
        N.xcq_check(fn(self.h, N.XCQ_FUNC[func], len(xyz), xyz.ctypes.data,
                       int(bool(weight_derivs))), "xcq_exc_grad")

    def set_ssf_neighbour_lists(self, enable=True):
        """Bounded (True) or all-atom loops in the gradient's weight part."""
        N.xcq_check(self.lib.xcq_set_ssf_neighbour_lists(self.h, int(bool(enable))),
                    "xcq_set_ssf_neighbour_lists")

    def results_grad(self):
        """(EXC, NEL, gradient (natoms, 3)) of the last EXC gradient."""
        exc = ctypes.c_double(0.)
        nel = ctypes.c_double(0.)
        g = np.zeros((getattr(self, "_grad_natoms", 0), 3))
        N.xcq_check(self.lib.xcq_download_exc_grad(self.h, ctypes.addressof(exc),
                                                   ctypes.addressof(nel), g.ctypes.data),
                    "xcq_download_exc_grad")
        return exc.value, nel.value, g

    def upload_trial_density(self, tP):
        """Trial density of the FXC contraction."""
        tP = np.asfortranarray(tP, dtype=np.float64)
        N.xcq_check(self.lib.xcq_upload_trial_density(self.h, tP.shape[0], tP.ctypes.data,
                                                      tP.shape[0]), "xcq_upload_trial_density")

    def fxc_contraction(self, func="PBE"):
        N.xcq_check(self.lib.xcq_fxc_contraction(self.h, N.XCQ_FUNC[func]),
                    "xcq_fxc_contraction")

    def upload_trial_density_z(self, tPz):
        """Trial spin density of the UKS FXC contraction."""
        tPz = np.asfortranarray(tPz, dtype=np.float64)
        N.xcq_check(self.lib.xcq_upload_trial_density_z(self.h, tPz.shape[0], tPz.ctypes.data,
                                                        tPz.shape[0]),
                    "xcq_upload_trial_density_z")

    def fxc_contraction_uks(self, func="PBE"):
        N.xcq_check(self.lib.xcq_fxc_contraction_uks(self.h, N.XCQ_FUNC[func]),
                    "xcq_fxc_contraction_uks")

    def results_fxc_uks(self):
        """(NEL, FXC_s, FXC_z) of the last UKS FXC contraction."""
        nel = ctypes.c_double(0.)
        fs = np.zeros((self.nbf, self.nbf), order="F")
        fz = np.zeros((self.nbf, self.nbf), order="F")
        N.xcq_check(self.lib.xcq_download_fxc_uks(self.h, ctypes.addressof(nel),
                                                  fs.ctypes.data, fz.ctypes.data, self.nbf),
                    "xcq_download_fxc_uks")
        return nel.value, fs, fz

    def results_fxc(self):
        """(NEL, FXC) of the last FXC contraction."""
        nel = ctypes.c_double(0.)
        fxc = np.zeros((self.nbf, self.nbf), order="F")
        N.xcq_check(self.lib.xcq_download_fxc(self.h, ctypes.addressof(nel), fxc.ctypes.data,
                                              self.nbf), "xcq_download_fxc")
        return nel.value, fxc

    def integrate_den(self):
        nel = ctypes.c_double(0.)
        N.xcq_check(self.lib.xcq_integrate_den(self.h, ctypes.addressof(nel)),
                    "xcq_integrate_den")
        return nel.value

    def collocation(self, task, deriv=True, precision="fp64"):
        nbe = int(self.tasks["nbe"][task])
        npts = int(self.tasks["npts"][task])
        ncomp = 4 if deriv else 1
        out = np.zeros((ncomp, nbe, npts))
        N.xcq_check(self.lib.xcq_task_collocation(self.h, task, int(deriv), N.XCQ_PREC[precision],
                                                  out.ctypes.data), "xcq_task_collocation")
        return out

    def stage_times(self):
        ms = np.zeros(7)
        self.lib.xcq_stage_times(self.h, ms.ctypes.data, 7)
        return ms

    def work_counters(self):
        c = np.zeros(4, dtype=np.int64)
        self.lib.xcq_work_counters(self.h, c.ctypes.data, 4)
        return c

    # ---- scheduling, timing and graph controls ----
    def set_tile_policy(self, dens32_max_nbe, syr2k32_max_nbe):
        N.xcq_check(self.lib.xcq_set_tile_policy(self.h, int(dens32_max_nbe),
                                                 int(syr2k32_max_nbe)), "xcq_set_tile_policy")

    def tile_policy(self):
        """(dens32_max_nbe, syr2k32_max_nbe, hiocc_max_nbe, nqueues, wide_min_nbe)."""
        out = np.zeros(5, dtype=np.int32)
        N.xcq_check(self.lib.xcq_get_tile_policy(self.h, out.ctypes.data),
                    "xcq_get_tile_policy")
        return tuple(int(x) for x in out)

    def set_vxc_drop_tolerance(self, tol):
        N.xcq_check(self.lib.xcq_set_vxc_drop_tolerance(self.h, float(tol)),
                    "xcq_set_vxc_drop_tolerance")

    def set_density_basis_eval(self, mode):
        """3 (default) = the density stage evaluates basis values/gradients
        for its fold and a separate launch for Z; 1 = the mixed density stage
        for both (FP64 runs 3); 0 (or False) = both read the planes the
        collocation stage stored."""
        N.xcq_check(self.lib.xcq_set_density_basis_eval(self.h, int(mode)),
                    "xcq_set_density_basis_eval")

    def density_basis_eval(self):
        """The basis evaluation mode in effect (see set_density_basis_eval)."""
        m = ctypes.c_int32(0)
        N.xcq_check(self.lib.xcq_get_density_basis_eval(self.h, ctypes.byref(m)),
                    "xcq_get_density_basis_eval")
        return int(m.value)

    def set_density_wide(self, wide_min_nbe):
        N.xcq_check(self.lib.xcq_set_density_wide(self.h, int(wide_min_nbe)),
                    "xcq_set_density_wide")

    def set_density_occupancy(self, hiocc_max_nbe):
        N.xcq_check(self.lib.xcq_set_density_occupancy(self.h, int(hiocc_max_nbe)),
                    "xcq_set_density_occupancy")

    def set_schedule(self, nqueues=0, dens_cost=None, syr2k_cost=None):
        if dens_cost is None:
            N.xcq_check(self.lib.xcq_set_schedule(self.h, int(nqueues), None, None),
                        "xcq_set_schedule")
            return
        dc = np.ascontiguousarray(dens_cost, dtype=np.float64)
        sc = np.ascontiguousarray(syr2k_cost, dtype=np.float64)
        N.xcq_check(self.lib.xcq_set_schedule(self.h, int(nqueues), dc.ctypes.data,
                                              sc.ctypes.data), "xcq_set_schedule")

    def enable_task_timing(self, on=True):
        N.xcq_check(self.lib.xcq_enable_task_timing(self.h, int(bool(on))),
                    "xcq_enable_task_timing")

    def task_times(self):
        nt = len(self.tasks["npts"])
        d = np.zeros(nt)
        s = np.zeros(nt)
        N.xcq_check(self.lib.xcq_task_times(self.h, d.ctypes.data, s.ctypes.data),
                    "xcq_task_times")
        return d, s

    def use_graph(self, on=True):
        N.xcq_check(self.lib.xcq_use_graph(self.h, int(bool(on))), "xcq_use_graph")

    def autotune(self, func="PBE", precision="mixed"):
        """Device autotuner (xcq_autotune): returns the chosen
        (dens32_max_nbe, syr2k32_max_nbe, nqueues, best total ms)."""
        d, s, q = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        ms = ctypes.c_double()
        N.xcq_check(self.lib.xcq_autotune(self.h, N.XCQ_FUNC[func], N.XCQ_PREC[precision],
                                          ctypes.addressof(d), ctypes.addressof(s),
                                          ctypes.addressof(q), ctypes.addressof(ms)),
                    "xcq_autotune")
        return d.value, s.value, q.value, ms.value


class BatchedDeviceContext(DeviceContext):
    """Several independent molecules evaluated in the same launches.

    ``upload_molecules`` takes one list of xcq_shell arrays per molecule;
    ``upload_tasks`` takes one task dict (LoadBalancer.get_tasks layout) per
    molecule and concatenates them, tagging each task with its molecule."""

    def upload_molecules(self, shell_arrays):
        counts = np.array([len(a) for a in shell_arrays], dtype=np.int32)
        allsh = (N.XcqShell * int(counts.sum()))()
        k = 0
        for a in shell_arrays:
            for sh in a:
                allsh[k] = sh
                k += 1
        N.xcq_check(self.lib.xcq_upload_molecules(self.h, len(counts), counts.ctypes.data, allsh),
                    "xcq_upload_molecules")
        self.mol_nshells = counts
        self.mol_nbf = [None] * len(counts)

    def upload_tasks(self, task_dicts):
        sh_base = np.concatenate([[0], np.cumsum(self.mol_nshells)[:-1]]).astype(np.int64)
        descs, mols, pts, ws, sls = [], [], [], [], []
        pt_base = 0
        sl_base = 0
        for m, t in enumerate(task_dicts):
            nt = len(t["npts"])
            descs.append(N.pack_tasks(t, pt_base, sl_base))
            mols.append(np.full(nt, m, dtype=np.int32))
            pts.append(np.asarray(t["points"], dtype=np.float64).reshape(-1, 3))
            ws.append(np.asarray(t["weights"], dtype=np.float64))
            sls.append(np.asarray(t["shell_list"], dtype=np.int64) + sh_base[m])
            pt_base += len(ws[-1])
            sl_base += len(sls[-1])
        desc = np.concatenate(descs)
        nt = len(desc)
        mol = np.ascontiguousarray(np.concatenate(mols))
        P = np.ascontiguousarray(np.concatenate(pts))
        W = np.ascontiguousarray(np.concatenate(ws))
        SL = np.ascontiguousarray(np.concatenate(sls).astype(np.int32))
        N.xcq_check(self.lib.xcq_upload_tasks_batched(self.h, nt, desc.ctypes.data,
                                                      mol.ctypes.data, len(W),
                                                      P.ctypes.data, W.ctypes.data, len(SL),
                                                      SL.ctypes.data), "xcq_upload_tasks_batched")
        self.tasks = {"npts": desc["npts"].copy(), "nbe": desc["nbe"].copy(), "molecule": mol}

    def upload_density(self, mol, P):
        P = np.asfortranarray(P, dtype=np.float64)
        self.mol_nbf[mol] = P.shape[0]
        N.xcq_check(self.lib.xcq_upload_density_mol(self.h, int(mol), P.shape[0], P.ctypes.data,
                                                    P.shape[0]), "xcq_upload_density_mol")

    def results(self, mol, want_vxc=True):
        n = self.mol_nbf[mol]
        exc = ctypes.c_double(0.)
        nel = ctypes.c_double(0.)
        vxc = np.zeros((n, n), order="F") if want_vxc else None
        N.xcq_check(self.lib.xcq_download_results_mol(
            self.h, int(mol), ctypes.addressof(exc), ctypes.addressof(nel),
            vxc.ctypes.data if want_vxc else None, n), "xcq_download_results_mol")
        return exc.value, nel.value, vxc


def eval_functional_device(func, rho, sigma=None, device=0):
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    n = len(rho)
    eps = np.zeros(n)
    vr = np.zeros(n)
    gga = func != "SVWN5"
    vs = np.zeros(n) if gga else None
    sig = np.ascontiguousarray(sigma, dtype=np.float64) if gga else None
    N.xcq_check(N.device_lib().xcq_eval_functional(
        device, N.XCQ_FUNC[func], n, rho.ctypes.data, sig.ctypes.data if gga else None,
        eps.ctypes.data, vr.ctypes.data, vs.ctypes.data if gga else None), "xcq_eval_functional")
    return eps, vr, vs


def ssf_weights_device(xyz, points, parent, weights, device=0):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    pts = np.ascontiguousarray(points, dtype=np.float64)
    par = np.ascontiguousarray(parent, dtype=np.int32)
    w = np.array(weights, dtype=np.float64, copy=True)
    N.xcq_check(N.device_lib().xcq_ssf_weights(device, len(xyz), xyz.ctypes.data, len(w),
                                               pts.ctypes.data, par.ctypes.data, w.ctypes.data),
                "xcq_ssf_weights")
    return w
