"""ctypes bindings of the two native libraries.

* ``libgauxc_amd.so``: the GauXC-compatible C API (include/gauxc/c/gauxc.h).
* ``libgauxc_amd_device.so``: the device C ABI (include/gauxc_amd/device.h),
  used directly by kernel-level tests and by the benchmark for timings.

Loading fails loudly when the libraries have not been built: there is no
Python or CPU fallback for any product entry point.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
HOST_LIB = os.path.join(LIB_DIR, "libgauxc_amd.so")
DEVICE_LIB = os.path.join(LIB_DIR, "libgauxc_amd_device.so")

c_i32, c_i64, c_dbl, c_vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
P_DBL = ctypes.POINTER(ctypes.c_double)
P_I32 = ctypes.POINTER(ctypes.c_int32)
P_I64 = ctypes.POINTER(ctypes.c_int64)


class GauXCStatus(ctypes.Structure):
    _fields_ = [("code", ctypes.c_int), ("message", ctypes.c_void_p)]


class GauXCHandle(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("ptr", ctypes.c_void_p)]


class GauXCRuntimeHandle(ctypes.Structure):
    """GauXCRuntimeEnvironment: {header, host environment, device environment}."""
    _fields_ = [("type", ctypes.c_int), ("ptr", ctypes.c_void_p), ("device_ptr", ctypes.c_void_p)]


class GauXCAtom(ctypes.Structure):
    _fields_ = [("Z", c_i64), ("x", c_dbl), ("y", c_dbl), ("z", c_dbl)]


class GauXCShell(ctypes.Structure):
    _fields_ = [("l", c_i32), ("pure", ctypes.c_bool), ("nprim", c_i32),
                ("exponents", c_dbl * 32), ("coefficients", c_dbl * 32),
                ("origin", c_dbl * 3), ("shell_tolerance", c_dbl)]


class GauXCMolecularWeightsSettings(ctypes.Structure):
    _fields_ = [("weight_alg", ctypes.c_int), ("becke_size_adjustment", ctypes.c_bool)]


class XcqShell(ctypes.Structure):
    _fields_ = [("nprim", c_i32), ("l", c_i32), ("pure", c_i32), ("pad_", c_i32),
                ("alpha", c_dbl * 32), ("coeff", c_dbl * 32), ("origin", c_dbl * 3),
                ("cutoff_radius", c_dbl)]


class XcqTask(ctypes.Structure):
    _fields_ = [("iparent", c_i32), ("npts", c_i32), ("nshells", c_i32), ("nbe", c_i32),
                ("point_offset", c_i64), ("shell_offset", c_i64), ("dist_nearest", c_dbl)]


# The same record as a numpy structured dtype (identical 40-byte layout), so
# task tables of a million entries are packed without a Python loop.
XCQ_TASK_DTYPE = np.dtype([("iparent", "<i4"), ("npts", "<i4"), ("nshells", "<i4"),
                           ("nbe", "<i4"), ("point_offset", "<i8"), ("shell_offset", "<i8"),
                           ("dist_nearest", "<f8")], align=True)
assert XCQ_TASK_DTYPE.itemsize == ctypes.sizeof(XcqTask)


def pack_tasks(t, point_base=0, shell_base=0):
    """xcq_task records of a task dict (LoadBalancer.get_tasks layout)."""
    nt = len(t["npts"])
    out = np.zeros(nt, dtype=XCQ_TASK_DTYPE)
    for k in ("iparent", "npts", "nshells", "nbe", "dist_nearest"):
        out[k] = t[k]
    out["point_offset"] = np.asarray(t["point_offset"], dtype=np.int64) + point_base
    out["shell_offset"] = np.asarray(t["shell_offset"], dtype=np.int64) + shell_base
    return out


# Enumerations (include/gauxc/c/gauxc.h, GauXC order)
RADIAL_MURA_KNOWLES = 1
GRID = {"FineGrid": 0, "UltraFineGrid": 1, "SuperFineGrid": 2, "GM3": 3, "GM5": 4}
PRUNING_UNPRUNED = 0
PRUNING = {"Unpruned": 0, "Robust": 1, "Treutler": 2}
WEIGHT_SSF = 2
EXEC_HOST, EXEC_DEVICE = 0, 1
XCQ_FUNC = {"SVWN5": 1, "PBE": 2, "PBE0": 3}
XCQ_PREC = {"fp64": 0, "mixed": 1}

# Every symbol each header declares (checked by the CPU test-suite).
HOST_SYMBOLS = [
    "gauxc_status_delete", "gauxc_molecule_new", "gauxc_molecule_new_from_atoms",
    "gauxc_molecule_delete", "gauxc_molecule_natoms", "gauxc_basisset_new",
    "gauxc_basisset_new_from_shells", "gauxc_basisset_delete", "gauxc_basisset_nbf",
    "gauxc_basisset_get_shells", "gauxc_molgrid_new_default", "gauxc_molgrid_delete",
    "gauxc_runtime_environment_new", "gauxc_device_runtime_environment_new",
    "gauxc_device_runtime_environment_new_mem", "gauxc_runtime_environment_comm_rank",
    "gauxc_runtime_environment_comm_size", "gauxc_object_delete", "gauxc_objects_delete",
    "gauxc_runtime_environment_delete", "gauxc_load_balancer_factory_new",
    "gauxc_load_balancer_factory_delete", "gauxc_load_balancer_factory_get_instance",
    "gauxc_load_balancer_delete", "gauxc_load_balancer_task_counts",
    "gauxc_load_balancer_export_tasks", "gauxc_load_balancer_weights_modified",
    "gauxc_molecular_weights_factory_new", "gauxc_molecular_weights_factory_delete",
    "gauxc_molecular_weights_factory_get_instance", "gauxc_molecular_weights_modify_weights",
    "gauxc_molecular_weights_delete", "gauxc_functional_from_string", "gauxc_functional_delete",
    "gauxc_integrator_new", "gauxc_integrator_delete", "gauxc_integrator_integrate_den",
    "gauxc_integrator_eval_exc_rks", "gauxc_integrator_eval_exc_vxc_rks",
    "gauxc_integrator_eval_exc_uks", "gauxc_integrator_eval_exc_vxc_uks",
    "gauxc_integrator_eval_fxc_contraction_rks", "gauxc_integrator_eval_exc_grad_rks",
    "gauxc_integrator_eval_exc_grad_hf_rks", "gauxc_integrator_eval_fxc_contraction_uks",
    "gauxc_integrator_eval_exc_grad_uks", "gauxc_integrator_eval_exc_grad_hf_uks",
    "gauxc_integrator_device_context",
]
DEVICE_SYMBOLS = [
    "xcq_device_count", "xcq_last_error", "xcq_context_create", "xcq_context_create_mem",
    "xcq_context_destroy",
    "xcq_synchronize", "xcq_upload_basis", "xcq_upload_tasks", "xcq_upload_density",
    "xcq_ssf_weights", "xcq_ssf_weights_tasks", "xcq_ssf_weights_noalloc",
    "xcq_upload_density_z", "xcq_exc_vxc_uks", "xcq_download_results_uks",
    "xcq_upload_trial_density", "xcq_fxc_contraction", "xcq_download_fxc",
    "xcq_exc_grad", "xcq_exc_grad_uks", "xcq_download_exc_grad", "xcq_collision_detection",
    "xcq_set_ssf_neighbour_lists",
    "xcq_upload_trial_density_z", "xcq_fxc_contraction_uks", "xcq_download_fxc_uks",
    "xcq_download_weights", "xcq_exc_vxc",
    "xcq_integrate_den", "xcq_download_results", "xcq_task_collocation",
    "xcq_eval_functional", "xcq_stage_times", "xcq_work_counters",
    "xcq_upload_molecules", "xcq_upload_tasks_batched", "xcq_upload_density_mol",
    "xcq_download_results_mol", "xcq_set_tile_policy", "xcq_set_density_occupancy",
    "xcq_get_tile_policy", "xcq_set_density_wide", "xcq_set_vxc_drop_tolerance",
    "xcq_set_density_basis_eval", "xcq_get_density_basis_eval",
    "xcq_set_schedule",
    "xcq_enable_task_timing", "xcq_task_times", "xcq_use_graph", "xcq_autotune",
    "xcq_mlp_create", "xcq_mlp_destroy", "xcq_mlp_num_params", "xcq_mlp_get_params",
    "xcq_mlp_set_params", "xcq_mlp_forward", "xcq_mlp_gradients", "xcq_mlp_train",
]

_host = None
_dev = None


def _load(path):
    if not os.path.exists(path):
        raise RuntimeError(
            f"{os.path.basename(path)} is not built; run `python -m gauxc_amd.build` "
            "(the HIP path has no fallback)")
    return ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


def device_lib():
    global _dev
    if _dev is None:
        lib = _load(DEVICE_LIB)
        H = ctypes.c_void_p
        sig = {
            "xcq_device_count": (ctypes.c_int, [P_I32]),
            "xcq_last_error": (ctypes.c_char_p, []),
            "xcq_context_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_size_t, c_dbl,
                                                  ctypes.POINTER(H)]),
            "xcq_context_create_mem": (ctypes.c_int, [ctypes.c_int, c_vp, ctypes.c_size_t,
                                                      ctypes.POINTER(H)]),
            "xcq_context_destroy": (ctypes.c_int, [H]),
            "xcq_synchronize": (ctypes.c_int, [H]),
            "xcq_upload_basis": (ctypes.c_int, [H, c_i32, c_vp]),
            "xcq_upload_tasks": (ctypes.c_int, [H, c_i32, c_vp, c_i64, c_vp, c_vp, c_i64,
                                                c_vp]),
            "xcq_upload_density": (ctypes.c_int, [H, c_i32, c_vp, c_i64]),
            "xcq_ssf_weights": (ctypes.c_int, [ctypes.c_int, c_i32, c_vp, c_i64, c_vp, c_vp,
                                               c_vp]),
            "xcq_ssf_weights_tasks": (ctypes.c_int, [H, c_i32, c_vp]),
            "xcq_upload_density_z": (ctypes.c_int, [H, c_i32, c_vp, c_i64]),
            "xcq_exc_vxc_uks": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_int]),
            "xcq_download_results_uks": (ctypes.c_int, [H, c_vp, c_vp, c_vp, c_vp, c_i64]),
            "xcq_upload_trial_density": (ctypes.c_int, [H, c_i32, c_vp, c_i64]),
            "xcq_fxc_contraction": (ctypes.c_int, [H, ctypes.c_int]),
            "xcq_download_fxc": (ctypes.c_int, [H, c_vp, c_vp, c_i64]),
            "xcq_exc_grad": (ctypes.c_int, [H, ctypes.c_int, c_i32, c_vp, ctypes.c_int]),
            "xcq_exc_grad_uks": (ctypes.c_int, [H, ctypes.c_int, c_i32, c_vp, ctypes.c_int]),
            "xcq_set_ssf_neighbour_lists": (ctypes.c_int, [H, c_i32]),
            "xcq_upload_trial_density_z": (ctypes.c_int, [H, c_i32, c_vp, c_i64]),
            "xcq_fxc_contraction_uks": (ctypes.c_int, [H, ctypes.c_int]),
            "xcq_download_fxc_uks": (ctypes.c_int, [H, c_vp, c_vp, c_vp, c_i64]),
            "xcq_download_exc_grad": (ctypes.c_int, [H, c_vp, c_vp, c_vp]),
            "xcq_collision_detection": (ctypes.c_int, [ctypes.c_int, c_i64, c_vp, c_vp, c_i32,
                                                       c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
            "xcq_ssf_weights_noalloc": (ctypes.c_int, [ctypes.c_int, c_i32, c_vp, c_i64, c_vp,
                                                       c_vp, c_vp]),
            "xcq_download_weights": (ctypes.c_int, [H, c_vp]),
            "xcq_exc_vxc": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "xcq_integrate_den": (ctypes.c_int, [H, c_vp]),
            "xcq_download_results": (ctypes.c_int, [H, c_vp, c_vp, c_vp, c_i64]),
            "xcq_task_collocation": (ctypes.c_int, [H, c_i32, ctypes.c_int, ctypes.c_int,
                                                    c_vp]),
            "xcq_eval_functional": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, c_i64, c_vp,
                                                   c_vp, c_vp, c_vp, c_vp]),
            "xcq_stage_times": (ctypes.c_int, [H, c_vp, ctypes.c_int]),
            "xcq_work_counters": (ctypes.c_int, [H, c_vp, ctypes.c_int]),
            "xcq_upload_molecules": (ctypes.c_int, [H, c_i32, c_vp, c_vp]),
            "xcq_upload_tasks_batched": (ctypes.c_int, [H, c_i32, c_vp, c_vp, c_i64, c_vp, c_vp,
                                                        c_i64, c_vp]),
            "xcq_upload_density_mol": (ctypes.c_int, [H, c_i32, c_i32, c_vp, c_i64]),
            "xcq_download_results_mol": (ctypes.c_int, [H, c_i32, c_vp, c_vp, c_vp, c_i64]),
            "xcq_set_tile_policy": (ctypes.c_int, [H, c_i32, c_i32]),
            "xcq_set_density_occupancy": (ctypes.c_int, [H, c_i32]),
            "xcq_get_tile_policy": (ctypes.c_int, [H, c_vp]),
            "xcq_set_density_wide": (ctypes.c_int, [H, c_i32]),
            "xcq_set_density_basis_eval": (ctypes.c_int, [H, c_i32]),
            "xcq_get_density_basis_eval": (ctypes.c_int, [H, ctypes.POINTER(c_i32)]),
            "xcq_set_vxc_drop_tolerance": (ctypes.c_int, [H, ctypes.c_double]),
            "xcq_set_schedule": (ctypes.c_int, [H, c_i32, c_vp, c_vp]),
            "xcq_enable_task_timing": (ctypes.c_int, [H, ctypes.c_int]),
            "xcq_task_times": (ctypes.c_int, [H, c_vp, c_vp]),
            "xcq_use_graph": (ctypes.c_int, [H, ctypes.c_int]),
            "xcq_autotune": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_int, c_vp, c_vp, c_vp,
                                            c_vp]),
            "xcq_mlp_create": (ctypes.c_int, [ctypes.c_int, c_i32, c_i32, c_i32, c_i32,
                                              ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]),
            "xcq_mlp_destroy": (ctypes.c_int, [H]),
            "xcq_mlp_num_params": (ctypes.c_int, [H, c_vp]),
            "xcq_mlp_get_params": (ctypes.c_int, [H, c_vp]),
            "xcq_mlp_set_params": (ctypes.c_int, [H, c_vp]),
            "xcq_mlp_forward": (ctypes.c_int, [H, c_i64, c_vp, c_vp]),
            "xcq_mlp_gradients": (ctypes.c_int, [H, c_i64, c_vp, c_vp, c_vp, c_vp]),
            "xcq_mlp_train": (ctypes.c_int, [H, c_i64, c_vp, c_vp, c_i32, c_i32, ctypes.c_float,
                                             ctypes.c_uint64, c_vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype, f.argtypes = res, args
        _dev = lib
    return _dev


def host_lib():
    global _host
    if _host is None:
        device_lib()
        lib = _load(HOST_LIB)
        S = ctypes.POINTER(GauXCStatus)
        Hd = GauXCHandle
        PH = ctypes.POINTER(GauXCHandle)
        Rt = GauXCRuntimeHandle
        sig = {
            "gauxc_status_delete": (None, [S]),
            "gauxc_molecule_new_from_atoms": (Hd, [S, c_vp, ctypes.c_size_t]),
            "gauxc_molecule_delete": (None, [S, PH]),
            "gauxc_molecule_natoms": (ctypes.c_size_t, [S, Hd]),
            "gauxc_basisset_new_from_shells": (Hd, [S, c_vp, ctypes.c_size_t, ctypes.c_bool]),
            "gauxc_basisset_delete": (None, [S, PH]),
            "gauxc_basisset_nbf": (c_i32, [S, Hd]),
            "gauxc_basisset_get_shells": (None, [S, Hd, c_vp, c_vp]),
            "gauxc_molgrid_new_default": (Hd, [S, Hd, ctypes.c_int, c_i64, ctypes.c_int,
                                               ctypes.c_int]),
            "gauxc_molgrid_delete": (None, [S, PH]),
            "gauxc_runtime_environment_new": (Rt, [S]),
            "gauxc_device_runtime_environment_new": (Rt, [S, c_dbl]),
            "gauxc_device_runtime_environment_new_mem": (Rt, [S, c_vp, ctypes.c_size_t]),
            "gauxc_runtime_environment_delete": (None, [S, ctypes.POINTER(Rt)]),
            "gauxc_runtime_environment_comm_rank": (ctypes.c_int, [S, Rt]),
            "gauxc_runtime_environment_comm_size": (ctypes.c_int, [S, Rt]),
            "gauxc_object_delete": (None, [S, ctypes.POINTER(c_vp)]),
            "gauxc_objects_delete": (None, [S, ctypes.POINTER(c_vp), ctypes.c_size_t]),
            "gauxc_load_balancer_factory_new": (Hd, [S, ctypes.c_int, ctypes.c_char_p]),
            "gauxc_load_balancer_factory_delete": (None, [S, PH]),
            "gauxc_load_balancer_factory_get_instance": (Hd, [S, Hd, Rt, Hd, Hd, Hd]),
            "gauxc_load_balancer_delete": (None, [S, PH]),
            "gauxc_load_balancer_task_counts": (None, [S, Hd, c_vp]),
            "gauxc_load_balancer_export_tasks": (None, [S, Hd, c_vp, c_vp, c_vp, c_vp, c_vp,
                                                        c_vp, c_vp, c_vp]),
            "gauxc_load_balancer_weights_modified": (ctypes.c_bool, [S, Hd]),
            "gauxc_molecular_weights_factory_new": (Hd, [S, ctypes.c_int, ctypes.c_char_p,
                                                         GauXCMolecularWeightsSettings]),
            "gauxc_molecular_weights_factory_delete": (None, [S, PH]),
            "gauxc_molecular_weights_factory_get_instance": (Hd, [S, Hd]),
            "gauxc_molecular_weights_modify_weights": (None, [S, Hd, Hd]),
            "gauxc_molecular_weights_delete": (None, [S, PH]),
            "gauxc_functional_from_string": (Hd, [S, ctypes.c_char_p, ctypes.c_bool]),
            "gauxc_functional_delete": (None, [S, PH]),
            "gauxc_integrator_new": (Hd, [S, Hd, Hd, ctypes.c_int, ctypes.c_char_p,
                                          ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]),
            "gauxc_integrator_delete": (None, [S, PH]),
            "gauxc_integrator_integrate_den": (None, [S, Hd, c_i64, c_i64, c_vp, c_i64, c_vp]),
            "gauxc_integrator_eval_exc_rks": (None, [S, Hd, c_i64, c_i64, c_vp, c_i64, c_vp]),
            "gauxc_integrator_eval_exc_vxc_rks": (None, [S, Hd, c_i64, c_i64, c_vp, c_i64, c_vp,
                                                         c_vp, c_i64]),
            "gauxc_integrator_eval_exc_uks": (None, [S, Hd, c_i64, c_i64, c_vp, c_i64, c_vp,
                                                     c_i64, c_vp]),
            "gauxc_integrator_eval_exc_vxc_uks": (None, [S, Hd, c_i64, c_i64, c_vp, c_i64, c_vp,
                                                         c_i64, c_vp, c_vp, c_i64, c_vp, c_i64]),
            "gauxc_integrator_eval_fxc_contraction_rks": (None, [S, Hd, c_i64, c_i64, c_vp,
                                                                 c_i64, c_vp, c_i64, c_vp,
                                                                 c_i64]),
            "gauxc_integrator_eval_exc_grad_rks": (None, [S, Hd, c_i64, c_i64, c_vp, c_i64,
                                                          c_vp]),
            "gauxc_integrator_eval_exc_grad_hf_rks": (None, [S, Hd, c_i64, c_i64, c_vp, c_i64,
                                                             c_vp]),
            "gauxc_integrator_eval_exc_grad_uks": (None, [S, Hd, c_i64, c_i64, c_vp, c_i64,
                                                          c_vp, c_i64, c_vp]),
            "gauxc_integrator_eval_exc_grad_hf_uks": (None, [S, Hd, c_i64, c_i64, c_vp, c_i64,
                                                             c_vp, c_i64, c_vp]),
            "gauxc_integrator_eval_fxc_contraction_uks": (None, [S, Hd, c_i64, c_i64] +
                                                          [c_vp, c_i64] * 6),
            "gauxc_integrator_device_context": (c_vp, [S, Hd]),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype, f.argtypes = res, args
        _host = lib
    return _host


class GauXCError(RuntimeError):
    pass


class Status:
    """Wraps a GauXCStatus; raise_if() turns a non-zero code into GauXCError."""

    def __init__(self):
        self.s = GauXCStatus(0, None)

    @property
    def ptr(self):
        return ctypes.byref(self.s)

    def raise_if(self):
        if self.s.code != 0:
            msg = ctypes.string_at(self.s.message).decode() if self.s.message else "error"
            host_lib().gauxc_status_delete(self.ptr)
            raise GauXCError(msg)


def xcq_check(rc, what="device call"):
    if rc != 0:
        raise GauXCError(f"{what} failed ({rc}): {device_lib().xcq_last_error().decode()}")


def device_count():
    n = ctypes.c_int32(0)
    rc = device_lib().xcq_device_count(ctypes.byref(n))
    return n.value if rc == 0 else 0
