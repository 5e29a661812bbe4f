"""gauxc_amd -- MI355X-native XC quadrature behind GauXC's eval_exc_vxc API.

Native libraries (built in-tree by ``python -m gauxc_amd.build``):
  lib/libgauxc_amd_device.so  HIP kernels for gfx950 + device C ABI
  lib/libgauxc_amd.so         GauXC-compatible C++ host library + C API
"""
from ._native import GauXCError, device_count  # noqa: F401
from .api import (  # noqa: F401
    BasisSet, BatchedDeviceContext, DeviceContext, DeviceRuntimeEnvironment, Functional, LoadBalancer,
    LoadBalancerFactory, MolecularWeights, MolecularWeightsFactory, MolGrid, MolGridFactory,
    Molecule, RuntimeEnvironment, XCIntegrator, XCIntegratorFactory, eval_functional_device,
    ssf_weights_device,
)
from . import workloads  # noqa: F401,E402
