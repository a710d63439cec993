"""MI355X-native ERT forward + Jacobian path with pyGIMLi's API.

    from pygimli_amd import ERTModelling, DataContainerERT, Mesh
    fop = ERTModelling(); fop.setMesh(mesh); fop.setData(data)
    rhoa = fop.response(rho); fop.createJacobian(rho); J = fop.jacobian()

ERTModelling(sr=False) is the total-field operator (topography, bypass
electrodes) and, with setComplex(True), the complex-resistivity one;
load / save_data / save_mesh read and write the unified data format and .bms
meshes; inversion.solveCGLSCDWWhtrans runs the Gauss-Newton step's CGLS.
"""
from .data import DataContainerERT  # noqa: F401
from .mesh import Mesh  # noqa: F401


def __getattr__(name):
    # torch-dependent parts load lazily
    if name == "ERTModelling":
        from .ertModelling import ERTModelling
        return ERTModelling
    if name in ("createStiffnessMatrix", "createMassMatrix"):
        from . import solver
        return getattr(solver, name)
    if name in ("load", "load_data", "save_data", "load_mesh", "save_mesh"):
        from . import fileio
        return getattr(fileio, name)
    raise AttributeError(name)
