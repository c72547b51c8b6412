"""semidiscreteot_amd -- MI355X (gfx950) regularized semi-discrete optimal
transport dual evaluation, a drop-in for the functional / gradient / Hessian
assembly behind SemiDiscreteOT's SotSolver and its BFGS driver.

The compute path is libsdot.so (HIP kernels + C ABI, include/sdot.h); this
package is the host-side mirror of the reference's SotSolver interface.
"""
from ._lib import LIB_PATH, SdotError, load  # noqa: F401
from .mesh import SourceArrays, simplex_source, structured_source  # noqa: F401
from .solver import SolverParameters, SotSolver  # noqa: F401
from .server import ServeResult, SotServer  # noqa: F401
from .epsilon_scaling import epsilon_schedule, level_schedules  # noqa: F401
from .lloyd import run_lloyd  # noqa: F401
from .transport_plan import OptimalTransportPlan  # noqa: F401
from .parameters import ParameterFile  # noqa: F401
from .multilevel import (MultilevelParameters, SoftmaxRefinement,  # noqa: F401
                         generate_target_hierarchy, run_combined_multilevel,
                         run_target_multilevel)

__all__ = ["SotSolver", "SolverParameters", "SotServer", "ServeResult", "SourceArrays", "structured_source",
           "simplex_source", "SdotError", "load", "LIB_PATH", "epsilon_schedule", "level_schedules",
           "SoftmaxRefinement", "MultilevelParameters", "generate_target_hierarchy",
           "run_target_multilevel", "run_combined_multilevel", "run_lloyd",
           "OptimalTransportPlan", "ParameterFile"]
