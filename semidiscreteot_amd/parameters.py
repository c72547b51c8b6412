"""A whole SotParameterManager parameter file (ParameterManager.cc:19-146,
fields and defaults ParameterManager.h:42-121; the `lloyd parameters` of
LloydParameterManager, ParameterManager.cc:412-422), read by the library's
C++ reader (sdot_prm_file_parse) into the objects the callers here take:

  solver          SolverParameters       -> SotSolver.configure / solve
  multilevel      MultilevelParameters   -> run_target_multilevel / run_combined_multilevel
  io_coding       "txt" | "bin"          -> the drivers' and formats' file mode
  transport_map_truncation_radius        -> OptimalTransportPlan.set_truncation_radius
  conditional_density                    -> SotSolver.get_potential_conditioned_density
  lloyd                                  -> run_lloyd
Mesh generation, the power diagram and the selected task are read and kept
for the caller; this library does not act on them.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

from . import _lib as L
from .multilevel import MultilevelParameters
from .solver import SolverParameters


def _s(v: bytes) -> str:
    return v.decode("utf-8", errors="replace")


@dataclass
class MeshGenerationParameters:
    """SotParameterManager::MeshParameters (ParameterManager.h:42-51)."""
    n_refinements: int = 0
    grid_generator_function: str = ""
    grid_generator_arguments: str = ""
    use_tetrahedral_mesh: bool = False
    use_custom_density: bool = False
    density_file_path: str = ""
    density_file_format: str = "vtk"
    density_field_name: str = "normalized_density"

    @classmethod
    def from_c(cls, m: L.MeshParams) -> "MeshGenerationParameters":
        return cls(m.n_refinements, _s(m.grid_generator_function), _s(m.grid_generator_arguments),
                   bool(m.use_tetrahedral_mesh), bool(m.use_custom_density),
                   _s(m.density_file_path), _s(m.density_file_format), _s(m.density_field_name))


@dataclass
class ConditionalDensityParameters:
    """ParameterManager.h:117-121."""
    indices: list = field(default_factory=list)
    potential_folder: str = ""
    truncation_radius: float = -1.0


@dataclass
class LloydParameters:
    """LloydParameterManager::LloydParameters (ParameterManager.h:202-205)."""
    max_iterations: int = 1000
    relative_tolerance: float = 1e-8


@dataclass
class ParameterFile:
    solver: SolverParameters = field(default_factory=SolverParameters)
    multilevel: MultilevelParameters = field(default_factory=MultilevelParameters)
    selected_task: str = ""
    io_coding: str = "txt"
    source_mesh: MeshGenerationParameters = field(default_factory=MeshGenerationParameters)
    target_mesh: MeshGenerationParameters = field(default_factory=MeshGenerationParameters)
    power_diagram_implementation: str = "dealii"
    transport_map_truncation_radius: float = -1.0
    conditional_density: ConditionalDensityParameters = field(
        default_factory=ConditionalDensityParameters)
    lloyd: LloydParameters = field(default_factory=LloydParameters)

    @classmethod
    def from_prm(cls, text: str | None = None, path: str | None = None) -> "ParameterFile":
        """Parse a parameter file (text or path); raises SdotError on a
        grammar error, an undeclared key of a declared subsection or a bad value."""
        lib = L.load()
        f = L.PrmFile()
        L.check(lib.sdot_prm_file_default(C.byref(f)))
        if path is not None:
            L.check(lib.sdot_prm_file_load(str(path).encode(), C.byref(f)))
        else:
            L.check(lib.sdot_prm_file_parse((text or "").encode(), C.byref(f)))
        m = f.multilevel
        ml = MultilevelParameters(
            source_enabled=bool(m.source_enabled), source_min_vertices=m.source_min_vertices,
            source_max_vertices=m.source_max_vertices,
            source_hierarchy_dir=_s(m.source_hierarchy_dir), target_enabled=bool(m.target_enabled),
            target_min_points=m.target_min_points, target_max_points=m.target_max_points,
            target_hierarchy_dir=_s(m.target_hierarchy_dir),
            use_softmax_potential_transfer=bool(m.use_softmax_potential_transfer),
            use_python_clustering=bool(m.use_python_clustering),
            python_script_name=_s(m.python_script_name), output_prefix=_s(m.output_prefix))
        return cls(
            solver=SolverParameters.from_c(f.solver), multilevel=ml,
            selected_task=_s(f.selected_task), io_coding="bin" if f.io_coding == 1 else "txt",
            source_mesh=MeshGenerationParameters.from_c(f.source),
            target_mesh=MeshGenerationParameters.from_c(f.target),
            power_diagram_implementation=_s(f.power_diagram_implementation),
            transport_map_truncation_radius=f.transport_map_truncation_radius,
            conditional_density=ConditionalDensityParameters(
                [int(f.density_indices[i]) for i in range(f.n_density_indices)],
                _s(f.density_potential_folder), f.density_truncation_radius),
            lloyd=LloydParameters(f.lloyd_max_iterations, f.lloyd_relative_tolerance))
