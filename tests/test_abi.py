"""CPU checks of the product library: it loads, exports every symbol that
include/sdot.h declares, and its host-only parts (parameter defaults, .prm
reader, FEValues stand-in) agree with the reference's definitions and the
oracle.  No GPU compute is called here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import semidiscreteot_amd as sd
from semidiscreteot_amd import _lib as L
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "sdot.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(sdot_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = L.load()
    syms = declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(L.SIGNATURES), set(syms) ^ set(L.SIGNATURES)


def test_version_and_error_channel():
    lib = L.load()
    assert b"gfx950" in lib.sdot_version()
    rc = lib.sdot_set_params(None, None)
    assert rc == -1 and lib.sdot_last_error()


def test_parameter_defaults_match_reference():
    p = sd.SolverParameters.from_prm("")
    # ParameterManager.h:56-72
    assert p.max_iterations == 1000 and p.tolerance == 1.0 and p.epsilon == 1e-3
    assert p.tau == 1e-8 and p.solver_control_type == "l1norm"
    assert p.distance_threshold_type == "pointwise" and p.quadrature_order == 3
    assert p.use_log_sum_exp_trick is False and p.n_threads == 0
    assert p.epsilon_scaling_factor == 2.0 and p.epsilon_scaling_steps == 5


PRM = """
subsection SotParameterManager
  set selected_task =  sot    # Options: mesh_generation, sot
  subsection rsot_solver
    set max_iterations = 10000
    set epsilon = 1e-2
    set tau = 1e-8  # Truncation error tolerance for integral radius bound
    set tolerance = 1e-2
    set distance_threshold_type = integral # Options: pointwise/integral
    set solver_control_type = componentwise
    set verbose_output = false
    set solver_type = BFGS
    set quadrature_order = 2
    set number_of_threads = 0
    set use_log_sum_exp_trick = true
    set use_epsilon_scaling = false
    set epsilon_scaling_factor = 10
    set epsilon_scaling_steps = 3
    set precision = fp32_lse
    set line_search = noise_aware
    set preconditioner = diagonal
  end
  subsection power_diagram_parameters
    set implementation = geogram
  end
end
"""


def test_prm_reader():
    p = sd.SolverParameters.from_prm(PRM)
    assert p.max_iterations == 10000 and p.epsilon == 1e-2 and p.tolerance == 1e-2
    assert p.distance_threshold_type == "integral" and p.solver_control_type == "componentwise"
    assert p.verbose_output is False and p.quadrature_order == 2
    assert p.use_log_sum_exp_trick is True and p.epsilon_scaling_factor == 10.0
    assert p.epsilon_scaling_steps == 3 and p.precision == "fp32_lse"
    assert p.line_search == "noise_aware" and p.preconditioner == "diagonal"
    assert sd.SolverParameters.from_prm("").preconditioner == "none"
    # the value bound: -1 (the precision's default) unless set
    assert p.prune_bits == -1.0 and sd.SolverParameters().prune_bits == -1.0
    q = sd.SolverParameters.from_prm("subsection rsot_solver\n set prune_bits = 64\nend\n")
    assert q.prune_bits == 64.0
    assert sd.SolverParameters.from_c(q.to_c()).prune_bits == 64.0


def test_prm_reader_errors():
    with pytest.raises(sd.SdotError):
        sd.SolverParameters.from_prm("subsection rsot_solver\n set bogus_key = 1\nend\n")
    with pytest.raises(sd.SdotError):
        sd.SolverParameters.from_prm("subsection rsot_solver\n set epsilon = abc\nend\n")
    with pytest.raises(sd.SdotError):
        sd.SolverParameters.from_prm("subsection rsot_solver\n set epsilon = 1\n")
    with pytest.raises(sd.SdotError):
        sd.SolverParameters.from_prm("subsection rsot_solver\n set preconditioner = jacobi\nend\n")
    for bad in ("many", "", "inf", "nan", "1e400", "12 bits"):
        with pytest.raises(sd.SdotError):
            sd.SolverParameters.from_prm(f"subsection rsot_solver\n set prune_bits = {bad}\nend\n")
    # unknown threshold names fall back to pointwise as in SotSolver.cc:406
    p = sd.SolverParameters.from_prm("subsection rsot_solver\n set distance_threshold_type = foo\nend\n")
    assert p.distance_threshold_type == "pointwise"


def test_reference_prm_file_parses():
    path = os.path.join(ROOT, "tests", "golden", "parameters_rsot.prm")
    p = sd.SolverParameters.from_prm(path=path)
    assert p.epsilon == 1e-2 and p.max_iterations == 10000


@pytest.mark.parametrize("dim,n,q", [(2, [7, 5], 3), (3, [4, 3, 5], 2), (3, [3, 3, 3], 3)])
def test_structured_mesh_matches_oracle(dim, n, q):
    rng = np.random.default_rng(dim * 10 + q)
    nv = int(np.prod([k + 1 for k in n]))
    rho = 0.5 + rng.random(nv)
    lo, hi = [-0.5] * dim, [1.5] * dim
    a = sd.structured_source(dim, n, lo, hi, q, rho)
    o = O.structured_source(dim, n, lo, hi, q, rho)
    assert np.array_equal(a.centers, o.center)
    assert np.max(np.abs(a.x - o.x)) <= 4e-16 * 2
    assert np.max(np.abs(a.w - o.w) / o.w) < 1e-14
    assert abs(a.w.sum() - 1.0) < 1e-13


def test_simplex_mesh_matches_oracle():
    from tests.cases import tet_case
    _, _, (p, cells, rho) = tet_case(200, 5, seed=1)
    a = sd.simplex_source(3, p.ravel(), cells.ravel(), 2, rho)
    o = O.simplex_source(3, p.ravel(), cells.ravel(), 2, rho)
    assert np.array_equal(a.centers, o.center)
    assert np.max(np.abs(a.x - o.x)) < 1e-15
    assert np.max(np.abs(a.w - o.w) / np.abs(o.w)) < 1e-13


def test_mesh_argument_errors():
    with pytest.raises(sd.SdotError):
        sd.structured_source(4, [2, 2, 2], [0] * 3, [1] * 3, 2)
    with pytest.raises(sd.SdotError):
        sd.simplex_source(3, np.zeros(12), np.array([0, 1, 2, 3]), 3)


def test_cpp_facade_compiles_and_links(tmp_path):
    """include/sdot_sot_solver.hpp (the SotSolver-shaped C++ facade) builds
    against the C ABI and reads the reference defaults without a GPU."""
    import subprocess
    src = tmp_path / "facade.cpp"
    src.write_text('#include "sdot_sot_solver.hpp"\n'
                   "std::vector<double> refine(const sdot::SotSolverGPU<3> &s,\n"
                   "    const std::vector<std::vector<std::vector<size_t>>> &ch) {\n"
                   "  sdot::SoftmaxRefinementGPU<3> r(s, 0.1, true);\n"
                   "  return r.compute_refinement({}, {}, {}, {}, {}, 0.1, 0, ch); }\n"
                   "void post(sdot::SotSolverGPU<2> &s, std::vector<double> &out) {\n"
                   "  s.evaluate_weighted_barycenters({}, out);\n"
                   "  s.get_potential_conditioned_density({}, {}, {}, {}, out);\n"
                   "  sdot::OptimalTransportPlanGPU<2> plan(\"barycentric\");\n"
                   "  plan.set_source_measure({}, {}); plan.set_target_measure({}, {});\n"
                   "  plan.set_potential({}, 0.1); plan.set_truncation_radius(0.2);\n"
                   "  plan.compute_map(); out = plan.mapped_points; }\n"
                   "int main() { auto p = sdot::SotSolverGPU<3>::default_parameters();\n"
                   "  return (p.max_iterations == 1000 && p.tau == 1e-8) ? 0 : 1; }\n")
    exe = tmp_path / "facade"
    libdir = os.path.dirname(L.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src),
                    "-L", libdir, "-lsdot", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    assert subprocess.run([str(exe)]).returncode == 0


def test_lockstep_line_search_matches_loop_form(tmp_path):
    """The resumable line search used by the batched server takes the same
    decisions as the loop form used by sdot_solve (bitwise iterates on
    Rosenbrock, tests/bfgs_sm_check.cpp)."""
    import subprocess
    exe = tmp_path / "sm"
    subprocess.run(["g++", "-std=c++17", "-O2", os.path.join(ROOT, "tests", "bfgs_sm_check.cpp"),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout


def test_served_solve_driver_matches_loop_form(tmp_path):
    """The step-by-step driver of a served solve, fed by a host emulation of
    the device actions, reproduces lbfgs_solve bitwise on Rosenbrock, also
    when evaluations are repeated (tests/serve_driver_check.cpp)."""
    import subprocess
    exe = tmp_path / "drv"
    subprocess.run(["g++", "-std=c++17", "-O2", os.path.join(ROOT, "tests", "serve_driver_check.cpp"),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout


def test_server_config_defaults():
    lib = L.load()
    cfg = L.ServerConfig()
    assert lib.sdot_server_config_default(C.byref(cfg)) == 0
    assert cfg.precision == 1 and cfg.max_slots == 256 and cfg.use_graph == 1
    assert cfg.hbm_budget_bytes == 16 << 30 and cfg.polish_factor == 2.0
    out = C.c_void_p()
    assert lib.sdot_server_create(None, C.byref(cfg), C.byref(out)) != 0
    assert not out.value


# ---- whole parameter files (ParameterManager.cc:19-146, 412-422) ----

def test_parameter_file_every_subsection():
    from semidiscreteot_amd.parameters import ParameterFile
    pf = ParameterFile.from_prm(path=os.path.join(ROOT, "tests", "golden", "parameters_full.prm"))
    assert pf.selected_task == "multilevel_sot" and pf.io_coding == "bin"
    s = pf.solver
    assert (s.max_iterations, s.epsilon, s.tau, s.tolerance) == (2500, 5e-3, 1e-10, 1e-6)
    assert s.distance_threshold_type == "geometric" and s.solver_control_type == "componentwise"
    assert s.quadrature_order == 2 and s.n_threads == 4 and s.use_log_sum_exp_trick
    assert s.use_epsilon_scaling and s.epsilon_scaling_factor == 4.0 and s.epsilon_scaling_steps == 3
    src, tgt = pf.source_mesh, pf.target_mesh
    assert (src.n_refinements, src.grid_generator_function) == (5, "hyper_cube")
    assert src.grid_generator_arguments == "-1 : 1 : false" and src.use_tetrahedral_mesh
    assert src.density_file_path == "output/density_field/source_density.vtk"
    assert (src.density_field_name, src.density_file_format) == ("rho", "h5")
    # continuation line joined; unset keys keep their defaults
    assert tgt.grid_generator_arguments.replace(" ", "") == "0,0,0:1:true"
    assert (tgt.density_file_format, tgt.density_field_name) == ("vtk", "normalized_density")
    ml = pf.multilevel
    assert ml.source_enabled and (ml.source_min_vertices, ml.source_max_vertices) == (500, 20000)
    assert ml.source_hierarchy_dir == "out/src_levels" and ml.target_enabled
    assert (ml.target_min_points, ml.target_max_points) == (250, 64000)
    assert ml.target_hierarchy_dir == "out/tgt_levels" and not ml.use_softmax_potential_transfer
    assert ml.use_python_clustering and ml.python_script_name == "multilevel_clustering_faiss_cpu.py"
    assert ml.output_prefix == "out/multilevel"
    assert pf.power_diagram_implementation == "geogram"
    assert pf.transport_map_truncation_radius == 0.25
    cd = pf.conditional_density
    assert cd.indices == [0, 3, 17] and cd.potential_folder == "out/potentials"
    assert cd.truncation_radius == 0.5
    assert (pf.lloyd.max_iterations, pf.lloyd.relative_tolerance) == (40, 1e-5)


def test_parameter_file_defaults_and_errors():
    from semidiscreteot_amd.parameters import ParameterFile
    d = ParameterFile.from_prm(text="")
    # ParameterManager.h:42-121, 202-205
    assert d.io_coding == "txt" and d.selected_task == ""
    assert d.multilevel.target_hierarchy_dir == "output/data_multilevel/target_multilevel"
    assert d.multilevel.source_hierarchy_dir == "output/data_multilevel/source_multilevel"
    assert d.multilevel.use_softmax_potential_transfer and not d.multilevel.target_enabled
    assert (d.multilevel.source_min_vertices, d.multilevel.source_max_vertices) == (1000, 10000)
    assert (d.multilevel.target_min_points, d.multilevel.target_max_points) == (100, 1000)
    assert d.multilevel.python_script_name == "multilevel_clustering_scipy.py"
    assert d.multilevel.output_prefix == "output/multilevel/sot"
    assert d.transport_map_truncation_radius == -1.0 and d.conditional_density.truncation_radius == -1.0
    assert d.conditional_density.indices == [] and d.power_diagram_implementation == "dealii"
    assert (d.lloyd.max_iterations, d.lloyd.relative_tolerance) == (1000, 1e-8)
    assert d.solver.epsilon == 1e-3 and d.solver.max_iterations == 1000
    # the manager's keys at the top level (no SotParameterManager subsection)
    t = ParameterFile.from_prm(text="set io_coding = bin\nsubsection transport_map_parameters\n"
                                    "  set truncation_radius = 2\nend\n")
    assert t.io_coding == "bin" and t.transport_map_truncation_radius == 2.0
    bad = [
        "subsection SotParameterManager\n set io_codng = txt\nend\n",               # undeclared key
        "subsection SotParameterManager\n set io_coding = vtk\nend\n",              # bad value
        "subsection SotParameterManager\n subsection multilevel_parameters\n"
        "  set target_min_points = many\n end\nend\n",                               # not an integer
        "subsection SotParameterManager\n subsection no_such_section\n set a = 1\n end\nend\n",
        "subsection SotParameterManager\n subsection conditional_density_parameters\n"
        "  set indices = 1, -2\n end\nend\n",                                        # negative index
        "subsection SotParameterManager\n set selected_task = sot\n",               # missing end
        "subsection SotParameterManager\n subsection transport_map_parameters\n"
        "  set radius = 1\n end\nend\n",                                             # undeclared key
    ]
    for text in bad:
        with pytest.raises(L.SdotError):
            ParameterFile.from_prm(text=text)


def test_parameter_file_reads_the_reference_run_files():
    """Every parameter file the reference ships parses (only in a tree where
    the reference checkout is present)."""
    from semidiscreteot_amd.parameters import ParameterFile
    import glob
    files = glob.glob("/root/reference/**/*.prm", recursive=True)
    if not files:
        pytest.skip("reference checkout not present")
    for f in files:
        pf = ParameterFile.from_prm(path=f)
        assert pf.io_coding in ("txt", "bin")
