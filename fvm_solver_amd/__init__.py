"""MI355X-native (gfx950) FVM step loop: ELL SpMV/SpMM, the compressible Navier-Stokes
residual and RK/Adams stage updates as hand-written HIP kernels behind the call
signatures of Merlotec/fvm_solver's ``time_fvm`` package.

Module map (reference module -> this package):
  time_fvm.utils.ell_kernel          -> fvm_solver_amd.utils.ell_kernel
  time_fvm.utils.sparse              -> fvm_solver_amd.utils.sparse
  time_fvm.mesh_utils.fvm_mesh       -> fvm_solver_amd.mesh_utils.fvm_mesh
  time_fvm.mesh_utils.mesh_store     -> fvm_solver_amd.mesh_utils.mesh_store
  time_fvm.fvm_stepping.facet_process-> fvm_solver_amd.fvm_stepping.facet_process
  time_fvm.fvm_equation              -> fvm_solver_amd.fvm_equation
  time_fvm.time_solvers.*            -> fvm_solver_amd.time_solvers.*
  time_fvm.config_fvm                -> fvm_solver_amd.config_fvm
"""
__version__ = "0.1.0"
