"""cup2d_amd -- MI355X-native pressure-Poisson solver for CubismUP-2D.

The solver (BiCGSTAB on the block-AMR Laplacian with the 64x64 block-Jacobi
preconditioner, reference source/Poisson/) lives in the HIP library
``libcup2d_poisson.so`` behind the C ABI of ``include/cup2d_poisson.h``;
``cup2d_amd.poisson`` mirrors the reference's ``PoissonSolver`` /
``makePoissonSolver`` / ``solve(input, output)`` interface on top of it.
"""

from .mesh import Mesh  # noqa: F401
