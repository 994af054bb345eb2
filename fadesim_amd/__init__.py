"""fadesim_amd -- MI355X-native IR-drop solver for memristive crossbars.

Drop-in surfaces of the FADESim node-based Gauss-Seidel/SOR path:

* :func:`fadesim_amd.solver.solve_batch` -- batched solve over the C ABI
  ``fadesim_solve_batch`` (include/fadesim.h);
* :func:`fadesim_amd.hook.ir_drop_gs` -- the PimTorch hook signature of
  ``phyArrParams.cpp_ir_drop_accelerate.ir_drop_gs``;
* ``fadesim_amd/lib/sim`` -- the ``sim <config>`` command-language CLI.
"""
from __future__ import annotations

__version__ = "0.1.0"

__all__ = ["solver", "hook", "build"]
