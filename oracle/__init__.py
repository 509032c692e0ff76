"""CPU FP64 restatement of the TOP_GPU finite-element hot path -- TEST INFRASTRUCTURE ONLY.

This package is the correctness oracle for the HIP library in ``top_gpu_on_matlab_amd``.
It restates, function by function, the MATLAB/CUDA reference (TOP_GPU.m and the six
``*_inplace.cu`` MEX kernels) in numpy/scipy, with each function citing the reference
``file:line`` it follows.  It is deliberately literal and slow.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import it, and only as the checker or the timed CPU baseline.  The product path
(``top_gpu_on_matlab_amd``) never imports, links or calls anything in here.

Modules
-------
constants : element stiffness Ke, trilinear transfer table RI, sub-element node tables,
            super-element sub-voxel matrices I_mat, PDE-filter element kernel.
mesh      : voxel discretisation, multigrid mesh hierarchy, boundary conditions and the
            BC-modified element matrices.
solver    : MEX-kernel restatements (gather / scatter / scatter_accum3 / Galerkin
            assembly), matrix-free K*u, V-cycle, PCG, Jacobi diagonals, coarsest solve.
topopt    : compliance, sensitivities, PDE Helmholtz filter, optimality-criteria driver.

Parity status of each function (how it is pinned) is listed in tests/test_oracle_*.py.
"""
