"""CPU oracle for the displacement-formulation assembly path -- TEST INFRASTRUCTURE.

Nothing in the product package may import, call or link anything under this
directory.  Only the tests, ``__graft_entry__.smoke()`` (as the checker) and the
``cpu_baseline`` leg of ``bench.py`` use it.

Contents
* ``df_numpy``  -- FP64 numpy restatement of the reference's low-level DF assembly
  (kinematics, Neo-Hookean / Mooney-Rivlin kinetics, constitutive + geometric
  stiffness, traction, sparsity pattern, CSR scatter); each function cites the
  reference lines it follows.
* ``df_ref.c``  -- the same algorithm as a plain C loop nest (the reference's element
  loop with dense B, H*B^T and B*(H*B^T) per quadrature point), built into
  ``liboracle_df.so`` by ``oracle/Makefile``; used as the timed CPU baseline and as a
  second, independently written restatement.
* ``newton``    -- the reference's incremental Newton-Raphson (StaticSolver /
  NewtonRaphson) driven by the numpy oracle and scipy's sparse direct solver.
* ``energy``    -- an independent check: total potential energy evaluated with torch
  float64 autograd, whose gradient and Hessian must equal the assembled residual and
  tangent.
* ``mlp``       -- numpy restatement of the warm-start network's forward pass, MSE
  gradient (with the BF16 operand rounding of the device products) and Adam step,
  pinned by finite differences, torch's bfloat16 conversion and torch.optim.Adam.

The reference itself cannot be compiled here (its kernels include the un-vendored
Fastor headers and cblas.h), so there is no ``oracle/_ref`` build.
"""
