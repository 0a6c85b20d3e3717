"""CPU oracle for the spdepy GMRF hot path -- TEST INFRASTRUCTURE ONLY.

This package is a numpy/scipy restatement of the reference algorithms
(berild/spdepy): regular space-time grids and their B-spline bases, the
finite-volume diffusion/advection stencils, the block assembly of the
space-time precision Q and of its parameter derivatives, the Gaussian
log-likelihood with its gradient, and sampling.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker / the timed CPU
baseline.  The product path (``spdepy_amd``) never imports it and never
falls back to it.

Pinning (what each part is checked against, see ``tests/test_oracle_*.py``):
  * grid / bases      -- closed forms (partition of unity, Cox-de Boor on
                         the stated knots, index maps).
  * stencils          -- golden vectors written by the reference C++ compiled
                         from ``/root/reference`` (``oracle/Makefile`` ->
                         ``oracle/_ref``), committed under ``tests/golden``.
  * Q                 -- the factorised form Q = F^T D F (closed form), SPD,
                         symmetry, and pattern semantics of scipy.
  * log|Q|            -- closed form from the spatial determinant.
  * log-likelihood    -- dense brute force on small grids, finite differences.
  * gradient          -- exact trace (dense inverse) and central finite
                         differences on interior / Dirichlet cases.
  * sampling          -- F-form exact sampler vs. dense Cholesky; covariance.
Parity unpinned: CHOLMOD-permutation-dependent sample values and CHOLMOD's
own rounding (CHOLMOD is not installed here); the reference's bc=2
constant-diffusion stencil (``AcH_2D_b2.cpp`` takes a scalar ``H`` while
its Python caller passes a pointer, so its output is undefined).
"""
