"""FP64 CPU oracle for the MSGB forward/adjoint path -- TEST INFRASTRUCTURE ONLY.

This package is a plain NumPy (plus one C++/OpenMP file) restatement of the
reference algorithm (elma16/beamax 0.2.2).  It exists to *check* the HIP
implementation in ``beamax_amd``; it is never the thing measured or shipped.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product path (``beamax_amd``) never imports
anything from here and fails loudly when its HIP library is missing.

Modules and what pins them (each function cites the reference file:line it
follows):

* ``decomposition`` -- dyadic Fourier tiling (``beamax/decomposition.py``).
  Pinned by the reference's locked widths/counts and mirror-pair tests.
* ``transforms`` -- windowed Gaussian filters, sum of squares, packed filters,
  MSWPT analysis/synthesis, frame atoms and frame phase
  (``beamax/transforms.py``, ``beamax/utils/fft.py``,
  ``beamax/utils/arrays.py``).  Pinned by inverse(forward(x)) == x,
  partition of unity, single-coefficient == frame atom, rho=2 phase == 1.
* ``beams`` -- coefficient positions, forward beam parameters,
  ``compute_coefficients``, the closed-form homogeneous propagators
  ``solve_hom_diag`` / ``solve_hom_TR`` and the real beam sums
  (``beamax/gb/*.py``, ``forward_solver_utils.py``).  Pinned by frame atom ==
  Gaussian beam at t=0, closed form == numerically integrated Riccati ODE,
  real sum == GB + conj(GB), a brute-force double loop, and (for the
  initial-velocity coefficients) the reference's c+ == +-c- symmetries,
  its 0.5i b / H known answer and its mirror-box conjugacy.
* ``topn`` -- the ``top_n`` selection rule (``forward_solver_utils.py:65-89``)
  in two independent formulations (an emulation of ``jnp.argpartition`` built
  on ``lax.top_k`` tie semantics, and a lexicographic sort), checked against
  each other and against brute force on tie-heavy inputs.
* ``adjoint`` -- adjoint source, boundary (TR) beam parameters, spacetime to
  spatial Hessian conversion, principal B^-1 symbol and the TR sum
  (``msgb_solver.py``, ``tr_solver_utils.py``, ``adjoint_solver_utils.py``).
  Pinned by the reference's known answers for B^-1, the spectral-derivative
  closed form, inward normals, tangential phase and ODE agreement.  The
  time-reversal pipeline composes these pinned pieces (0.5 MSWPT, top_n,
  TR parameters, solve_hom_TR, TR sum, times 2) and adds no new arithmetic.
* ``cpu_beamsum.cpp`` -- the same forward/adjoint real beam sums in C++17 with
  OpenMP, FP64, threads owning disjoint output columns.  Used as the CPU
  baseline (kind "port") and as a fast checker for larger cases; itself
  checked against the NumPy restatement.

Parity unpinned: the exact-tie order of ``top_n`` against the real
``jnp.argpartition`` (jax 0.10) cannot be observed here; the rule implemented
is inferred from ``lax.top_k``'s documented lower-index-first tie order.
"""
