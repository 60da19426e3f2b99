"""FP64 CPU oracle of the GBGHomogeneousTurbulence.jl DNS time step.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library imports, links or calls this
package: only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker the HIP path is compared against.

The reference (Julia, plus the un-vendored field library GBGPeriodicFields 0.1.x) cannot run
in this image, so the oracle restates its algorithm from the source text.  Each function
cites the reference file:line it follows.  The field-library operations the reference calls
(transforms, cross/multiply, Helmholtz projection, divergence, exponential convolution,
reality projection, weighted inner products, Catmull-Rom interpolation, displacement with
wrap) are restated from their documented contracts; where that contract leaves a bit-level
choice open (shell binning, the cell-index expression, the summation order) the oracle
states the canonical choice it makes, and the HIP library makes the same one.

Modules:
  dns      -- domain tables, 2/3 mask, right-hand sides, IF low-storage RK3/RK4 step,
              field deposit, reductions and spectra (FP64, numpy/pocketfft transforms).
  tracers  -- Catmull-Rom gather, the Heun (ABM(1)) tracer push, the wrap rule, the
              Adams weights and the tracer cell keys.
"""
