"""CPU oracle for the MembraneAleFem.jl per-time-step hot path.

TEST INFRASTRUCTURE ONLY.  This package is a restatement, written from the
reference's source text, of the algorithm the reference runs on its CPU path:

* ``spline``  - knot vectors, Cox--de Boor values/derivatives, Gauss points and
                the tabulated basis functions
                (src/input/Spline.jl, GaussPoint.jl, GpBasisFn.jl);
* ``mesh``    - the F_PULL mesh: IX, ID, ID_inv, LM, boundary lists, Dirichlet
                and Neumann data, plus the F_PULL initial state
                (src/input/Mesh.jl, Bc.jl, src/Input.jl);
* ``fem``     - ``calc_r_K`` (literal complex-step tangent, evaluated by the C
                library in ``csrc/oracle.c``), the sparse direct solve and the
                Newton loop ``time_step`` (src/analysis/FiniteElement.jl,
                GeoDynStress.jl, src/Analysis.jl).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or the timed CPU baseline.
The product path (``membranealefem.jl_amd``) never imports it.

Pinning (what each function is checked against) is listed in the header of each
module; a function with no independent check says "parity unpinned" there.
"""
