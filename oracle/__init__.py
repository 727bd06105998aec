"""FP64 CPU oracle for JANC's single-block 2D reacting-Euler step.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``janc_amd``) imports this
package; only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` use it, and only as the checker / the timed CPU baseline.

The oracle is a from-scratch NumPy restatement of the reference (JAX) code path
``src/janc_v2`` of luofx23/JANC.  Every function cites the reference file:line
it follows and keeps the reference's operation order and quirks (off-by-one
periodic padding, ghost cells inside the Lax-Friedrichs max, global Halley stop
test, clip order of the chemistry increment).  The reference cannot be imported
here (no JAX, no Cantera), so each function is pinned by independent checks in
``tests/test_oracle_*.py``:

===========================  ==================================================
oracle function              pinned by
===========================  ==================================================
mech.load_mechanism          closed-form unit conversions, IUPAC molecular
                             weights, NASA-7 textbook values (cp, h_f, s at
                             298.15 K)
thermo.*                     NASA-7 closed form; T -> e -> T round trip;
                             scipy brentq root of e(T) = e
reaction.*                   element and mass conservation; detailed balance
                             kf/kb = Kc from raw NASA-7 Gibbs energies;
                             Jacobian vs central finite differences;
                             dt -> 0 limit vs the explicit rate; stiff limit vs
                             scipy BDF integration of the same rates
flux.weno5_*                 exactness for cell averages of quadratics,
                             5th-order convergence, L/R mirror symmetry
flux.split_flux_lf           F+ + F- = F, F+ - F- = alpha U, alpha = global
                             NaN-ignoring max over the padded field
boundary.*                   labelled-cell index tables written from the
                             reference's slicing; sign flips per BC kind
step.*                       telescoping conservation on the reference's
                             (period N-1) periodic padding; uniform-flow
                             preservation; SSP-RK3 stability polynomial on a
                             linear right-hand side
===========================  ==================================================

Parity with the JAX reference itself (XLA fusion/FMA order, Cantera's exact
unit-conversion rounding, LAPACK pivot order inside ``jnp.linalg.solve``) is
unpinned: no JAX or Cantera exists in this environment.
"""
