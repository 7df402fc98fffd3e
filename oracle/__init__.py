"""FP64 CPU oracle of the OCEANSAR raw-data generation path.

TEST INFRASTRUCTURE ONLY. This package is the checker, never the product:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, call it, or load its C library. Nothing in
``oceansar_amd`` imports it, and the product path raises if the HIP library
is missing instead of falling back to anything here.

Each function restates the reference algorithm from its source text
(file:line citations are to the reference repository). The random draws and
the host set-up (coefficients, geometry tables, closure series, Bragg
spectrum table, range-response table) are inputs, shared with the GPU path.

Contents and how each is pinned (see tests/test_oracle_*.py):
  surface.synth       a2  numpy complex128 ifft2 restatement; pinned by a
                          single-mode closed form (elevation, displacements,
                          slopes, curvatures, choppy forms) and Parseval.
  bragg.rcs_romeiser97 a9 facet_terms (k_b, theta, T_hh, T_vv, bad) and the
                          Elfouhaily spreading (Delta(k) included) pinned by
                          golden vectors written by the reference's own
                          _rcs_numba and spread.elfouhaily source text on tilted
                          facets (tests/golden/make_golden_bragg.py); also by
                          the flat-surface SPM closed form, SPM at the local
                          incidence for a pure range tilt, the tangent-plane
                          Bragg vector of a pure azimuth tilt, slope symmetry,
                          the bad-facet mask and spread half plane;
                          rcs_from_terms by scipy's interp1d and the k_b = 0
                          special case (inf / NaN before the mask).
  ka.field_spa        a10 pinned by the geometrical-optics RCS of a specular
                          point, pi R1 R2 |R(0)|^2 (caps, rotated axes, a
                          saddle, both polarisations), the facet-boundary cut
                          and the exp(i q.delta) phase of the stationary point.
  echo.profile_integrator / chan_profile  a13/a14  pinned by brute-force
                          sinc evaluation, a point target, edge dropping,
                          linearity and a one-bin shift.
  sar.lerp_fields     the balancer's float32 interpolation, pinned by a case
                          where FP64 evaluation rounds differently.
  sar.simulate        a1/a6/a11/a12  per-pulse loop with integer cadence;
                          built from the pinned parts above; the C
                          restatement (c/oracle_echo.c) agrees with it to 1e-12.
                          Its own geometry and phase terms are pinned by the
                          radar geometry (tests/test_oracle_pulse.py): the
                          displacement terms are the line-of-sight projections
                          of an exact 3-D range, the migration term the
                          hyperbolic range history to second order, the Bragg
                          phase rate the capillary-gravity frequency at k_b,
                          the beam a product of uniform-aperture patterns
                          (quadrature), the along-track channel phase the
                          one-way path difference b sin(az), the cross-track
                          phase b sin(inc - inc0) and its vertical-displacement
                          turn. facet_range (sr_surface, wave_dinc, including
                          the reference's horizontal term Dx sin(inc)/sr0 that no
                          closed form fixes) is pinned bit for bit by golden
                          vectors of the reference's own two statements.
  insar.coregister / l1b / ati_stats  (InSAR / ATI processors) the azimuth
                          shift pinned by a roll for integer shifts and by the
                          tone eigenvalue at the reference's rolled frequency
                          index (the half-bin offset of odd lengths included);
                          smooth1d / smooth by brute-force window sums; the
                          coherence by identical and phase-offset images;
                          uwphase by the left-edge centre of the fullest bin;
                          sr_to_geo (host) bit for bit by golden vectors of
                          the reference's own source text.
"""
