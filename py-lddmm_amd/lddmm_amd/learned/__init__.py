"""Learned warm start for surface matching: a point-cloud network predicts the
initial momentum a0 of a template-target pair from the initial data-term
gradient field; a0 is shot with EPDiff into a time-dependent control that
starts the plain BFGS registration.  Zero-momentum plain registration stays
the fallback and the judge of the final varifold energy."""
