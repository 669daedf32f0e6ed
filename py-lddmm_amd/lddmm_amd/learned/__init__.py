"""Learned warm start for surface matching: a point-cloud network predicts,
from features that need no vertex correspondence (the signed height of the
target along each template vertex normal, raw and smoothed), the
time-averaged velocity of every vertex in its local frame; a truncated
conjugate-gradient solve of K(x0) a = v turns it into a time-constant
control that starts the plain BFGS registration.  The zero control stays the
fallback, and the plain registration's final varifold energy is the level
the warm-started run is measured against."""
