"""Engines of the MI355X backend: conventional (LSQ-ML, ePIE) and gradient."""
