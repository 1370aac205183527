"""Theory-agnostic driver layer: parameters and the ``Simulation`` class."""
