"""Dose-protocol genetic algorithm: the reference's operators
(python_ga/*.py) in `operators`, the simulator-fitness and
surrogate-screened loops in `driver`."""
from . import operators  # noqa: F401
from .driver import (SimulatorFitness, run_plain_ga, run_surrogate_ga,  # noqa: F401
                     surrogate_fitness, FITNESS_OFFSET)
