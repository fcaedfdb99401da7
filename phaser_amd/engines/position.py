"""Position solvers (phaser/engines/common/position_correction.py:13-61) with the
PositionSolver protocol of phaser/hooks/solver.py:20-75: constructed as `cls(args, props)`
(args is None, as PositionSolverHook passes it), `init_state(sim)` on the ReconsState and
`perform_update(positions, gradients, state) -> (update, state)`.  They act on the
(n_pos, 2) per-iteration update on the host, once per iteration."""
import numpy


class SteepestDescentPositionSolver:
    def __init__(self, args, props):
        self.step_size = props.step_size
        # the reference sets the maximum step from step_size (position_correction.py:16)
        self.max_step_size = props.step_size

    def init_state(self, sim):
        return None

    def perform_update(self, positions, gradients, state):
        update = self.step_size * gradients
        if self.max_step_size is not None:
            mag = numpy.linalg.norm(update, axis=-1, keepdims=True)
            update *= numpy.minimum(mag, self.max_step_size) / mag
        return update, state


class MomentumPositionSolver:
    def __init__(self, args, props):
        self.step_size = props.step_size
        self.max_step_size = props.max_step_size
        self.momentum = props.momentum

    def init_state(self, sim):
        return numpy.zeros_like(numpy.asarray(sim.scan))

    def perform_update(self, positions, gradients, state):
        update = self.step_size * gradients + self.momentum * state
        if self.max_step_size is not None:
            mag = numpy.linalg.norm(update, axis=-1, keepdims=True)
            update *= numpy.minimum(mag, self.max_step_size) / mag
        return update, update


POSITION_SOLVER_CLASSES = {'steepest_descent': SteepestDescentPositionSolver, 'momentum': MomentumPositionSolver}


def make_position_solver(props):
    """The solver a plan's `position_solver` entry names, built as the hook builds it
    (`props.position_solver(None)`, conventional/run.py:54)."""
    return POSITION_SOLVER_CLASSES[props.name](None, props)
