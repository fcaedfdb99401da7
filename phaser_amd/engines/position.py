"""Position solvers (phaser/engines/common/position_correction.py:13-61).  They act on the
(n_pos, 2) per-iteration update on the host, once per iteration."""
import numpy


class SteepestDescentPositionSolver:
    def __init__(self, props):
        self.step_size = props.step_size
        # the reference sets the maximum step from step_size (position_correction.py:16)
        self.max_step_size = props.step_size

    def init_state(self, scan):
        return None

    def perform_update(self, positions, gradients, state):
        update = self.step_size * gradients
        if self.max_step_size is not None:
            mag = numpy.linalg.norm(update, axis=-1, keepdims=True)
            update *= numpy.minimum(mag, self.max_step_size) / mag
        return update, state


class MomentumPositionSolver:
    def __init__(self, props):
        self.step_size = props.step_size
        self.max_step_size = props.max_step_size
        self.momentum = props.momentum

    def init_state(self, scan):
        return numpy.zeros_like(scan)

    def perform_update(self, positions, gradients, state):
        update = self.step_size * gradients + self.momentum * state
        if self.max_step_size is not None:
            mag = numpy.linalg.norm(update, axis=-1, keepdims=True)
            update *= numpy.minimum(mag, self.max_step_size) / mag
        return update, update


def make_position_solver(props):
    return {'steepest_descent': SteepestDescentPositionSolver, 'momentum': MomentumPositionSolver}[props.name](props)
