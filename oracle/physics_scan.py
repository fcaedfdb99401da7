"""Electron wavelength and raster scans (reference phaser/utils/physics.py, scan.py).
Test infrastructure only."""
import math

import numpy

# physics.py:7-24
E_REST_ENERGY = 5.1099906e5   # eV
E_REST_MASS = 9.1093837e-31   # kg
HC = 1.23984244e4             # eV * angstrom
C_LIGHT = 299792458


class Electron:
    """Relativistic electron at kinetic energy `energy` [eV] (physics.py:27-81)."""

    def __init__(self, energy):
        self.energy = energy

    @property
    def rest_energy(self):
        return E_REST_ENERGY

    @property
    def total_energy(self):
        return self.energy + E_REST_ENERGY

    @property
    def gamma(self):
        return self.energy / E_REST_ENERGY + 1.

    @property
    def mass(self):
        return self.gamma * E_REST_MASS

    @property
    def rest_mass(self):
        return E_REST_MASS

    @property
    def momentum(self):
        return math.sqrt(self.energy * (2 * E_REST_ENERGY + self.energy))

    @property
    def wavelength(self):
        return HC / self.momentum

    @property
    def beta(self):
        return math.sqrt(1 - self.gamma**-2)

    @property
    def velocity(self):
        return self.beta * C_LIGHT

    @property
    def interaction_param(self):
        m0_h2 = (E_REST_ENERGY / HC**2)
        return 2 * math.pi * self.wavelength * (self.gamma * m0_h2)


def make_raster_scan(shape, scan_step, rotation=0., *, dtype=numpy.float64):
    """Raster scan of (y, x) positions (scan.py:23-50)."""
    yy = numpy.arange(shape[0], dtype=dtype) - numpy.array(shape[0] / 2., dtype=dtype)
    xx = numpy.arange(shape[1], dtype=dtype) - numpy.array(shape[1] / 2., dtype=dtype)
    pts = numpy.stack(numpy.meshgrid(yy, xx, indexing='ij'), axis=-1)
    if rotation != 0.:
        theta = rotation * numpy.pi / 180.
        mat = numpy.array([[numpy.cos(theta), -numpy.sin(theta)],
                           [numpy.sin(theta), numpy.cos(theta)]], dtype=dtype)
        pts = pts @ mat.T
    return pts * numpy.broadcast_to(numpy.array(scan_step), (2,)).astype(dtype)
