"""Far-field noise models (reference phaser/engines/common/noise_models.py).
Test infrastructure only."""
import numpy


class AmplitudeNoiseModel:
    """Amplitude (offset 0) and Anscombe (offset 0.375) models (noise_models.py:13-72)."""

    def __init__(self, eps=1e-3, gaussian_variance=0.1, offset=0.0):
        self.offset = offset
        self.gaussian_variance = gaussian_variance
        self.var = 1 + gaussian_variance
        self.eps = eps

    def name(self):
        return "amplitude" if self.offset == 0.0 else "anscombe"

    def calc_loss(self, model_wave, model_intensity, exp_patterns, mask):
        """2 sum mask (sqrt(D+c) - sqrt(I+c) - eps)^2 / var (noise_models.py:29-45)."""
        patterns = numpy.maximum(exp_patterns, 0.0)
        return (2. * numpy.sum(mask * (numpy.sqrt(patterns + self.offset)
                                       - numpy.sqrt(model_intensity + self.offset) - self.eps)**2)
                / self.var).astype(exp_patterns.dtype)

    def calc_wave_update(self, model_wave, model_intensity, exp_patterns, mask):
        """chi_hat = (sqrt(D+c)/(sqrt(I+c)+eps) - 1) * mask * Psi (noise_models.py:47-63)."""
        patterns = numpy.maximum(exp_patterns, 0.0)
        update = numpy.sqrt(patterns + self.offset) / (numpy.sqrt(model_intensity + self.offset) + self.eps) - 1.0
        update *= mask
        return update * model_wave


def AnscombeNoiseModel(eps=1e-3, gaussian_variance=0.1, offset=0.375):
    return AmplitudeNoiseModel(eps=eps, gaussian_variance=gaussian_variance, offset=offset)


class PoissonNoiseModel:
    """Poisson loss; no wave update (noise_models.py:75-111)."""

    def __init__(self, eps=1e-3):
        self.eps = eps

    def name(self):
        return "poisson"

    def calc_loss(self, model_wave, model_intensity, exp_patterns, mask):
        patterns = numpy.maximum(exp_patterns, 0.0)
        return numpy.sum(mask * (model_intensity + self.eps + patterns * (
            numpy.log(patterns + self.eps) - numpy.log(model_intensity + self.eps) - 1.0
        ))).astype(exp_patterns.dtype)

    def calc_wave_update(self, *args):
        raise NotImplementedError()
