"""Host sampling helpers of the reference's ``pysofft.wigner`` (make_wigner module)."""
import numpy as np


def create_beta_samples(n):
    """pysofft/wigner.f90:34-39: (2i+1) pi / (2n), i < n."""
    n = int(n)
    return (2 * np.arange(n) + 1) * (np.pi / (2 * n))


def create_alpha_gamma_samples(n):
    """pysofft/wigner.f90:94-100: 2 pi i / n, i < n."""
    n = int(n)
    return np.arange(n) * (2.0 * np.pi / n)


def create_trig_samples(bw):
    """pysofft/wigner.f90:51-58: columns cos(beta), cos(beta/2) sin(beta/2), cos(beta/2)^2."""
    b = create_beta_samples(2 * int(bw))
    return np.stack([np.cos(b), np.cos(b / 2) * np.sin(b / 2), np.cos(b / 2) ** 2], axis=1)
