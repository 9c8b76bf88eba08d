"""Named sample configurations used by the bench, the tests and the drivers."""
import numpy as np

from .config import Layer, SampleConfig, Sphere, cap_port_fraction

# spheres: radius 4 cm, sample port 0.5 cm, detector port 0.25 cm, wall 0.98;
# the source (R) / optional (T) port is taken equal to the sample port.
SPHERE_RADIUS = 4.0
SAMPLE_PORT = 0.5
DETECTOR_PORT = 0.25
WALL = 0.98


def standard_sphere():
    fs = cap_port_fraction(SAMPLE_PORT, SPHERE_RADIUS)
    fd = cap_port_fraction(DETECTOR_PORT, SPHERE_RADIUS)
    return Sphere(present=True, pw=WALL, fs=fs, fp=fs, fd=fd)


def slab_between_slides():
    """mu_a=1/cm, mu_s=100/cm, g=0.9, n=1.40, 0.1 cm, between two 0.1 cm
    glass slides (n=1.52) in air, dual spheres; roulette 1e-4, 1 in 10."""
    return SampleConfig(
        layers=[Layer.glass(1.52, 0.1), Layer(n=1.40, g=0.9, t=0.1, mu_a=1.0, mu_s=100.0),
                Layer.glass(1.52, 0.1)],
        R=standard_sphere(), T=standard_sphere(), roulette_threshold=1e-4, roulette_m=10)


def three_layer_sample():
    """Three turbid layers (no slides) between dual spheres:
    (mu_a, mu_s, g, n, t) = (0.5, 150, 0.8, 1.37, 0.01), (2, 80, 0.9, 1.40, 0.1),
    (0.2, 120, 0.85, 1.44, 0.2)."""
    return SampleConfig(
        layers=[Layer(n=1.37, g=0.8, t=0.01, mu_a=0.5, mu_s=150.0),
                Layer(n=1.40, g=0.9, t=0.1, mu_a=2.0, mu_s=80.0),
                Layer(n=1.44, g=0.85, t=0.2, mu_a=0.2, mu_s=120.0)],
        R=standard_sphere(), T=standard_sphere(), roulette_threshold=1e-4, roulette_m=10)


def single_layer(mu_a, mu_s, g=0.8, n=1.40, t=0.1, roulette_threshold=1e-4):
    """One turbid layer, no slides, dual standard spheres (lookup-grid point)."""
    return SampleConfig(layers=[Layer(n=n, g=g, t=t, mu_a=mu_a, mu_s=mu_s)],
                        R=standard_sphere(), T=standard_sphere(),
                        roulette_threshold=roulette_threshold, roulette_m=10)


def lookup_grid(n_mua=64, n_mus=64):
    """mu_a log-spaced 0.01..10 /cm, mu_s log-spaced 1..1000 /cm, g=0.8,
    n=1.40, t=0.1 cm, dual spheres.  Returns (mu_a[i], mu_s[j], configs
    in row-major (i, j) order)."""
    mua = 0.01 * 1000.0 ** (np.arange(n_mua) / max(n_mua - 1, 1))
    mus = 1.0 * 1000.0 ** (np.arange(n_mus) / max(n_mus - 1, 1))
    cfgs = [single_layer(float(a), float(s)) for a in mua for s in mus]
    return mua, mus, cfgs
