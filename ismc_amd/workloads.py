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


def reference_example_spheres():
    """The spheres of the reference's --example runs (Cook et al. 2019;
    CUDA/ISMLMC.cu:123-138, CUDA/ISMLMC-lookup.cu:559-574): wall 0.99, R
    sphere ports (sample, source, detector) 0.023510 / 0.054638 / 0.023510,
    T sphere 0.054638 / 0.023510 / 0.023510; with a sphere absent the
    angle thresholds are 0 (R) and 0.22 degrees (T)."""
    R = Sphere(present=True, pw=0.99, fs=0.023510, fp=0.054638, fd=0.023510, angle=0.0)
    T = Sphere(present=True, pw=0.99, fs=0.054638, fp=0.023510, fd=0.023510,
               angle=0.22 * np.pi / 180)
    return R, T


def wang_three_layer_example():
    """The reference's multilayer --example (CUDA/ISMLMC.cu:140-157, sample
    of Wang et al. 1995): three layers of n 1.37, (g, t, mu_a, mu_s') =
    (0.9, 0.1, 1, 10), (0, 0.1, 1, 10), (0.7, 0.2, 2, 3), mu_s = mu_s' /
    (1 - g) (:324-331: an isotropic middle layer keeps mu_s = mu_s'),
    between the example spheres."""
    R, T = reference_example_spheres()
    rows = ((0.9, 0.1, 1.0, 10.0), (0.0, 0.1, 1.0, 10.0), (0.7, 0.2, 2.0, 3.0))
    return SampleConfig(layers=[Layer.from_reduced(1.37, g, t, a, s) for g, t, a, s in rows],
                        R=R, T=T, roulette_threshold=1e-4, roulette_m=10)


def lookup_example():
    """The reference's lookup --example (CUDA/ISMLMC-lookup.cu:576-591):
    keyword arguments of lookup.sample_space for a 13 x 11 grid, mu_a =
    0.5 * 1.2^i, mu_s' = 10 * 1.2^j, n 1.4, g 0.5, t 0.135 cm, 1e5 packets
    x 10 runs per point, separate R-sphere and T-sphere simulations."""
    R, T = reference_example_spheres()
    return {"n": 1.4, "g": 0.5, "t": 0.135, "mua": 0.5 * 1.2 ** np.arange(13),
            "mus_reduced": 10.0 * 1.2 ** np.arange(11), "R": R, "T": T, "N": 100_000, "runs": 10,
            "separate": True}
