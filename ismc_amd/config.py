"""Sample / sphere description and conversion to the C-ABI struct.

Units follow the reference: lengths in cm, coefficients in 1/cm, angles in
radians at this level (the .ini loader converts degrees).  ``mu_s`` is the
scattering coefficient; ``Layer.from_reduced`` converts the reduced one
(mu_s = mu_s' / (1 - g), mu_s = mu_s' when g == 1; CUDA/ISMLMC.cu:327-333).
A layer with mu_s == 0 is a non-scattering glass slide
(CUDA/CUDAISMLMC.h:102,155).
"""
import math
from dataclasses import dataclass, field
from typing import List, Optional

from . import _lib


def cap_port_fraction(port_radius, sphere_radius):
    """Port area as a fraction of the sphere wall for a circular port of the
    given radius: the spherical-cap area (1 - sqrt(1 - (r/R)^2)) / 2.
    It reproduces the .ini values R_fs = 0.023510 and R_fp = 0.054638 for
    1.27 cm and 1.905 cm ports on the 8.382 cm sphere (Lookup/example.ini:21,25)."""
    x = port_radius / sphere_radius
    return (1.0 - math.sqrt(1.0 - x * x)) / 2.0


def mus_from_reduced(mus_reduced, g):
    """mu_s from mu_s' (CUDA/ISMLMC.cu:329-330)."""
    return mus_reduced if g == 1 else mus_reduced / (1.0 - g)


@dataclass
class Layer:
    n: float
    g: float
    t: float
    mu_a: float
    mu_s: float
    mu_t: Optional[float] = None

    @classmethod
    def from_reduced(cls, n, g, t, mu_a, mus_reduced):
        return cls(n=n, g=g, t=t, mu_a=mu_a, mu_s=mus_from_reduced(mus_reduced, g))

    @classmethod
    def glass(cls, n=1.52, t=0.1):
        return cls(n=n, g=0.0, t=t, mu_a=0.0, mu_s=0.0)

    @property
    def is_glass(self):
        return self.mu_s == 0.0


@dataclass
class Sphere:
    """An integrating sphere; with present=False only ``angle`` is used
    (the specular / unscattered acceptance half-angle, radians)."""
    present: bool = False
    pw: float = 0.0
    fs: float = 0.0
    fp: float = 0.0
    fd: float = 0.0
    f: Optional[float] = None
    angle: float = 0.0

    @property
    def f_total(self):
        return self.f if self.f is not None else self.fs + self.fp + self.fd

    def return_probability(self):
        """Power-balance probability that light in the sphere comes back to
        the sample: pw fs / (1 - pw (1 - f)) (reincidence, CUDAISMLMC.h:126-129)."""
        if not self.present:
            return 0.0
        return self.pw * self.fs / (1.0 - self.pw * (1.0 - self.f_total))


@dataclass
class SampleConfig:
    layers: List[Layer]
    R: Sphere = field(default_factory=Sphere)
    T: Sphere = field(default_factory=Sphere)
    specular_included: bool = False
    n_ambient: float = 1.0
    roulette_threshold: float = 1e-4
    roulette_m: float = 10.0

    @property
    def thickness(self):
        return sum(L.t for L in self.layers)

    def to_c(self):
        if not 1 <= len(self.layers) <= _lib.MAX_LAYERS:
            raise ValueError("need 1..%d layers" % _lib.MAX_LAYERS)
        c = _lib.Config()
        c.n_layers = len(self.layers)
        c.specular_included = int(bool(self.specular_included))
        c.n_ambient = float(self.n_ambient)
        c.roulette_threshold = float(self.roulette_threshold)
        c.roulette_m = float(self.roulette_m)
        for i, L in enumerate(self.layers):
            c.n[i] = float(L.n)
            c.g[i] = float(L.g)
            c.t[i] = float(L.t)
            c.mu_a[i] = float(L.mu_a)
            c.mu_s[i] = float(L.mu_s)
            c.mu_t[i] = float(L.mu_t if L.mu_t is not None else L.mu_a + L.mu_s)
        for name in ("R", "T"):
            s = getattr(self, name)
            cs = getattr(c, name)
            cs.present = int(bool(s.present))
            cs.pw, cs.fs, cs.fp, cs.fd = float(s.pw), float(s.fs), float(s.fp), float(s.fd)
            cs.f = float(s.f_total)
            cs.angle = float(s.angle)
        return c

    @classmethod
    def from_spec(cls, spec):
        """Build from the plain-dict form used by fixtures and tests."""
        layers = [Layer(n=L["n"], g=L["g"], t=L["t"], mu_a=L["mu_a"], mu_s=L["mu_s"],
                        mu_t=L.get("mu_t")) for L in spec["layers"]]

        def sph(d):
            d = d or {}
            return Sphere(present=bool(d.get("present", 0)), pw=d.get("pw", 0.0) or 0.0,
                          fs=d.get("fs", 0.0) or 0.0, fp=d.get("fp", 0.0) or 0.0,
                          fd=d.get("fd", 0.0) or 0.0, f=d.get("f"), angle=d.get("angle", 0.0) or 0.0)

        return cls(layers=layers, R=sph(spec.get("R")), T=sph(spec.get("T")),
                   specular_included=bool(spec.get("specular_included", 0)),
                   n_ambient=spec.get("n_ambient", 1.0),
                   roulette_threshold=spec.get("roulette_threshold", 1e-4),
                   roulette_m=spec.get("roulette_m", 10.0))


def dual_sphere_slab(mu_a, mu_s, g, n, t, slides=None, pw=0.98, sphere_radius=4.0,
                     sample_port=0.5, detector_port=0.25, source_port=None,
                     roulette_threshold=1e-4):
    """A slab between a reflection and a transmission sphere of the same
    geometry; source/optional port = sample port unless given.  ``slides`` is
    (n_glass, t_glass) for glass slides on both faces, or None."""
    fs = cap_port_fraction(sample_port, sphere_radius)
    fd = cap_port_fraction(detector_port, sphere_radius)
    fp = cap_port_fraction(source_port, sphere_radius) if source_port else fs
    sph = Sphere(present=True, pw=pw, fs=fs, fp=fp, fd=fd)
    layers = [Layer(n=n, g=g, t=t, mu_a=mu_a, mu_s=mu_s)]
    if slides:
        ng, tg = slides
        layers = [Layer.glass(ng, tg)] + layers + [Layer.glass(ng, tg)]
    return SampleConfig(layers=layers, R=sph, T=Sphere(**vars(sph)),
                        roulette_threshold=roulette_threshold)
