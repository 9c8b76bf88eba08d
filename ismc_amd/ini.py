"""Reader for the reference's .ini input files (Lookup/example.ini ==
Minimization/example.ini), sections [sample] [spheres] [experiment]
[lookup] [minimization].

Differences from the reference drivers, each deliberate:
  * sphere angles are given in degrees and converted to radians here (the
    CUDA command lines convert, CUDA/ISMLMC.cu:220,239; the Python drivers
    pass degrees straight into the radians-expecting propagator,
    Minimization/minimization.py:111,265);
  * a missing M_U_unc leaves M_U set (minimization.py:79-82 clears M_U);
  * a measured value of 0.0 counts as measured (minimization.py:145-152
    tests truthiness);
  * the lookup grid's mu_s' axis uses mus_base (Lookup/lookup.py:126 uses
    mua_base).
"""
import configparser
import math
from dataclasses import dataclass, field
from typing import List, Optional

from .config import Layer, SampleConfig, Sphere


def _floats(raw):
    return [float(x) for x in raw.replace(",", " ").split()]


@dataclass
class SphereSpec:
    present: bool
    pw: float
    pw_unc: float
    fs: float
    fs_unc: float
    fp: float
    fp_unc: float
    fd: float
    fd_unc: float
    angle_deg: float
    angle_unc_deg: float

    def sphere(self, pw=None, fs=None, fp=None, fd=None, angle_deg=None):
        pw = self.pw if pw is None else pw
        fs = self.fs if fs is None else fs
        fp = self.fp if fp is None else fp
        fd = self.fd if fd is None else fd
        a = self.angle_deg if angle_deg is None else angle_deg
        return Sphere(present=self.present, pw=pw, fs=fs, fp=fp, fd=fd, f=fs + fp + fd,
                      angle=math.radians(a))


@dataclass
class Measurement:
    values: List[float]
    uncs: List[float]

    @property
    def first(self):
        return self.values[0] if self.values else None

    @property
    def first_unc(self):
        return self.uncs[0] if self.uncs else 0.0


@dataclass
class IniFile:
    n: float
    n_unc: float
    g: float
    t: float
    R: SphereSpec
    T: SphereSpec
    specular_included: bool
    separate: bool
    M_R: Optional[Measurement] = None
    M_T: Optional[Measurement] = None
    M_U: Optional[Measurement] = None
    M_S: Optional[Measurement] = None
    lookup: dict = field(default_factory=dict)
    minimization: dict = field(default_factory=dict)

    def sample(self, mu_a, mus_reduced, n=None, R=None, T=None, roulette_threshold=1e-4):
        """Single-layer configuration at (mu_a, mu_s'), as the lookup and
        minimization drivers build it (CUDA/ISMLMC-lookup.cu:244-250,
        CUDA/ISMLMC-minimization.cu:92-101)."""
        return SampleConfig(layers=[Layer.from_reduced(self.n if n is None else n, self.g, self.t,
                                                       mu_a, mus_reduced)],
                            R=R or self.R.sphere(), T=T or self.T.sphere(),
                            specular_included=self.specular_included,
                            roulette_threshold=roulette_threshold)

    def lookup_grid(self):
        L = self.lookup
        mua = [L["mua_start"] * L["mua_base"] ** i for i in range(int(L["N_mua"]))]
        mus = [L["mus_start"] * L["mus_base"] ** j for j in range(int(L["N_mus"]))]
        return mua, mus

    def lookup_photons(self):
        """(N, runs) of the lookup sweep: given directly, or N =
        ceil(estimate / uncertainty^2) with runs = 0 (Poisson errors;
        CUDA/ISMLMC-lookup.cu:659-673)."""
        L = self.lookup
        if "N" in L and "runs" in L:
            return int(L["N"]), int(L["runs"])
        return int(math.ceil(L.get("estimate", 1.0) / L["uncertainty"] ** 2)), 0


def _sphere(sec, p):
    g = lambda k: float(sec[p + k])
    return SphereSpec(present=bool(int(sec[p + "sphere"])), pw=g("pw"), pw_unc=g("pw_unc"),
                      fs=g("fs"), fs_unc=g("fs_unc"), fp=g("fp"), fp_unc=g("fp_unc"),
                      fd=g("fd"), fd_unc=g("fd_unc"), angle_deg=g("angle"),
                      angle_unc_deg=g("angle_unc"))


def load(path_or_text):
    cp = configparser.ConfigParser()
    cp.optionxform = str   # keys keep their case (N_mua, M_R, R_sphere, ...)
    if "\n" in path_or_text or "[" in path_or_text:
        cp.read_string(path_or_text)
    else:
        with open(path_or_text) as f:
            cp.read_file(f)
    s = cp["sample"]
    sp = cp["spheres"]
    ini = IniFile(n=float(s["n"]), n_unc=float(s.get("n_unc", "0")), g=float(s["g"]),
                  t=float(s["t"]), R=_sphere(sp, "R_"), T=_sphere(sp, "T_"),
                  specular_included=bool(int(sp.get("specular_included", "0"))),
                  separate=bool(int(sp.get("separate", "0"))))
    if cp.has_section("experiment"):
        e = cp["experiment"]
        for key, attr in (("M_R", "M_R"), ("M_T", "M_T"), ("M_U", "M_U"), ("M_S", "M_S")):
            if key in e and e[key].strip():
                unc = _floats(e[key + "_unc"]) if (key + "_unc") in e and e[key + "_unc"].strip() else []
                setattr(ini, attr, Measurement(_floats(e[key]), unc))
    for name in ("lookup", "minimization"):
        if cp.has_section(name):
            d = {}
            for k, v in cp[name].items():
                v = v.strip()
                if not v:
                    continue
                try:
                    d[k] = float(v)
                except ValueError:
                    d[k] = v
            setattr(ini, name, d)
    return ini
