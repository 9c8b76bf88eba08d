"""Drop-in replacements for the reference's Python propagators.

Same positional signatures, argument meanings (angles in radians, mu_s the
non-reduced scattering coefficient, mu_t passed explicitly) and return
tuple (A, R_diffuse, R_specular, T_diffuse, T_direct) as fractions of N:

    ISMC(N, n, g, t, mu_a, mu_s, mu_t, R, R_pw, R_fs, R_fp, R_fd, R_f, R_angle,
         specular_included, T, T_pw, T_fs, T_fp, T_fd, T_f, T_angle)     # ISMC.py:164-406
    ISMLMC(N, n, g, t, mu_a, mu_s, mu_t, ...same sphere arguments...)    # ISMLMC.py:101-347

ISMC is the single-slab propagator (ISMC.py:87-160: one turbid layer,
roulette threshold 1e-3); ISMLMC takes per-layer lists (MLMC.py:76-202,
roulette threshold 1e-4, mu_s == 0 marks a glass slide).  Both run on the
GPU through libismc; the reference's unseeded `random` becomes a fresh
64-bit Philox seed per call unless ``seed=`` is given.
"""
from .config import Layer, SampleConfig, Sphere
from .engine import default_engine, fresh_seed


def _spheres(R, R_pw, R_fs, R_fp, R_fd, R_f, R_angle, T, T_pw, T_fs, T_fp, T_fd, T_f, T_angle):
    def f(x):
        return float(x) if x is not None else 0.0
    Rs = Sphere(present=bool(R), pw=f(R_pw), fs=f(R_fs), fp=f(R_fp), fd=f(R_fd),
                f=f(R_f) if R_f is not None else None, angle=f(R_angle))
    Ts = Sphere(present=bool(T), pw=f(T_pw), fs=f(T_fs), fp=f(T_fp), fd=f(T_fd),
                f=f(T_f) if T_f is not None else None, angle=f(T_angle))
    return Rs, Ts


def _run(cfg, N, seed, engine, fp64):
    eng = engine or default_engine()
    N = int(N)
    if N <= 0:
        raise ValueError("N must be a positive number of photons")
    res = eng.simulate(cfg, N, seed=fresh_seed() if seed is None else int(seed), fp64=fp64)
    return res.tuple5


def ISMC(N, n, g, t, mu_a, mu_s, mu_t, R, R_pw, R_fs, R_fp, R_fd, R_f, R_angle,
         specular_included, T, T_pw, T_fs, T_fp, T_fd, T_f, T_angle, *, seed=None,
         engine=None, fp64=False, return_result=False):
    """Single slab in air between optional spheres (ISMC.py:164)."""
    Rs, Ts = _spheres(R, R_pw, R_fs, R_fp, R_fd, R_f, R_angle, T, T_pw, T_fs, T_fp, T_fd, T_f, T_angle)
    cfg = SampleConfig(layers=[Layer(n=float(n), g=float(g), t=float(t), mu_a=float(mu_a),
                                     mu_s=float(mu_s), mu_t=float(mu_t))],
                       R=Rs, T=Ts, specular_included=bool(specular_included),
                       roulette_threshold=1e-3)
    if float(mu_s) == 0.0:
        raise ValueError("ISMC needs a scattering slab (mu_s > 0)")
    if return_result:
        eng = engine or default_engine()
        return eng.simulate(cfg, int(N), seed=fresh_seed() if seed is None else int(seed), fp64=fp64)
    return _run(cfg, N, seed, engine, fp64)


def ISMLMC(N, n, g, t, mu_a, mu_s, mu_t, R, R_pw, R_fs, R_fp, R_fd, R_f, R_angle,
           specular_included, T, T_pw, T_fs, T_fp, T_fd, T_f, T_angle, *, seed=None,
           engine=None, fp64=False, return_result=False):
    """Multilayer sample (lists from the illuminated face; mu_s == 0 is a
    glass slide) between optional spheres (ISMLMC.py:101)."""
    lists = [list(x) for x in (n, g, t, mu_a, mu_s, mu_t)]
    if len({len(x) for x in lists}) != 1:
        raise ValueError("n, g, t, mu_a, mu_s, mu_t must have the same length")
    layers = [Layer(n=float(a), g=float(b), t=float(c), mu_a=float(d), mu_s=float(e), mu_t=float(f))
              for a, b, c, d, e, f in zip(*lists)]
    Rs, Ts = _spheres(R, R_pw, R_fs, R_fp, R_fd, R_f, R_angle, T, T_pw, T_fs, T_fp, T_fd, T_f, T_angle)
    cfg = SampleConfig(layers=layers, R=Rs, T=Ts, specular_included=bool(specular_included),
                       roulette_threshold=1e-4)
    if return_result:
        eng = engine or default_engine()
        return eng.simulate(cfg, int(N), seed=fresh_seed() if seed is None else int(seed), fp64=fp64)
    return _run(cfg, N, seed, engine, fp64)
