"""ORACLE — pure-Python restatement of the reference's CPU propagator.

Test/benchmark infrastructure only (bench.py times it as the CPU baseline;
tests check it against the golden fixtures).  It restates ISMLMC.py:101-347
with MLMC.py:76-202 as the propagator, in the same single-threaded
`math` + `random` style as the reference, so its speed stands for the
reference's CPU path on the same host.  Unlike the C oracle it draws from
Python's MT19937 (random.random), as the reference does.
"""
import math
import random


def _fresnel(ni, nt, mu_z):
    # fresnel_snell, MLMC.py:45-54
    if abs(mu_z) > 0.99999:
        return ((nt - ni) / (nt + ni)) ** 2
    ti = math.acos(abs(mu_z))
    st = ni * math.sin(ti) / nt
    if st >= 1.0:
        return 1.0
    tt = math.asin(st)
    return 0.5 * (math.sin(ti - tt) ** 2 / math.sin(ti + tt) ** 2
                  + math.tan(ti - tt) ** 2 / math.tan(ti + tt) ** 2)


def _snell(ni, nt, mx, my, mz):
    # multilayer_snell, MLMC.py:39-42
    ti = math.acos(abs(mz))
    tt = math.asin(ni * math.sin(ti) / nt)
    return mx * ni / nt, my * ni / nt, math.copysign(math.cos(tt), mz)


def _u():
    # (0, 1] like curand_uniform; random.random() is [0, 1)
    return 1.0 - random.random()


def _mc(n, mu_a, mu_s, mu_t, g, bounds, namb, thr, m, w, x, y, z, mx, my, mz):
    # MC, MLMC.py:76-202
    A = Rf = Tf = 0.0
    L = len(n)
    layer = 0 if mz > 0 else L - 1
    s = 0.0
    while w > 0:
        if mu_s[layer] != 0:
            s = -math.log(_u()) / mu_t[layer]
        if mz < 0:
            d, nxt = (bounds[layer] - z) / mz, layer - 1
        elif mz > 0:
            d, nxt = (bounds[layer + 1] - z) / mz, layer + 1
        else:
            d, nxt = math.inf, layer
        if mu_s[layer] == 0:
            s = d
        while s >= d:
            x += d * mx
            y += d * my
            z += d * mz
            s -= d
            if nxt == L or nxt == -1:
                if _u() < _fresnel(n[layer], namb, mz):
                    mz = -mz
                else:
                    if nxt == L:
                        Tf += w
                    else:
                        Rf += w
                    mx, my, mz = _snell(n[layer], namb, mx, my, mz)
                    w = 0.0
                    break
            else:
                if _u() < _fresnel(n[layer], n[nxt], mz):
                    mz = -mz
                else:
                    mx, my, mz = _snell(n[layer], n[nxt], mx, my, mz)
                    if mu_s[nxt] != 0 and s != 0:
                        s *= mu_t[layer] / mu_t[nxt]
                    layer = nxt
            if mz < 0:
                d, nxt = (bounds[layer] - z) / mz, layer - 1
            elif mz > 0:
                d, nxt = (bounds[layer + 1] - z) / mz, layer + 1
            else:
                d, nxt = math.inf, layer
            if mu_s[layer] == 0:
                s = d
        if w == 0:
            break
        x += s * mx
        y += s * my
        z += s * mz
        if mu_s[layer] != 0:
            dW = w * mu_a[layer] / mu_t[layer]
            w -= dW
            A += dW
        if w <= thr:
            w = m * w if _u() <= 1.0 / m else 0.0
        if w > 0 and mu_s[layer] != 0:
            gg = g[layer]
            if gg == 0:
                ct = 2 * _u() - 1
            else:
                ct = (1 / (2 * gg)) * (1 + gg * gg - ((1 - gg * gg) / (1 - gg + 2 * gg * _u())) ** 2)
            phi = 2 * math.pi * _u()
            cp, sp = math.cos(phi), math.sin(phi)
            st = math.sqrt(max(0.0, 1.0 - ct * ct))
            if abs(mz) > 0.99999:
                mx, my, mz = st * cp, st * sp, (1.0 if mz > 0 else -1.0) * ct
            else:
                zs = math.sqrt(1 - mz * mz)
                mx, my, mz = (st / zs * (mx * mz * cp - my * sp) + mx * ct,
                              st / zs * (my * mz * cp + mx * sp) + my * ct,
                              -st * cp * zs + mz * ct)
    return A, Rf, Tf, mx, my, mz


def simulate(spec, N, seed=None):
    """ISMLMC(...) of ISMLMC.py:101 on a plain-dict spec; returns the
    (A, R_d, R_s, T_d, T_u) fractions of N."""
    if seed is not None:
        random.seed(seed)
    layers = spec["layers"]
    n = [L["n"] for L in layers]
    g = [L["g"] for L in layers]
    mu_a = [L["mu_a"] for L in layers]
    mu_s = [L["mu_s"] for L in layers]
    mu_t = [L.get("mu_t", L["mu_a"] + L["mu_s"]) for L in layers]
    bounds = [0.0]
    for L in layers:
        bounds.append(bounds[-1] + L["t"])
    t_tot = bounds[-1]
    namb = spec.get("n_ambient", 1.0)
    thr = spec.get("roulette_threshold", 1e-4)
    m = spec.get("roulette_m", 10)
    sph = []
    for key in ("R", "T"):
        d = spec.get(key) or {}
        fs, fp, fd = d.get("fs", 0.0), d.get("fp", 0.0), d.get("fd", 0.0)
        f = d.get("f") if d.get("f") is not None else fs + fp + fd
        pres = bool(d.get("present", 0))
        pret = d.get("pw", 0.0) * fs / (1 - d.get("pw", 0.0) * (1 - f)) if pres else -1.0
        cs = math.sqrt(1 - fp) if pres else math.sqrt(1 - 0.5 * (1 - math.cos(2 * d.get("angle", 0.0))))
        sph.append((pres, pret, cs))
    Lc = len(layers)
    tal = [0.0] * 5
    for _ in range(int(N)):
        w = 1.0
        x = y = z = 0.0
        if spec.get("specular_included", 0):
            mx, my, mz = math.sqrt(1 - 0.99 ** 2), 0.0, 0.99
        else:
            mx, my, mz = 0.0, 0.0, 1.0
        side = 0
        while w > 0:
            layer = 0 if mz > 0 else Lc - 1
            # incident_reflection, ISMLMC.py:79-93
            if mu_s[layer] == 0.0:
                nb = layer + 1 if mz > 0 else layer - 1
                n3 = n[nb] if 0 <= nb < Lc else namb
                r1 = _fresnel(namb, n[layer], mz)
                r2 = _fresnel(n[layer], n3, mz)
                Rent = r1 + (1 - r1) ** 2 * r2 / (1 - r1 * r2)
            else:
                Rent = _fresnel(namb, n[layer], mz)
            if _u() < Rent:
                leaving = True
            else:
                mx, my, mz = _snell(namb, n[layer], mx, my, mz)
                a, r, t, mx, my, mz = _mc(n, mu_a, mu_s, mu_t, g, bounds, namb, thr, m, w, x, y, z,
                                         mx, my, mz)
                tal[0] += a
                if r:
                    w, side, leaving = r, 0, True
                elif t:
                    w, side, leaving = t, 1, True
                else:
                    break
            if leaving:
                pres, pret, cs = sph[side]
                if abs(mz) >= cs:
                    tal[2 if side == 0 else 4] += w
                    break
                if pres and _u() < pret:
                    mz = _u() if side == 0 else -_u()
                    phi = 2 * math.pi * _u()
                    mx = math.cos(phi) * math.sqrt(1 - mz * mz)
                    my = math.sin(phi) * math.sqrt(1 - mz * mz)
                    x, y, z = 0.0, 0.0, (0.0 if side == 0 else t_tot)
                    continue
                tal[1 if side == 0 else 3] += w
                break
    return tuple(v / N for v in tal)
