"""Brent's principal-axis minimizer (PRAXIS), written as a reverse-
communication generator so that many independent minimizations can share
batched function evaluations.

The algorithm is Brent, "Algorithms for Minimization without Derivatives"
(1973), ch. 7: line searches along a set of directions that are made
conjugate by replacing the direction of greatest improvement with the net
step of an inner loop, quadratic extrapolation along a curve through the
last three points, and periodic re-orthogonalisation of the directions by
the singular value decomposition of the scaled direction matrix.  The
reference calls it as praxis(t0, h0, n, prin, x, f) with fixed settings
ktm = 1, scbd = 1, illc = false and a Park-Miller random-step generator
seeded with 123456789 (CUDA/praxis.cpp:879-1536); the same settings are
used here.

    gen = praxis(x0, t0, h0)
    x = next(gen)              # first point to evaluate
    while True:
        try:
            x = gen.send(f(x))    # value of the last point, next point
        except StopIteration as e:
            x_min, f_min, n_eval = e.value
            break

`minimize(f, x0, t0, h0)` runs one generator against a plain function;
`minimize_batch(starts, evaluate)` advances many generators in lock step,
calling `evaluate(list_of_points)` once per round.
"""
import math

import numpy as np

MACHEP = 2.220446049250313e-16


class _ParkMiller:
    """Minimal-standard LCG (Park & Miller 1988): seed <- 16807 seed mod (2^31-1)."""

    def __init__(self, seed=123456789):
        self.seed = seed

    def __call__(self):
        k = self.seed // 127773
        self.seed = 16807 * (self.seed - k * 127773) - k * 2836
        if self.seed < 0:
            self.seed += 2147483647
        return self.seed * 4.656612875e-10


def _norm(v):
    """Euclidean norm, squares summed in index order."""
    s = 0.0
    for a in v:
        s = s + float(a) * float(a)
    return math.sqrt(s)


def _curve(l, d0, d1):
    """Weights of q0, x, q1 for the point at arc length l on the quadratic
    through q0 (at -d0), x (at 0) and q1 (at d1): Lagrange interpolation,
    with the divisions in a fixed order."""
    qa = l * (l - d1) / (d0 + d1) / d0
    qb = -(l + d0) * (l - d1) / d1 / d0
    qc = (l + d0) * l / d1 / (d0 + d1)
    return qa, qb, qc


def _hypot(x, y):
    """sqrt(x^2 + y^2) as a sqrt(1 + (b/a)^2) with a = max(|x|, |y|)."""
    a, b = (abs(y), abs(x)) if abs(x) < abs(y) else (abs(x), abs(y))
    return 0.0 if a == 0.0 else a * math.sqrt(1.0 + (b / a) * (b / a))


def _minfit(a, tol):
    """Singular value decomposition by Golub and Reinsch's method.

    Sources: G. H. Golub and C. Reinsch, "Singular value decomposition and
    least squares solutions", Numer. Math. 14 (1970) 403-420 (the procedure
    that EISPACK publishes as `minfit`), in the form R. P. Brent gives for
    PRAXIS (Algorithms for Minimization without Derivatives, Prentice-Hall
    1973, ch. 7, procedure minfit).  The arithmetic-order target is the
    reference's C++ minfit (CUDA/praxis.cpp:130-556), whose iterates
    tests/test_praxis.py compares bit for bit.

    Householder reduction of the n x n matrix `a` (list of rows) to
    bidiagonal form, accumulation of the right-hand transformations, and
    implicitly shifted QR on the bidiagonal.  On return `a` holds the right
    singular vectors as columns; the singular values (not sorted, all >= 0)
    are returned.  Rotations are applied in the same order and with the same
    sign conventions as that procedure, so its output is reproduced to the
    last bit."""
    n = len(a)
    if n == 1:
        q = [a[0][0]]
        a[0][0] = 1.0
        return q
    e = [0.0] * n
    q = [0.0] * n
    g = x = 0.0
    # Householder reduction to bidiagonal form
    for i in range(n):
        e[i] = g
        l = i + 1
        s = 0.0
        for r in range(i, n):
            s = s + a[r][i] * a[r][i]
        g = 0.0
        if tol <= s:
            f = a[i][i]
            g = math.sqrt(s)
            if 0.0 <= f:
                g = -g
            h = f * g - s
            a[i][i] = f - g
            for j in range(l, n):
                f = 0.0
                for r in range(i, n):
                    f = f + a[r][i] * a[r][j]
                f = f / h
                for r in range(i, n):
                    a[r][j] = a[r][j] + f * a[r][i]
        q[i] = g
        s = 0.0
        for j in range(l, n):
            s = s + a[i][j] * a[i][j]
        g = 0.0
        if tol <= s:
            if i < n - 1:
                f = a[i][i + 1]
            g = math.sqrt(s)
            if 0.0 <= f:
                g = -g
            h = f * g - s
            if i < n - 1:
                a[i][i + 1] = f - g
                for c in range(l, n):
                    e[c] = a[i][c] / h
                for j in range(l, n):
                    s = 0.0
                    for c in range(l, n):
                        s = s + a[j][c] * a[i][c]
                    for c in range(l, n):
                        a[j][c] = a[j][c] + s * e[c]
        y = abs(q[i]) + abs(e[i])
        x = x if y < x else y
    # accumulate the right-hand transformations
    a[n - 1][n - 1] = 1.0
    g = e[n - 1]
    l = n - 1
    for i in range(n - 2, -1, -1):
        if g != 0.0:
            h = a[i][i + 1] * g
            for r in range(l, n):
                a[r][i] = a[i][r] / h
            for j in range(l, n):
                s = 0.0
                for c in range(l, n):
                    s = s + a[i][c] * a[c][j]
                for r in range(l, n):
                    a[r][j] = a[r][j] + s * a[r][i]
        for c in range(l, n):
            a[i][c] = 0.0
        for r in range(l, n):
            a[r][i] = 0.0
        a[i][i] = 1.0
        g = e[i]
        l = i
    # diagonalise the bidiagonal form
    eps = MACHEP * x
    for k in range(n - 1, -1, -1):
        its = 0
        while True:
            its += 1
            if its > 30:
                raise RuntimeError("minfit: the QR iteration did not converge")
            split = False
            for l in range(k, -1, -1):
                if abs(e[l]) <= eps:
                    split = True
                    break
                if l > 0 and abs(q[l - 1]) <= eps:
                    break
            if not split:
                # cancel e[l] (q[l - 1] is negligible)
                c, s = 0.0, 1.0
                for i in range(l, k + 1):
                    f = s * e[i]
                    e[i] = c * e[i]
                    if abs(f) <= eps:
                        break
                    g = q[i]
                    h = _hypot(f, g)
                    q[i] = h
                    if h == 0.0:
                        g = h = 1.0
                    c = g / h
                    s = -f / h
            z = q[k]
            if l == k:
                if z < 0.0:
                    q[k] = -z
                    for r in range(n):
                        a[r][k] = -a[r][k]
                break
            # shift from the bottom 2 x 2 minor
            x = q[l]
            y = q[k - 1]
            g = e[k - 1]
            h = e[k]
            f = ((y - z) * (y + z) + (g - h) * (g + h)) / (2.0 * h * y)
            g = _hypot(f, 1.0)
            f = ((x - z) * (x + z) + h * (y / (f - g if f < 0.0 else f + g) - h)) / x
            # QR transformation
            c = s = 1.0
            for i in range(l + 1, k + 1):
                g = e[i]
                y = q[i]
                h = s * g
                g = g * c
                z = _hypot(f, h)
                e[i - 1] = z
                if z == 0.0:
                    f = z = 1.0
                c = f / z
                s = h / z
                f = x * c + g * s
                g = -x * s + g * c
                h = y * s
                y = y * c
                for r in range(n):
                    xr = a[r][i - 1]
                    zr = a[r][i]
                    a[r][i - 1] = xr * c + zr * s
                    a[r][i] = -xr * s + zr * c
                z = _hypot(f, h)
                q[i - 1] = z
                if z == 0.0:
                    f = z = 1.0
                c = f / z
                s = h / z
                f = c * g + s * y
                x = -s * g + c * y
            e[l] = 0.0
            e[k] = f
            q[k] = x
    return q


def praxis(x0, t0, h0, max_evals=100000):
    """Generator form of PRAXIS; yields points, receives f at each point,
    returns (x, f(x), evaluations) when converged."""
    x = np.array(x0, dtype=np.float64)
    n = x.size
    if n < 2:
        raise ValueError("PRAXIS needs n >= 2")
    small = MACHEP * MACHEP
    vsmall = small * small
    large = 1.0 / small
    vlarge = 1.0 / vsmall
    m2 = math.sqrt(MACHEP)
    m4 = math.sqrt(m2)
    ldfac = 0.01
    ktm = 1
    rnd = _ParkMiller()

    st = {"nf": 0, "nl": 0}
    V = np.eye(n)
    d = np.zeros(n)
    z = np.zeros(n)
    q0 = x.copy()
    q1 = x.copy()
    Q = {"qa": 0.0, "qb": 0.0, "qc": 0.0, "qd0": 0.0, "qd1": 0.0, "qf1": 0.0}

    def evaluate(point):
        st["nf"] += 1
        val = yield point.copy()
        return float(val)

    fx = yield from evaluate(x)
    Q["qf1"] = fx
    t = small + abs(t0)
    t2 = t
    dmin = small
    h = max(h0, 100 * t)
    ldt = h
    illc = False
    kt = 0

    # ---- line search along direction j (j < 0: along the quadratic curve)
    def flin(j, lam):
        if j >= 0:
            pt = x + lam * V[:, j]
        else:
            qa, qb, qc = _curve(lam, Q["qd0"], Q["qd1"])
            Q["qa"], Q["qb"], Q["qc"] = qa, qb, qc
            pt = (qa * q0 + qb * x) + qc * q1
        return (yield from evaluate(pt))

    def line_min(j, nits, d2, x1, f1, fk):
        """Brent's procedure `min`: returns (d2, x1), updates x and fx."""
        nonlocal fx, x
        sf1, sx1 = f1, x1
        k = 0
        xm = 0.0
        fm = f0 = fx
        dz = d2 < MACHEP
        s = _norm(x)
        temp = dmin if dz else d2
        t2_ = m4 * math.sqrt(abs(fx) / temp + s * ldt) + m2 * ldt
        s = m4 * s + t
        if dz and t2_ > s:
            t2_ = s
        t2_ = max(t2_, small)
        t2_ = min(t2_, 0.01 * h)
        if fk and f1 <= fm:
            xm, fm = x1, f1
        if not fk or abs(x1) < t2_:
            x1 = t2_ if x1 >= 0 else -t2_
            f1 = yield from flin(j, x1)
        if f1 <= fm:
            xm, fm = x1, f1
        while True:
            if dz:
                # evaluate at another point to estimate the second derivative
                x2 = 2.0 * x1 if f0 >= f1 else -x1
                f2 = yield from flin(j, x2)
                if f2 <= fm:
                    xm, fm = x2, f2
                d2 = (x2 * (f1 - f0) - x1 * (f2 - f0)) / ((x1 * x2) * (x1 - x2))
            d1 = (f1 - f0) / x1 - x1 * d2
            dz = True
            if d2 <= small:
                x2 = -h if d1 >= 0 else h
            else:
                x2 = (-0.5 * d1) / d2
            if abs(x2) > h:
                x2 = h if x2 > 0 else -h
            # evaluate at the predicted minimum; halve the step until success
            restart = False
            while True:
                f2 = yield from flin(j, x2)
                if k >= nits or f2 <= f0:
                    break
                k += 1
                if f0 < f1 and x1 * x2 > 0.0:
                    restart = True     # back to the second-derivative estimate
                    break
                x2 = 0.5 * x2
            if not restart:
                break
        st["nl"] += 1
        if f2 <= fm:
            fm = f2
        else:
            x2 = xm
        if abs(x2 * (x2 - x1)) > small:
            d2 = (x2 * (f1 - f0) - x1 * (fm - f0)) / ((x1 * x2) * (x1 - x2))
        elif k > 0:
            d2 = 0.0
        if d2 <= small:
            d2 = small
        x1 = x2
        fx = fm
        if sf1 < fx:
            fx = sf1
            x1 = sx1
        if j >= 0:
            x = x + x1 * V[:, j]
        return d2, x1

    def quad():
        """Quadratic extrapolation along the curve through q0, q1 and x."""
        nonlocal fx, x, q0, q1
        s = fx
        fx = Q["qf1"]
        Q["qf1"] = s
        old_x = x.copy()
        x = q1.copy()
        q1 = old_x
        Q["qd1"] = _norm(x - q1)
        lam = Q["qd1"]
        if Q["qd0"] > 0.0 and Q["qd1"] > 0.0 and st["nl"] >= 3 * n * n:
            _, lam = yield from line_min(-1, 2, 0.0, lam, Q["qf1"], True)
            qa, qb, qc = _curve(lam, Q["qd0"], Q["qd1"])
        else:
            fx = Q["qf1"]
            qa, qb, qc = 0.0, 0.0, 1.0
        Q["qd0"] = Q["qd1"]
        s_old = q0.copy()
        q0 = x.copy()
        x = (qa * s_old + qb * x) + qc * q1

    while True:
        # ---- main loop
        sf = d[0]
        d[0] = 0.0
        d[0], s = yield from line_min(0, 2, d[0], 0.0, fx, False)
        if s <= 0.0:
            V[:, 0] = -V[:, 0]
        if not (sf > 0.9 * d[0] and 0.9 * sf < d[0]):
            d[1:] = 0.0
        for k in range(1, n):
            y = x.copy()
            sf = fx
            if kt > 0:
                illc = True
            while True:
                kl = k
                df = 0.0
                if illc:
                    # random step to leave resolution valleys
                    for i in range(n):
                        s = (0.1 * ldt + t2 * (10 ** kt)) * (rnd() - 0.5)
                        z[i] = s
                        x = x + s * V[:, i]
                    fx = yield from evaluate(x)
                for k2 in range(k, n):
                    sl = fx
                    d[k2], s = yield from line_min(k2, 2, d[k2], 0.0, fx, False)
                    s = d[k2] * (s + z[k2]) ** 2 if illc else sl - fx
                    if df <= s:
                        df = s
                        kl = k2
                if not illc and df < abs((100 * MACHEP) * fx):
                    illc = True
                    continue
                break
            for k2 in range(0, k):
                d[k2], _ = yield from line_min(k2, 2, d[k2], 0.0, fx, False)
            f1 = fx
            fx = sf
            step = x - y
            x = y.copy()
            y = step
            lds = _norm(step)
            if lds > small:
                for i in range(kl - 1, k - 1, -1):
                    V[:, i + 1] = V[:, i]
                    d[i + 1] = d[i]
                d[k] = 0.0
                V[:, k] = y / lds
                d[k], lds = yield from line_min(k, 4, d[k], lds, f1, True)
                if lds <= 0.0:
                    lds = -lds
                    V[:, k] = -V[:, k]
            ldt = ldfac * ldt
            if ldt < lds:
                ldt = lds
            t2 = m2 * _norm(x) + t
            if ldt > 0.5 * t2:
                kt = -1
            kt += 1
            if kt > ktm or st["nf"] > max_evals:
                return x.copy(), fx, st["nf"]
        yield from quad()
        dn = 0.0
        for i in range(n):
            d[i] = 1.0 / math.sqrt(d[i])
            dn = max(dn, d[i])
        V = V * (d / dn)[None, :]
        # new orthogonal directions: singular vectors of the scaled direction
        # matrix (minfit on its transpose returns them as columns)
        A = [[float(V[j, i]) for j in range(n)] for i in range(n)]
        d = np.array(_minfit(A, vsmall))
        V = np.array(A)
        for i in range(n):
            dni = dn * d[i]
            if dni > large:
                d[i] = vsmall
            elif dni < small:
                d[i] = vlarge
            else:
                d[i] = 1.0 / dni / dni
        # order by decreasing value: selection sort, swapping the first
        # largest remaining entry into place
        for j1 in range(n - 1):
            j3 = j1
            for j2 in range(j1 + 1, n):
                if d[j3] < d[j2]:
                    j3 = j2
            if j1 != j3:
                d[[j1, j3]] = d[[j3, j1]]
                V[:, [j1, j3]] = V[:, [j3, j1]]
        dmin = max(d[n - 1], small)
        illc = m2 * d[0] > dmin


def minimize(f, x0, t0, h0, max_evals=100000):
    """PRAXIS on a plain function; returns (x, f(x), evaluations)."""
    gen = praxis(x0, t0, h0, max_evals=max_evals)
    pt = next(gen)
    while True:
        try:
            pt = gen.send(f(pt))
        except StopIteration as e:
            return e.value


def minimize_batch(starts, evaluate, t0, h0, max_evals=100000, on_round=None):
    """Run one PRAXIS per start point in lock step.  `evaluate(points, ids)`
    receives the list of pending points (and their solver indices) and
    returns their values; it is called once per round.  Returns a list of
    (x, f, evaluations)."""
    gens = [praxis(x0, t0, h0, max_evals=max_evals) for x0 in starts]
    pending = {i: next(g) for i, g in enumerate(gens)}
    results = [None] * len(gens)
    rounds = 0
    while pending:
        ids = sorted(pending)
        vals = evaluate([pending[i] for i in ids], ids)
        rounds += 1
        nxt = {}
        for i, v in zip(ids, vals):
            try:
                nxt[i] = gens[i].send(float(v))
            except StopIteration as e:
                results[i] = e.value
        pending = nxt
        if on_round:
            on_round(rounds, len(pending))
    return results
