"""The surrogate-seeded inverse solver on the GPU with the packaged network:
what the trained network gains over the same solver without it, and the
MC judge with its fallback."""
import numpy as np
import pytest

from ismc_amd import inverse as I
from ismc_amd.surrogate import Surrogate, packaged_weights

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(engine):
    sur = Surrogate.load(packaged_weights())
    meas = I.make_measurements(256, 1_000_000, seed=7, engine=engine)
    probs = [I.problem(r, t) for r, t in zip(meas["R_d"], meas["T_d"])]
    return sur, meas, probs


def test_network_gain(engine, setup):
    sur, meas, probs = setup
    truth = np.stack([meas["mu_a"], meas["mus_reduced"]], 1)
    # with the network: its seeds, then 6 batched MC launches of defect
    # correction with its Jacobian
    seeds = I.surrogate_seeds(sur, meas["R_d"], meas["T_d"])
    sol_s, nev_s, _ = I.refine_newton(probs, seeds, sur, packets=100_000, engine=engine, stream_base=1000)
    # without it: the same batched MC minimization (PRAXIS on the MC cost,
    # 1e5 packets per evaluation) from the generic guess, up to 400 evaluations
    sol_g, nev_g, _ = I.refine_batch(probs, [I.GENERIC_GUESS] * len(probs), packets=100_000, step=1.0,
                                     max_evals=400, engine=engine, stream_base=1000)
    ea_s, es_s = I.errors(sol_s, truth[:, 0], truth[:, 1])
    ea_g, es_g = I.errors(sol_g, truth[:, 0], truth[:, 1])
    cost_truth = I.fit_cost(probs, truth, 1_000_000, engine)
    fit_s = np.mean(I.fit_cost(probs, sol_s, 1_000_000, engine) <= cost_truth + I.FIT_TOL)
    fit_g = np.mean(I.fit_cost(probs, sol_g, 1_000_000, engine) <= cost_truth + I.FIT_TOL)
    print("with the network: %.0f MC evaluations per solve, median mu_a err %.3f, mu_s' err %.3f, fit %.3f"
          % (nev_s.mean(), np.median(ea_s), np.median(es_s), fit_s))
    print("without: %.0f MC evaluations per solve, median mu_a err %.3f, mu_s' err %.3f, fit %.3f"
          % (nev_g.mean(), np.median(ea_g), np.median(es_g), fit_g))
    assert nev_s.mean() * 10 < nev_g.mean()
    assert np.median(es_s) < np.median(es_g) and np.median(ea_s) < np.median(ea_g)
    assert fit_s > fit_g and fit_s > 0.7


@pytest.mark.parametrize("fallback", ["mc_newton", "praxis"])
def test_judge_flags_misses_and_fallback_repairs_them(engine, setup, fallback):
    sur, meas, probs = setup
    k = 32
    sub = probs[:k]
    seeds = I.surrogate_seeds(sur, meas["R_d"][:k], meas["T_d"][:k])
    sol, _, _ = I.refine_newton(sub, seeds, sur, packets=100_000, engine=engine, stream_base=2000)
    # corrupt every other solution: absorption x5, scattering / 3
    bad = sol.copy()
    bad[::2] *= [5.0, 1 / 3.0]
    miss, rR, rT, _ = I.check(sub, bad, engine=engine)
    assert miss[::2].all()
    # the fallback from the judged point brings them back: Gauss-Newton on
    # the MC residuals (the pipeline's) or batched MC PRAXIS
    if fallback == "mc_newton":
        fb, nev, _ = I.refine_mc_newton([sub[i] for i in range(0, k, 2)], bad[::2], packets=100_000,
                                        engine=engine, stream_base=3000)
        assert nev.max() <= 24
    else:
        fb, _, _ = I.refine_batch([sub[i] for i in range(0, k, 2)], bad[::2], packets=100_000, step=0.3,
                                  tol=3e-3, max_evals=150, engine=engine, stream_base=3000)
    before = I.fit_cost([sub[i] for i in range(0, k, 2)], bad[::2], 1_000_000, engine)
    after = I.fit_cost([sub[i] for i in range(0, k, 2)], fb, 1_000_000, engine)
    print("fit cost of the corrupted solutions: median %.3f -> %.4f" % (np.median(before), np.median(after)))
    assert np.median(after) < 0.1 * np.median(before)


def test_grid_argmin_equals_numpy(engine):
    """The GPU global seed search returns numpy's argmin of the same FP64
    cost for every measurement: random grids, exact ties (first index
    wins), a NaN cost (first NaN wins) and a T of 0 (floored at 1e-6)."""
    from test_inverse_host import np_grid_argmin
    rng = np.random.default_rng(11)
    G, M = 128 * 128, 3000
    pR = rng.uniform(0.0, 0.6, G).astype(np.float32)
    pT = rng.uniform(0.0, 0.6, G).astype(np.float32)
    R = rng.uniform(0.02, 0.6, M)
    T = rng.uniform(0.0, 0.6, M)
    T[:5] = 0.0
    # ties: cells 7000 and 9000 equal to cell 100 exactly, and a
    # measurement sitting on them
    pR[7000] = pR[9000] = pR[100]
    pT[7000] = pT[9000] = pT[100]
    R[10], T[10] = float(pR[100]), float(pT[100])
    # a tie placed before the first occurrence: the lower index must win
    pR[50], pT[50] = pR[9000], pT[9000]
    k = I.grid_argmin(pR, pT, R, T, device=engine.device)
    ref = np_grid_argmin(pR, pT, R, T)
    assert np.array_equal(k, ref)
    assert k[10] == 50
    # a NaN cost: numpy's argmin takes the first NaN
    pR2 = pR.copy()
    pR2[4000] = np.nan
    pR2[12000] = np.nan
    assert np.array_equal(I.grid_argmin(pR2, pT, R[:64], T[:64], device=engine.device),
                          np_grid_argmin(pR2, pT, R[:64], T[:64]))
    assert np.all(I.grid_argmin(pR2, pT, R[:64], T[:64], device=engine.device) == 4000)


def test_full_size_pipeline(engine):
    # The inverse workload at its full size: 4096 synthetic measurements
    # (1e6 packets, 1 % noise) through the surrogate's seeds, up to six
    # batched MC defect-correction launches (a solve whose step falls below
    # 0.1 % stops earlier), the two-stage MC judge and the
    # fallback on its misses.  Checked: the solutions fit the measurements
    # as well as the truth does in most cases, the recovered scattering is
    # accurate and the judge's misses are few and mostly repaired.
    sur = Surrogate.load(packaged_weights())
    r = I.run(sur, count=4096, methods=("newton",), with_ablation=False, plain_count=1, engine=engine,
              log=lambda *a: None)
    s = r["seeded_newton"]
    print({k: s[k] for k in ("solves_per_s", "misses", "misses_repaired_by_fallback", "fits_as_well_as_truth")},
          s["errors"])
    assert s["launches"] <= 6 and 2 < s["mean_evals"] < 6
    assert s["fits_as_well_as_truth"] > 0.78
    assert s["errors"]["median_mus_reduced"] < 0.025 and s["errors"]["median_mu_a"] < 0.12
    assert s["misses"] < 100 and s["misses_repaired_by_fallback"] >= s["misses"] // 3
