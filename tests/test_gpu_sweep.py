"""GPU parity sweep: 24 seeded inputs across the parameters the filters and
emptiness tests depend on -- solvent radius, alpha (negative to large),
packing density, coordinate rounding (0.1 A to unrounded FP32), radius
scale -- under both cell policies and both precision ladders, each list
against the oracle row for row.  Cases are drawn from one fixed generator,
so a failure names a reproducible case.
"""
import sys

import numpy as np
import pytest

import pac
import oracle

pytestmark = pytest.mark.gpu

SOLVENTS = [0.0, 0.5, 1.4, 2.0]
ALPHAS = [-2.0, -0.5, 0.0, 0.7, 2.5, 8.0]
DECIMALS = [1, 2, 3, -1]
REACH2_MAX = 15.0  # (largest radius)^2 + alpha: keeps the clique counts modest


def _cases():
    rng = np.random.default_rng(20240)
    out = []
    for k in range(24):
        solvent = float(SOLVENTS[k % 4])
        rscale = float(rng.choice([0.6, 1.0, 1.3]))
        rmax2 = (1.8 * rscale + solvent) ** 2
        while rmax2 - 2.0 > REACH2_MAX:
            rscale -= 0.1
            rmax2 = (1.8 * rscale + solvent) ** 2
        want = ALPHAS[(k * 5) % 6]
        alpha = max([a for a in ALPHAS if rmax2 + a <= REACH2_MAX] + [-2.0])
        alpha = float(min(want, alpha))
        out.append(dict(
            seed=k + 100,
            n=int(rng.integers(400, 4000)),
            min_sep=float(rng.choice([1.0, 1.3, 1.6])),
            solvent=solvent, alpha=alpha,
            decimals=int(DECIMALS[(k // 2) % 4]),
            rscale=round(rscale, 2),
            policy=pac.CELL_EXACT if k % 3 else pac.CELL_REFERENCE_RAW,
            prec=pac.PREC_STAGED if k % 2 else pac.PREC_PLAIN_FP64))
    return out


@pytest.mark.parametrize("case", _cases(), ids=lambda c: "seed%d" % c["seed"])
def test_sweep_vs_oracle(case):
    print("case: %s" % case, flush=True)
    sys.stdout.flush()
    b = pac.synth_globule(case["n"], None, case["min_sep"], seed=case["seed"],
                          decimals=case["decimals"])
    b[:, 3] = (b[:, 3] * np.float32(case["rscale"])).astype(np.float32)
    # the grid step 2 sqrt(maxR_raw^2 + alpha) in FP32 (the reference's grid,
    # computed under both policies): not positive -> both refuse the input
    mr = np.float32(b[:, 3].max())
    if not np.float32(np.float32(mr * mr) + np.float32(case["alpha"])) > 0:
        with pytest.raises(RuntimeError):
            oracle.run(b, case["solvent"], case["alpha"], cell_policy=case["policy"])
        with pytest.raises(pac.PacError) as e:
            pac.compute(b, case["solvent"], case["alpha"], precision=case["prec"],
                        cell_policy=case["policy"])
        assert e.value.code == pac.PAC_ERR_INVALID
        return
    o, ost, _ = oracle.run(b, case["solvent"], case["alpha"],
                           cell_policy=case["policy"])
    r = pac.compute(b, case["solvent"], case["alpha"], precision=case["prec"],
                    cell_policy=case["policy"])
    L, st = r.lists(), r.stats()
    for d in (1, 2, 3):
        assert np.array_equal(L[d], o[d]), "dim %d: gpu %s oracle %s" % (
            d, L[d].shape, o[d].shape)
    assert st.edge_candidates == ost["E_cand"]
    if case["policy"] == pac.CELL_EXACT:
        assert st.tet_alpha == ost["K_alpha"]
