"""CPU: the FP64 gradient reference used by the adjoint tests is itself
pinned: its rho matches the exact (Fraction) alpha values of the golden
predicate vectors, and its closed-form gradient matches central finite
differences on random simplices of every dimension."""
import json
import os

import numpy as np

import adjoint_ref as ref


def test_rho_matches_golden(golden_dir):
    data = json.load(open(os.path.join(golden_dir, "predicates.json")))
    checked = 0
    for v in data["filters"]:
        if v["rho"] is None:
            continue
        b = np.array([[float.fromhex(x) for x in ball] for ball in v["balls"]])
        c, lam, rho = ref.ortho(b[None, :, :3], b[None, :, 3] ** 2)
        # some vectors are nearly degenerate (rho up to ~1e7): FP64 solve
        # conditioning, far below what a wrong term or sign would give
        assert abs(rho[0] - v["rho"]) <= 1e-7 * (abs(v["rho"]) + b[:, 3].max() ** 2)
        assert abs(lam.sum() - 1.0) < 1e-12 + 1e-15 * np.abs(lam).sum()
        checked += 1
    assert checked >= 100


def test_gradient_matches_fd():
    rng = np.random.default_rng(5)
    for k in (1, 2, 3):
        done = 0
        while done < 20:
            P = rng.normal(0, 2.0, (k + 1, 3))
            r = rng.uniform(1.0, 3.0, k + 1)
            rows = np.arange(k + 1)[None, :]
            # keep well-conditioned simplices: the roundoff of a central
            # difference grows with |rho| / step
            if abs(ref.rho(P, r, rows)[0]) > 50:
                continue
            done += 1
            g = {k: np.array([1.0])}
            gp, gr = ref.gradient(P, r, {k: rows}, g)
            h = 1e-6
            for q in range(k + 1):
                for a in range(4):
                    def f(d):
                        P2, r2 = P.copy(), r.copy()
                        if a < 3:
                            P2[q, a] += d
                        else:
                            r2[q] += d
                        return ref.rho(P2, r2, rows)[0]
                    fd = (f(h) - f(-h)) / (2 * h)
                    an = gp[q, a] if a < 3 else gr[q]
                    assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (k, q, a, fd, an)
