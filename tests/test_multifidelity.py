"""CPU: the multi-fidelity designer's foundry rules against per-pixel loop
restatements of their definition, and its screen -> verify selection
against brute force over the same evaluations (toy fidelities)."""

import numpy as np
import pytest
import torch

from diffnano_amd.multifidelity import (CostModel, FidelityOracle, FoundryConstraints,
                                        MultiFidelityDesigner)


def _loop_pool(line, k, op):
    # out[j] = op(line[max(0, j - k//2) : j + k//2 + 1] > 0.5); k <= 1: unchanged
    if k <= 1:
        return line.copy()
    h = k // 2
    return np.array([1.0 if op(line[max(0, j - h):j + h + 1] > 0.5) else 0.0
                     for j in range(len(line))])


def _loop_project(d, kf, ks):
    d = d.copy()
    for axis in (0, 1):
        for i in range(d.shape[1 - axis]):
            line = d[:, i] if axis == 0 else d[i, :]
            line = _loop_pool(_loop_pool(line, kf, np.all), kf, np.any)
            line = _loop_pool(_loop_pool(line, ks, np.any), ks, np.all)
            if axis == 0:
                d[:, i] = line
            else:
                d[i, :] = line
    return d


def _loop_check(d, kf, ks):
    b = d > 0.5
    viol = total = 0
    for lines in (b, b.T):
        for row in lines:
            j = 0
            while j < len(row):
                k = j
                while k < len(row) and row[k] == row[j]:
                    k += 1
                total += 1
                viol += int((k - j) < (kf if row[j] else ks))
                j = k
    return viol, total


@pytest.mark.parametrize("kf_nm,ks_nm,px", [(3.0, 2.0, 1.0), (40.0, 40.0, 10.0), (1.0, 5.0, 1.0),
                                            (2.0, 1.0, 1.0)])
def test_foundry_projection_and_check_match_loops(kf_nm, ks_nm, px):
    fc = FoundryConstraints(kf_nm, ks_nm, px)
    rng = np.random.default_rng(int(kf_nm * 10 + ks_nm))
    for shape in [(9, 13), (16, 16), (1, 7)]:
        d = rng.random(shape)
        got = fc.project(torch.from_numpy(d)).numpy()
        ref = _loop_project(d, fc.min_feature_px, fc.min_space_px)
        assert np.array_equal(got, ref)
        for arr in (d, ref):
            r = fc.check(torch.from_numpy(arr))
            v, n = _loop_check(arr, fc.min_feature_px, fc.min_space_px)
            assert r["violations"] == v and r["violation_rate"] == pytest.approx(v / n)
            assert r["passed"] == (v == 0)
    # batch and (1, H, W) forms
    stack = torch.from_numpy(rng.random((3, 8, 10)))
    pb = fc.project_batch(stack)
    for i in range(3):
        assert torch.equal(pb[i], fc.project(stack[i]))
    assert fc.project(stack[:1]).shape == (1, 8, 10)


def test_foundry_pixel_rounding():
    fc = FoundryConstraints(min_feature_nm=45.0, min_space_nm=40.0, pixel_size_nm=20.0)
    assert fc.min_feature_px == 3 and fc.min_space_px == 2
    assert FoundryConstraints(0.1, 0.1, 1.0).min_feature_px == 1


def test_cost_model():
    cm = CostModel(cost_low=1.0, cost_high=10.0, budget=25.0)
    assert cm.can_afford("high", 2) and not cm.can_afford("high", 3)
    cm.consume("high", 2)
    cm.consume("low", 5)
    assert cm.spent == 25.0 and cm.remaining == 0.0 and cm.counts == {"low": 5, "high": 2}
    with pytest.raises(ValueError, match="budget exceeded"):
        cm.consume("low", 1)


class _Toy:
    """Deterministic toy fidelities of a (S, S) design: low = a smooth
    response, high = low plus a design-dependent correction."""

    def __init__(self, batched, high=False):
        self.batched = batched
        self.high = high
        self.calls = 0

    def _one(self, d):
        x = torch.as_tensor(d, dtype=torch.float64)
        base = torch.stack([x.mean(0), x.mean(1), (x * x).mean(0)])
        return base + (0.3 * torch.sin(7 * x).mean(0) if self.high else 0)

    def __call__(self, d):
        self.calls += 1
        if self.batched:
            return torch.stack([self._one(v) for v in d])
        return self._one(d)


@pytest.mark.parametrize("batched", [False, True])
def test_designer_selection_matches_brute_force(batched):
    lo, hi = _Toy(batched), _Toy(batched, high=True)
    fc = FoundryConstraints(2.0, 2.0, 1.0)
    cm = CostModel(1.0, 20.0, budget=1e9)
    des = MultiFidelityDesigner(FidelityOracle(lo, hi), cm, fc)
    target = torch.linspace(0.2, 0.8, 12).repeat(3, 1).to(torch.float64)
    torch.manual_seed(11)
    out = des.design(target, n_initial=40, n_top=6, grid_size=12)
    # brute force with the same candidates, evaluated one by one
    torch.manual_seed(11)
    cand = torch.rand(40, 12, 12, dtype=torch.float64)
    assert torch.equal(cand, out["candidates"])
    proj = np.stack([_loop_project(c.numpy(), 2, 2) for c in cand])
    low = np.array([-((lo._one(p).numpy() - target.numpy()) ** 2).mean() for p in proj])
    order = np.argsort(-low, kind="stable")[:6]
    assert sorted(order.tolist()) == sorted(np.argsort(-low)[:6].tolist())   # distinct scores
    high = np.array([-((hi._one(proj[i]).numpy() - target.numpy()) ** 2).mean() for i in order])
    best = proj[order[int(np.argmax(high))]]
    assert np.array_equal(out["best"].numpy(), best)
    assert float(out["best_score"]) == pytest.approx(high.max(), rel=1e-6)
    assert np.allclose(out["screen_scores"].numpy(), low[order])
    assert out["screened"].shape == (6, 12, 12) and out["hf_responses"].shape == (6, 3, 12)
    # (one pass along each axis need not satisfy both axes' rules at once;
    # the reports are the check of each projected design)
    assert [r["violations"] for r in out["foundry_reports"]] == \
        [_loop_check(p, 2, 2)[0] for p in proj]
    assert cm.counts == {"low": 40, "high": 6}
    if batched:                      # one call per stage for batched evaluators
        assert lo.calls == 1 and hi.calls == 1
    # screen_candidates ranks by the response's sum
    sc, ss = des.screen_candidates(cand[:10], 3)
    sums = np.array([lo._one(p).numpy().sum() for p in proj[:10]])
    assert np.allclose(ss.numpy(), np.sort(sums)[::-1][:3])


def test_designer_budget_is_only_consumed_when_affordable():
    lo, hi = _Toy(True), _Toy(True, high=True)
    cm = CostModel(1.0, 50.0, budget=100.0)
    des = MultiFidelityDesigner(FidelityOracle(lo, hi), cm, FoundryConstraints(1.0, 1.0, 1.0))
    torch.manual_seed(0)
    des.design(torch.zeros(3, 8), n_initial=20, n_top=4, grid_size=8)
    # low: 20 fit; high: 4 x 50 = 200 exceed what is left, so nothing is charged
    assert cm.counts == {"low": 20, "high": 0}
    with pytest.raises(ValueError, match="fidelity"):
        FidelityOracle(lo, hi).evaluate(torch.zeros(8, 8), "medium")
