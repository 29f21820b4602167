"""Host-side pieces of the learned start-table depth (no GPU): the training
loss gradient against central finite differences, the size rule restated
for the A/B against the library's rule (pack_query_layout), and the cost
vector layout gahhg_set_table_model takes."""
import numpy as np

from gahhg_amd import learned_tables as lt, TABLE_BINS


def test_block_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    ng = 3 * lt.NC
    y = rng.normal(0.0, 1.0, ng * 42)
    S = rng.uniform(0.0, 1.0, (ng, 42))
    S /= S.sum(axis=1, keepdims=True)
    T = rng.normal(0.0, 1.0, ng)
    W = rng.uniform(0.5, 2.0, ng)
    _, dy = lt._block_loss_grad(y, S, T, W=W)
    h = 1e-6
    for i in rng.choice(len(y), 40, replace=False):
        yp, ym = y.copy(), y.copy()
        yp[i] += h
        ym[i] -= h
        num = (lt._block_loss_grad(yp, S, T, W=W)[0] - lt._block_loss_grad(ym, S, T, W=W)[0]) / (2 * h)
        assert abs(num - dy.reshape(-1)[i]) <= 1e-6 + 1e-5 * abs(num)


def test_block_loss_zero_at_exact_fit():
    rng = np.random.default_rng(1)
    ng = 2 * lt.NC
    y = rng.normal(0.0, 1.0, (ng, 42))
    S = np.full((ng, 42), 1.0 / 42)
    T = np.log2((S * np.exp2(y)).sum(axis=1))
    loss, dy = lt._block_loss_grad(y.reshape(-1), S, T)
    assert loss < 1e-20 and np.abs(dy).max() < 1e-12


def test_size_rule_candidate():
    # 6 B^2 <= 8 N, B in 16..128: N = 192 -> 16; N = 768 -> 32; N = 3072 -> 64; large -> 128
    assert TABLE_BINS[lt.size_rule_candidate(10)] == 16
    assert TABLE_BINS[lt.size_rule_candidate(767)] == 16
    assert TABLE_BINS[lt.size_rule_candidate(768)] == 32
    assert TABLE_BINS[lt.size_rule_candidate(3072)] == 64
    assert TABLE_BINS[lt.size_rule_candidate(12288)] == 128
    assert TABLE_BINS[lt.size_rule_candidate(10 ** 7)] == 128


def test_cost_vector_layout():
    cost = {"a_ns": list(range(lt.NC)), "b_ns_per_edge": [0.5 * c for c in range(lt.NC)], "batch_ns_per_bin": 0.25}
    v = lt.cost_vector(cost, 1e6, 2.0)
    assert len(v) == 3 + 2 * lt.NC
    assert v[:2] == [1e6, 2.0] and v[2 + 2 * 3] == 3 and v[3 + 2 * 3] == 1.5 and v[-1] == 0.25
