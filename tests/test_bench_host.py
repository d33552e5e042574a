"""Host logic of bench.py: the unit count and the algorithmic flop formula."""
import math

import numpy as np
import pytest

import bench


@pytest.mark.parametrize("ng", [1, 2, 5, 9, 17])
def test_quartet_class_count_is_q_ref(ng):
    L = list(np.random.default_rng(ng).integers(0, 4, size=ng))
    cnt = bench.quartet_classes(L)
    P = ng * (ng + 1) // 2
    assert sum(cnt.values()) == P * (P + 1) // 2 + math.comb(ng, 3)


def test_md_terms_s_and_p():
    # s s s s: one GTO quadruple with one MD term per axis
    assert bench._class_md_terms(0, 0, 0, 0) == 1
    # p s s s: 3 quadruples, the p axis has (1+0+1) terms
    assert bench._class_md_terms(1, 0, 0, 0) == 3 * 2


def test_xy6_flops():
    """XY6 (X 4s3p2d1f, Y 4s2p2d1f, 64 shells), 512 q."""
    X = [0] * 4 + [1] * 3 + [2] * 2 + [3]
    Y = [0] * 4 + [1] * 2 + [2] * 2 + [3]
    L = X + Y * 6
    assert len(L) == 64
    W, herm, ser, Q = bench.algorithmic_flops(L, 512)
    assert Q == 2205904
    assert W == herm + ser
    assert herm == pytest.approx(6.3e11, rel=0.02)
    assert ser == pytest.approx(4.7e10, rel=0.02)


def test_profile_figures_are_read():
    """The bench line's k_contract traffic and executed flops come from the
    committed XY6 profile; other workloads report none."""
    import bench
    c = bench.contract_counters("xy6")
    assert c is not None
    assert c["bytes"] > 2.798e9  # at least the Z triangle
    assert 1e10 < c["flops"] < 6.811e11  # below the algorithmic count
    assert bench.contract_counters("h2") is None


def test_cpu_model_is_named():
    assert isinstance(bench._cpu_model(), str) and bench._cpu_model()
