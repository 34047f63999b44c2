"""Host logic of bench.py (no GPU): the algorithmic byte models behind
`roofline.achieved`, checked against the per-unit figures they state."""
import bench
from angelo_amd.workloads import WORKLOADS


def test_encode_forward_bytes_per_sample():
    spec = WORKLOADS["neuralangelo-fp32"]
    L, F = 16, 2
    per_eval = 12 + 8 * L * F * 4            # position + 8 corners x 2 fp32 features
    store = (13 * L * F + 6 * 3 * L * F) * 4  # stored encodings + curvature derivatives
    assert per_eval == 1036
    want = 13 * per_eval + 39 * 4 + store
    assert bench.algorithmic_bytes("geo_fwd", spec, 1, 0, 0) == want
    # the mixed form reads fp16 rows: half the corner bytes
    mixed = bench.algorithmic_bytes("geo_fwd", spec, 1, 0, 0, mixed=True)
    assert want - mixed == 13 * 8 * L * F * 2


def test_analytic_mode_counts_seven_evaluations():
    spec = WORKLOADS["neuralangelo-analytic"]
    assert bench.evals_per_sample(spec) == 7
    assert bench.evals_per_sample(WORKLOADS["neuralangelo-fp32"]) == 13
    L, F = 16, 2
    store = (7 * L * F + 7 * 3 * L * F) * 4
    assert bench.algorithmic_bytes("geo_fwd", spec, 1, 0, 0) == \
        7 * (12 + 8 * L * F * 4) + 39 * 4 + store


def test_pair_kernels_scale_with_pairs():
    spec = WORKLOADS["neuralangelo-fp32"]
    assert bench.algorithmic_bytes("radix_sort", spec, 0, 0, 1000) == 1000 * 12 * 2
    assert bench.algorithmic_bytes("radix_sort", spec, 0, 0, 1000, mixed=True) == 1000 * 8 * 2
    assert bench.algorithmic_bytes("seg_reduce", spec, 0, 0, 1000) == 1000 * (12 + 8)
    assert bench.algorithmic_bytes("composite_fwd", spec, 10, 2, 0) == 10 * 40 + 2 * 32
    assert bench.algorithmic_bytes("march_count", spec, 10, 2, 0) is None
