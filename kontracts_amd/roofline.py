"""Algorithmic VALU work of the path kernels, for roofline accounting.

Unit: lane-ops, one VALU instruction executed by one lane.  The peak is
MI355X's published FP32 vector rate, 157.3 TFLOP/s = 78.6e12 FMA lane-ops/s
(256 CUs x 128 FP32 lanes per clock x 2.4 GHz).  The kernels issue no packed
FP32 instructions; a plain loop of independent non-packed FMAs
(kt_debug_valu, `bench.py --valu-calibration`) measures the issue rate they
can reach, and the bench reports the path kernel against both.  Not every
instruction issues at the FMA rate: the 64-bit integer multiply-add of Philox
and the transcendentals (log2, exp2, sqrt, reciprocal) take several issue
cycles each, which the lane-op count does not weight.

The counts below are the minimal instruction sequence of each step on
gfx950 (the sequence the kernels' straight-line loops execute), not the
compiler's total; the measured instruction count (rocprof SQ_INSTS_VALU,
which adds per-path set-up, the payoff and the FP64 reduction) is reported
beside them.

  Philox4x32-10: 10 rounds x (2 mad_u64_u32 + 2 three-input xor) per 4 words
                 -> 10 per 32-bit word
  normal from a word: 3 (cell -> x: shift, convert, one exact FMA)
                 + 1 (FMA 1 - x^2) + 1 (log2) + 1 (FMA to Giles' w - 2.5)
                 + 8 (central polynomial) + 1 (tail test) + 1 (scale by x)
                 = 16, plus the tail branch (12 instructions), which a wave
                 runs when any of its 64 lanes has |x| >= 0.99663
                 (P = 1 - (1 - 0.00337)^64 = 0.194) -> 18.3
  one normal: 28.3
  GBM advance (exact log step) per log-spot state: 2 (add, FMA)
  Heston advance: max, sqrt, 4 for the log-spot, 7 for the variance = 13
  basket advance, asset i (0-based): (i+1) Cholesky FMAs + 2
  knock-out monitor: running extreme, 0.5 per point and state (one
                 three-input max folds two points), compared with the level
                 once per run of points (0 per point)
  spot read (exp2 + mul) 2; Kahan sum 4; running max/min 1
"""

PEAK_LANE_OPS = 256 * 128 * 2.4e9  # 78.6e12
PEAK_FP32_TFLOPS = 157.3
PHILOX_PER_WORD = 10.0
TAIL_WAVE_PROBABILITY = 1.0 - (1.0 - 0.00337) ** 64
INVERSE_NORMAL = 16.0 + 12.0 * TAIL_WAVE_PROBABILITY
NORMAL = PHILOX_PER_WORD + INVERSE_NORMAL  # 28.3


def gbm_european() -> float:
    return NORMAL + 2


def gbm_barrier(scenarios: int = 1, states: int | None = None) -> float:
    """One path-step of a knock-out contract under GBM with `scenarios` CRN
    scenarios sharing the normal, `states` distinct log-spot paths (the
    scenarios' compares with their levels happen once per path)."""
    states = scenarios if states is None else states
    return NORMAL + 2 * states + 0.5 * states


def heston_european() -> float:
    return 2 * NORMAL + 13


def basket_european(k: int) -> float:
    return k * NORMAL + sum(i + 1 + 2 for i in range(k))


def gbm_asian() -> float:
    return NORMAL + 2 + 2 + 4
