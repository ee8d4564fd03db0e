"""Algorithmic VALU work of the path kernels, for roofline accounting.

Unit: lane-ops, one VALU instruction executed by one lane.  The peak is
MI355X's published FP32 vector rate, 157.3 TFLOP/s = 78.6e12 FMA lane-ops/s
(256 CUs x 128 FP32 lanes per clock x 2.4 GHz).  The kernels issue no packed
FP32 instructions, so this is the ceiling for their instruction stream.

The counts below are fixed from the algorithm (the minimal instruction
sequence for each step), not from the compiler's output; the measured
instruction count (rocprof SQ_INSTS_VALU) is reported beside them.

  Philox4x32-10: 10 rounds x (2 mad_u64_u32 + 2 three-input xor) per 4 words
                 -> 10 per 32-bit word
  normal from a word: 5 (cell -> x) + 1 (fma 1-x^2) + 1 (log2) + 1 (fma)
                 + 8 (central polynomial) + 1 (scale), plus the tail branch
                 (13 instructions) in the ~20% of wave-normals that need it
                 -> 20
  one normal: 30
  GBM advance (exact log step) per log-spot state: 2
  Heston advance: sqrt + 11 FMA/mul/max = 12
  basket advance, asset i (0-based): (i+1) Cholesky FMAs + 2
  knock-out monitor: running extreme 1 per point and state, plus compare and
                 select once per 4 points and scenario (0.5 per point)
  spot read (exp2 + mul) 2; Kahan sum 4; running max/min 1
"""

PEAK_LANE_OPS = 256 * 128 * 2.4e9  # 78.6e12
# Issue ceiling without packed FP32 (256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz):
# the rate one VALU instruction per lane per cycle reaches.
PEAK_NONPACKED_LANE_OPS = 256 * 64 * 2.4e9  # 39.3e12
PEAK_FP32_TFLOPS = 157.3
NORMAL = 30


def gbm_european() -> float:
    return NORMAL + 2


def gbm_barrier(scenarios: int = 1, states: int | None = None) -> float:
    """One path-step of a knock-out contract under GBM with `scenarios` CRN
    scenarios sharing the normal, `states` distinct log-spot paths."""
    states = scenarios if states is None else states
    return NORMAL + 2 * states + 1 * states + 0.5 * scenarios


def heston_european() -> float:
    return 2 * NORMAL + 12


def basket_european(k: int) -> float:
    return k * NORMAL + sum(i + 1 + 2 for i in range(k))


def gbm_asian() -> float:
    return NORMAL + 2 + 2 + 4
