"""Algorithmic VALU work of the path kernels, for roofline accounting.

Units are lane-op equivalents: one full-rate VALU instruction executed by one
lane counts 1; quarter-rate instructions (32x32->64 integer multiply-add,
v_exp/v_log/v_sqrt) count 4.  The peak is the VALU issue rate of MI355X:
256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz = 39.3e12 lane-ops/s (78.6 TFLOP/s when
every op is an FP32 FMA; AMD's 157.3 TFLOP/s FP32 figure counts packed FP32).

The counts are fixed from the algorithm, not from the compiler's output:
  Philox4x32-10 block: 10 rounds x (2 mad_u64_u32 x 4 + 4 xor) = 120 -> 30 / word
  normal from a word (central branch): 5 (cell index -> x) + 3 ((1-x)(1+x))
      + 4 (log) + 2 + 8 FMA (polynomial) + 2 (scale) + 2 (branch select) = 26
  one normal: 30 + 26 = 56
  GBM advance (exact log step): add + FMA = 2
  Heston advance: sqrt 4 + 9 FMA/mul/max = 13
  basket advance per asset i (0-based): (i+1) Cholesky FMAs + add + FMA
  spot read (exp2 + mul) = 5; barrier compare + knock-out update = 3
"""

PEAK_LANE_OPS = 256 * 4 * 16 * 2.4e9  # 39.3e12
NORMAL = 56
SPOT = 5
BARRIER_UPDATE = 3


def gbm_step(monitored: bool) -> int:
    return NORMAL + 2 + (SPOT + BARRIER_UPDATE if monitored else 0)


def heston_step(monitored: bool = False) -> int:
    return 2 * NORMAL + 13 + (SPOT + BARRIER_UPDATE if monitored else 0)


def basket_step(k: int) -> int:
    return k * NORMAL + sum(i + 1 + 2 for i in range(k))
