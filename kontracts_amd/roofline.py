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


# Pathwise-Greek and book counts.  A fixing point of a running sum adds the
# spot read and the Kahan update (FIXING); with pathwise Greeks it adds the
# spot's sensitivities into the accumulator's (GBM: d/ds0, d/dsigma, d/dmu,
# 4 FMAs) or, under Heston, the reverse sweep's seed (exp2, mul, 2 FMAs).
FIXING = 2 + 4
FIXING_PATHWISE = FIXING + 4
# Heston reverse sweep per point.  The recompute of a segment regenerates
# both normals and, without running accumulators, steps the variance alone:
# max, sqrt and the 7 variance instructions = 9 (with accumulators the full
# step, 13).  The sweep on a grid of equal steps without accumulators, its
# per-point constants folded out of the loop: vp = sv^2 (1), the correlated
# normal (2), the floor's cut of vbar (2), vbar sv (1), the kappa, theta,
# sigma_v and rho sums (2 + 1 + 1 + 2), half of d/d sqrt(v) (3), its
# reciprocal with the cut (3), d/dv of the step (2), vbar (2) = 22.  The
# general sweep (tables read per point, or running accumulators) forms every
# term per point: 28, plus the log-spot walk-back (4) with accumulators.
HESTON_RECOMPUTE_VARIANCE = 9.0
HESTON_REVERSE = 22.0
HESTON_REVERSE_GENERAL = 28.0
HESTON_REVERSE_FEEDS = 4.0


def heston_barrier() -> float:
    return heston_european() + 0.5


def gbm_pathwise() -> float:
    """Price plus d/d{s0, sigma, mu}: the Brownian sum W += sqrt(dt) w per point."""
    return gbm_european() + 1


def basket_pathwise(k: int) -> float:
    return basket_european(k) + k


def heston_pathwise(feeds: bool = False) -> float:
    """Forward pass, recompute of every segment from its checkpoint (both
    normals regenerated) and the reverse sweep: the equal-step sweep without
    running accumulators, or the general sweep with the log-spot recomputed
    and walked back."""
    if feeds:
        return 2 * heston_european() + HESTON_REVERSE_GENERAL + HESTON_REVERSE_FEEDS
    return heston_european() + 2 * NORMAL + HESTON_RECOMPUTE_VARIANCE + HESTON_REVERSE


def book_job(kind: str, model: str, n_steps: int, n_fixings: int = 0, greeks: bool = False) -> float:
    """Lane-ops per path-step of one job of the 1,000-contract book
    (kontracts_amd.book): kind european / asian / barrier / basket, model
    gbm / heston / basket5; with greeks the Lipschitz kinds run the pathwise
    kernels and the barriers the price-only kernel."""
    fix = 0.0
    if kind == "asian":
        fix = n_fixings * (FIXING_PATHWISE if greeks else FIXING) / max(n_steps, 1)
    if kind == "barrier":
        return (gbm_barrier() if model == "gbm" else heston_barrier())
    if kind == "basket":
        return basket_pathwise(5) if greeks else basket_european(5)
    if model == "gbm":
        return (gbm_pathwise() if greeks else gbm_european()) + fix
    return (heston_pathwise(feeds=kind == "asian") if greeks else heston_european()) + fix
