// kt_program.h — the flat per-path op program shared by the host compiler and
// the gfx950 path kernels.
//
// A contract is lowered (kt_lower.cpp) into one short op list per time-grid
// step.  The path kernel runs the list for step k after advancing the model to
// grid point k.  Every op reads and writes a small register file of float
// slots that lives in VGPRs; slot and op indices are wave-uniform, so the
// hardware indexes the register file with s_set_gpr_idx and never touches
// scratch memory.
//
// Slot conventions: slot 0 holds 0.0, slot 1 holds 1.0.  Booleans are 0.0/1.0.
#pragma once
#include <stdint.h>

namespace kt {

enum OpCode : uint8_t {
  OP_NOP = 0,
  OP_CONST,   // r[d] = imm
  OP_SPOT,    // r[d] = spot of program asset a at this grid point
  OP_NEG,     // r[d] = -r[a]
  OP_ADD,     // r[d] = r[a] + r[b]
  OP_SUB,
  OP_MUL,
  OP_DIV,
  OP_MAX,     // fmaxf: a NaN operand yields the other (Rust f64::max)
  OP_MIN,
  OP_ADDI,    // r[d] = r[a] + imm
  OP_SUBI,    // r[d] = r[a] - imm
  OP_RSUBI,   // r[d] = imm - r[a]
  OP_MULI,
  OP_DIVI,    // r[d] = r[a] / imm
  OP_RDIVI,   // r[d] = imm / r[a]
  OP_MAXI,
  OP_MINI,
  OP_GE,      // r[d] = r[a] >= r[b]
  OP_GT,
  OP_LE,
  OP_LT,
  OP_GEI,     // r[d] = r[a] >= imm
  OP_GTI,
  OP_LEI,
  OP_LTI,
  OP_AND,     // r[d] = r[a] * r[b]
  OP_OR,      // r[d] = max(r[a], r[b])
  OP_NOT,     // r[d] = 1 - r[a]
  OP_WHEN,    // acquisition: n = r[a]*(1-r[d])*r[b]; r[d] += n; r[d+1] = n
  OP_UNTIL,   // knock-out:   r[d] = max(r[d], r[a]*r[b])
  OP_EMIT,    // pv[out] += sign * r[a] * r[b] * df[k]   (out = d, sign = imm)
  OP_RMAX,    // r[d] = max(r[d], r[a])
  OP_RMIN,    // r[d] = min(r[d], r[a])
  OP_AVG,     // Kahan: r[d] (sum) and r[d+1] (compensation) accumulate r[a]
  OP_FLAGERR, // if r[a] != 0 on any live path: raise error code (int)imm
  OP_COUNT_
};

struct KOp {
  uint8_t code, d, a, b;
  float imm;
};
static_assert(sizeof(KOp) == 8, "KOp must pack into 8 bytes");

// Per grid point: where its op list starts and how long it is.
struct KStep {
  int32_t begin;
  int32_t count;
};

// Groups of 4 grid points.  When all four points of a group run the same
// single monitor (or nothing), the kernel executes the group as straight-line
// code with no interpretation; every other group is interpreted point by point.
enum GroupKind : int32_t {
  G_GENERIC = 0,  // interpret each point's op list
  G_NONE = 1,     // no ops at any of the four points; slot = index of the first
                  // group after this run of G_NONE groups (the group count when
                  // the run reaches the end)
  G_KO_GE = 2,    // r[slot] = max(r[slot], S_asset >= imm) at each point (until/knock-out)
  G_KO_GT = 3,
  G_KO_LE = 4,
  G_KO_LT = 5,
  G_AVG = 6,      // Kahan sum of S_asset into r[slot], r[slot+1]
  G_RMAX = 7,     // r[slot] = max(r[slot], S_asset)
  G_RMIN = 8,
  G_AVG_AT = 9,   // G_AVG at the points whose bit (1 << point in the group)
                  // is set in imm's bits; the other points have no ops
};

struct KGroup {
  int32_t kind;
  int32_t slot;
  int32_t asset;
  float imm;
};

// Number of leading groups that repeat group 0's fast kind (and, for a
// monitor, its slot, asset and level).  The path kernels run such a prefix
// in a loop specialised to that kind, with no per-group dispatch; 0 when
// group 0 is interpreted.
inline int uniform_group_prefix(const KGroup* g, int n) {
  if (n <= 0 || g[0].kind == G_GENERIC) return 0;
  int i = 1;
  while (i < n && g[i].kind == g[0].kind &&
         (g[0].kind == G_NONE || (g[i].slot == g[0].slot && g[i].asset == g[0].asset && g[i].imm == g[0].imm)))
    ++i;
  return i;
}

// Groups the GBM scenario kernels can interpret one scenario at a time after
// the uniform prefix: at most this many.
constexpr int kSerialTailGroups = 2;

// 1 when the GBM scenario kernels run the program with its uniform prefix
// followed by a scenario-serial tail: the prefix is a kind with a loop of
// its own (no ops, knock-out >= or <=, running sum, running max) and at most
// kSerialTailGroups groups follow it.
inline int scenario_serial_tail(const KGroup* g, int n) {
  const int nu = uniform_group_prefix(g, n);
  if (nu <= 0) return 0;
  const int k = g[0].kind;
  const bool looped = k == G_NONE || k == G_KO_GE || k == G_KO_LE || k == G_AVG || k == G_RMAX;
  return looped && n - nu <= kSerialTailGroups ? 1 : 0;
}

// Result statistics of one output over a set of paths: count, mean and the
// sum of squared deviations (combined with Chan et al.'s pairwise formula).
struct KStats {
  double n, mean, m2;
};

constexpr int kSlotZero = 0;
constexpr int kSlotOne = 1;
constexpr int kMaxSlots = 32;   // register file size of the path kernels
constexpr int kMaxOutputs = 8;  // pv accumulators per path in one launch
constexpr int kMaxAssets = 8;

}  // namespace kt
