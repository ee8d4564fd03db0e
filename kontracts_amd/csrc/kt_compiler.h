// kt_compiler.h — host side of the engine: contract AST, JSON reader, plan and
// time grid, and the lowering of a contract to the per-step op program.
#pragma once
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "kt_program.h"

namespace kt {

// Error raised by the host compiler; code is a kt_status value.
struct Error {
  int code;
  std::string msg;
};
[[noreturn]] void raise(int code, const std::string& msg);

// ---- AST (kontract src/ast.rs:34-206) -------------------------------------
struct Obs {
  // AverageAt is an extension (no reference counterpart): the arithmetic mean
  // of obs over fixing dates, {"AverageAt":{"obs":...,"dates":[t1,...]}}.
  enum Kind { Const, Spot, Neg, Add, Sub, Mul, Div, Max, Min, Average, RunningMax, RunningMin, AverageAt } kind;
  double value = 0.0;    // Const
  std::string name;      // Spot
  std::unique_ptr<Obs> a, b;
  bool has_from = false, has_to = false;  // Average window
  double from = 0.0, to = 0.0;
  std::vector<double> dates;  // AverageAt fixings
};

struct Cond {
  enum Kind { Bool, At, Ge, Gt, Le, Lt, And, Or, Not } kind;
  bool flag = false;  // Bool
  double t = 0.0;     // At
  std::unique_ptr<Obs> lhs, rhs;
  std::unique_ptr<Cond> a, b;
};

struct Con {
  enum Kind { Zero, One, Give, And, Or, Scale, When, Anytime, Until } kind;
  std::string ccy;
  std::unique_ptr<Obs> obs;
  std::unique_ptr<Cond> cond;
  std::unique_ptr<Con> a, b;
};

// Contract::from_json (src/ast.rs:210-218): serde's externally tagged form.
std::unique_ptr<Con> parse_contract(const char* json);

// ---- Plan / grid (src/compiler.rs:25-189) -------------------------------
struct Plan {
  std::vector<std::string> assets;  // sorted, unique
  std::vector<double> fixed_dates;  // sorted, deduplicated at 1e-12
  double horizon = 0.0;
  bool needs_fine_grid = false;
};
Plan compile_plan(const Con& c);
std::vector<double> time_grid_forced(const Plan& p, uint64_t steps_per_year, bool force_fine);
std::vector<double> sort_dedup(std::vector<double> v);
bool condition_is_price_dependent(const Cond& c);
void validate_grid(const std::vector<double>& grid);

// ---- Lowering ------------------------------------------------------------
struct Program {
  std::vector<KOp> init;       // run once before grid point 0
  std::vector<KOp> ops;        // concatenated per-step lists
  std::vector<KStep> steps;    // one per grid point
  std::vector<KGroup> groups;  // one per 4 grid points
  int n_slots = 2;             // slots used (<= kMaxSlots)
  int n_outputs = 1;           // pv accumulators used
  bool time_only = true;       // no price-dependent condition anywhere
  int max_ops_per_step = 0;
  // TimeOutOfRange's payload (KontractError::TimeOutOfRange(f64)): the first
  // AverageAt fixing date found off the grid, NaN if none
  double time_payload = __builtin_nan("");
  std::string describe() const;
};

// Lower contracts (one pv output each, in order) for simulation on `grid`
// with the model's assets `model_assets` (program asset index == model asset
// index).  Raises Error on contracts the engine cannot run.
Program lower(const std::vector<const Con*>& contracts, const std::vector<double>& grid,
              const std::vector<std::string>& model_assets);

// LSM: the exercised payoff evaluated at every grid point k >= 1 (flows at
// the acquisition step only; raises Unsupported otherwise).
Program lower_exercise(const Con& contract, const std::vector<double>& grid,
                       const std::vector<std::string>& model_assets);

// Classify groups of 4 grid points (fills prog.groups).
void classify_groups(Program& prog);

// Pathwise-Greek plan of a program: every flow on the last grid point and
// paid unconditionally, only arithmetic ops, and each running sum / extreme
// fed directly by one asset's spot and read only on the last point.  Raises
// Unsupported (with the reason) otherwise: such contracts need bump Greeks.
struct AdjPlan {
  int n_acc = 0;
  int slot[4] = {0, 0, 0, 0};
  int kind[4] = {0, 0, 0, 0};
  int asset[4] = {0, 0, 0, 0};
  std::vector<uint32_t> feed;  // per grid point: bit q set if accumulator q is fed there
};
AdjPlan classify_adjoint(const Program& prog);

// For LSM: true iff every flow of the contract happens at its acquisition
// step (no When/Anytime/Until/Or anywhere in the tree).
bool flows_at_acquisition(const Con& c);

}  // namespace kt
