// kt_compiler.cpp — contract JSON reader, plan/grid, and lowering to the
// per-step op program interpreted by the path kernels.
#include "kt_compiler.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <sstream>
#include <unordered_map>

#include "../../include/kontracts.h"

namespace kt {

void raise(int code, const std::string& msg) { throw Error{code, msg}; }

// ===========================================================================
// JSON reader for serde's externally tagged enums (src/ast.rs:34-218).
// Builds the AST directly while reading.
// ===========================================================================
namespace {

class Reader {
 public:
  explicit Reader(const char* s) : p_(s), e_(s + strlen(s)) {}

  std::unique_ptr<Con> contract() {
    ws();
    auto c = std::make_unique<Con>();
    if (peek('"')) {
      std::string unit = str();
      if (unit == "Zero") {
        c->kind = Con::Zero;
        return c;
      }
      bad("unknown variant `" + unit + "`, expected a Contract");
    }
    std::string tag = open_tag();
    if (tag == "One") {
      c->kind = Con::One;
      c->ccy = str();
    } else if (tag == "Give") {
      c->kind = Con::Give;
      c->a = contract();
    } else if (tag == "And" || tag == "Or") {
      c->kind = tag == "And" ? Con::And : Con::Or;
      expect('[');
      c->a = contract();
      expect(',');
      c->b = contract();
      expect(']');
    } else if (tag == "Scale") {
      c->kind = Con::Scale;
      expect('[');
      c->obs = observable();
      expect(',');
      c->a = contract();
      expect(']');
    } else if (tag == "When" || tag == "Anytime" || tag == "Until") {
      c->kind = tag == "When" ? Con::When : tag == "Anytime" ? Con::Anytime : Con::Until;
      expect('[');
      c->cond = condition();
      expect(',');
      c->a = contract();
      expect(']');
    } else {
      bad("unknown variant `" + tag + "`, expected a Contract");
    }
    expect('}');
    return c;
  }

  std::unique_ptr<Obs> observable() {
    auto o = std::make_unique<Obs>();
    std::string tag = open_tag();
    auto pair = [&](Obs::Kind k) {
      o->kind = k;
      expect('[');
      o->a = observable();
      expect(',');
      o->b = observable();
      expect(']');
    };
    if (tag == "Const") {
      o->kind = Obs::Const;
      o->value = number();
    } else if (tag == "Spot") {
      o->kind = Obs::Spot;
      o->name = str();
    } else if (tag == "Neg") {
      o->kind = Obs::Neg;
      o->a = observable();
    } else if (tag == "Add") pair(Obs::Add);
    else if (tag == "Sub") pair(Obs::Sub);
    else if (tag == "Mul") pair(Obs::Mul);
    else if (tag == "Div") pair(Obs::Div);
    else if (tag == "Max") pair(Obs::Max);
    else if (tag == "Min") pair(Obs::Min);
    else if (tag == "RunningMax" || tag == "RunningMin") {
      o->kind = tag == "RunningMax" ? Obs::RunningMax : Obs::RunningMin;
      o->a = observable();
    } else if (tag == "Average") {
      o->kind = Obs::Average;
      expect('{');
      bool first = true;
      while (!peek('}')) {
        if (!first) expect(',');
        first = false;
        ws();
        std::string key = str();
        expect(':');
        if (key == "obs") {
          o->a = observable();
        } else if (key == "from_year" || key == "to_year") {
          ws();
          bool is_null = literal("null");
          double v = is_null ? 0.0 : number();
          if (key == "from_year") {
            o->has_from = !is_null;
            o->from = v;
          } else {
            o->has_to = !is_null;
            o->to = v;
          }
        } else {
          skip_value();  // serde ignores unknown fields by default
        }
      }
      expect('}');
      if (!o->a) bad("missing field `obs`");
    } else if (tag == "AverageAt") {
      o->kind = Obs::AverageAt;
      expect('{');
      bool first = true, seen_dates = false;
      while (!peek('}')) {
        if (!first) expect(',');
        first = false;
        ws();
        std::string key = str();
        expect(':');
        if (key == "obs") {
          o->a = observable();
        } else if (key == "dates") {
          seen_dates = true;
          expect('[');
          bool f2 = true;
          while (!peek(']')) {
            if (!f2) expect(',');
            f2 = false;
            ws();
            o->dates.push_back(number());
          }
          expect(']');
        } else {
          skip_value();
        }
      }
      expect('}');
      if (!o->a) bad("missing field `obs`");
      if (!seen_dates) bad("missing field `dates`");
    } else {
      bad("unknown variant `" + tag + "`, expected an Observable");
    }
    expect('}');
    return o;
  }

  std::unique_ptr<Cond> condition() {
    auto c = std::make_unique<Cond>();
    std::string tag = open_tag();
    auto cmp = [&](Cond::Kind k) {
      c->kind = k;
      expect('[');
      c->lhs = observable();
      expect(',');
      c->rhs = observable();
      expect(']');
    };
    if (tag == "Bool") {
      c->kind = Cond::Bool;
      ws();
      if (literal("true")) c->flag = true;
      else if (literal("false")) c->flag = false;
      else bad("expected a boolean");
    } else if (tag == "At") {
      c->kind = Cond::At;
      c->t = number();
    } else if (tag == "Ge") cmp(Cond::Ge);
    else if (tag == "Gt") cmp(Cond::Gt);
    else if (tag == "Le") cmp(Cond::Le);
    else if (tag == "Lt") cmp(Cond::Lt);
    else if (tag == "And" || tag == "Or") {
      c->kind = tag == "And" ? Cond::And : Cond::Or;
      expect('[');
      c->a = condition();
      expect(',');
      c->b = condition();
      expect(']');
    } else if (tag == "Not") {
      c->kind = Cond::Not;
      c->a = condition();
    } else {
      bad("unknown variant `" + tag + "`, expected a Condition");
    }
    expect('}');
    return c;
  }

  void finish() {
    ws();
    if (p_ != e_) bad("trailing characters");
  }

 private:
  const char* p_;
  const char* e_;

  [[noreturn]] void bad(const std::string& m) { raise(KT_ERR_SERDE, m); }
  void ws() {
    while (p_ < e_ && (*p_ == ' ' || *p_ == '\n' || *p_ == '\r' || *p_ == '\t')) ++p_;
  }
  bool peek(char c) {
    ws();
    return p_ < e_ && *p_ == c;
  }
  void expect(char c) {
    ws();
    if (p_ >= e_ || *p_ != c) raise(KT_ERR_SERDE, std::string("expected `") + c + "`");
    ++p_;
  }
  bool literal(const char* lit) {
    size_t n = strlen(lit);
    if ((size_t)(e_ - p_) >= n && strncmp(p_, lit, n) == 0) {
      p_ += n;
      return true;
    }
    return false;
  }
  std::string open_tag() {
    expect('{');
    ws();
    std::string tag = str();
    expect(':');
    return tag;
  }
  std::string str() {
    ws();
    if (p_ >= e_ || *p_ != '"') raise(KT_ERR_SERDE, "expected a string");
    ++p_;
    std::string out;
    while (p_ < e_ && *p_ != '"') {
      char ch = *p_++;
      if (ch != '\\') {
        out += ch;
        continue;
      }
      if (p_ >= e_) break;
      char esc = *p_++;
      switch (esc) {
        case '"': out += '"'; break;
        case '\\': out += '\\'; break;
        case '/': out += '/'; break;
        case 'b': out += '\b'; break;
        case 'f': out += '\f'; break;
        case 'n': out += '\n'; break;
        case 'r': out += '\r'; break;
        case 't': out += '\t'; break;
        case 'u': {
          if (e_ - p_ < 4) raise(KT_ERR_SERDE, "bad unicode escape");
          unsigned cp = 0;
          for (int i = 0; i < 4; ++i) {
            char h = *p_++;
            cp <<= 4;
            if (h >= '0' && h <= '9') cp |= (unsigned)(h - '0');
            else if (h >= 'a' && h <= 'f') cp |= (unsigned)(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') cp |= (unsigned)(h - 'A' + 10);
            else raise(KT_ERR_SERDE, "bad unicode escape");
          }
          if (cp < 0x80) {
            out += (char)cp;
          } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
          } else {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
          }
          break;
        }
        default: raise(KT_ERR_SERDE, "bad escape");
      }
    }
    if (p_ >= e_) raise(KT_ERR_SERDE, "unterminated string");
    ++p_;
    return out;
  }
  double number() {
    ws();
    char* end = nullptr;
    double v = strtod(p_, &end);
    if (end == p_) raise(KT_ERR_SERDE, "expected a number");
    p_ = end;
    return v;
  }
  void skip_value() {
    ws();
    if (p_ >= e_) raise(KT_ERR_SERDE, "unexpected end");
    if (*p_ == '"') {
      str();
    } else if (*p_ == '{' || *p_ == '[') {
      char open = *p_, close = open == '{' ? '}' : ']';
      ++p_;
      bool first = true;
      while (!peek(close)) {
        if (!first) expect(',');
        first = false;
        if (open == '{') {
          str();
          expect(':');
        }
        skip_value();
      }
      expect(close);
    } else if (!literal("true") && !literal("false") && !literal("null")) {
      number();
    }
  }
};

}  // namespace

std::unique_ptr<Con> parse_contract(const char* json) {
  if (!json) raise(KT_ERR_INVALID_ARGUMENT, "null contract json");
  Reader r(json);
  auto c = r.contract();
  r.finish();
  return c;
}

// ===========================================================================
// Plan and grid (src/compiler.rs:25-189)
// ===========================================================================
std::vector<double> sort_dedup(std::vector<double> v) {
  std::stable_sort(v.begin(), v.end(), [](double a, double b) { return a < b; });
  std::vector<double> out;
  out.reserve(v.size());
  for (double x : v)
    if (out.empty() || !(std::fabs(out.back() - x) < 1e-12)) out.push_back(x);
  return out;
}

bool condition_is_price_dependent(const Cond& c) {
  switch (c.kind) {
    case Cond::Bool:
    case Cond::At: return false;
    case Cond::And:
    case Cond::Or: return condition_is_price_dependent(*c.a) || condition_is_price_dependent(*c.b);
    case Cond::Not: return condition_is_price_dependent(*c.a);
    default: return true;
  }
}

namespace {
struct PlanAcc {
  std::vector<std::string> assets;
  std::vector<double> dates;
  bool fine = false;
};
void plan_obs(const Obs& o, PlanAcc& acc) {
  switch (o.kind) {
    case Obs::Const: return;
    case Obs::Spot:
      if (std::find(acc.assets.begin(), acc.assets.end(), o.name) == acc.assets.end()) acc.assets.push_back(o.name);
      return;
    case Obs::Neg: plan_obs(*o.a, acc); return;
    case Obs::Average:
    case Obs::RunningMax:
    case Obs::RunningMin:
      acc.fine = true;
      plan_obs(*o.a, acc);
      return;
    case Obs::AverageAt:  // the fixings join the grid as fixed dates
      acc.dates.insert(acc.dates.end(), o.dates.begin(), o.dates.end());
      plan_obs(*o.a, acc);
      return;
    default:
      plan_obs(*o.a, acc);
      plan_obs(*o.b, acc);
  }
}
void plan_cond(const Cond& c, PlanAcc& acc) {
  switch (c.kind) {
    case Cond::Bool: return;
    case Cond::At:
      if (!std::isfinite(c.t) || c.t < 0.0) {
        char buf[96];
        snprintf(buf, sizeof buf, "invalid acquisition date: %.17g", c.t);
        raise(KT_ERR_MALFORMED_CONTRACT, buf);
      }
      acc.dates.push_back(c.t);
      return;
    case Cond::And:
    case Cond::Or:
      plan_cond(*c.a, acc);
      plan_cond(*c.b, acc);
      return;
    case Cond::Not: plan_cond(*c.a, acc); return;
    default:
      plan_obs(*c.lhs, acc);
      plan_obs(*c.rhs, acc);
  }
}
void plan_con(const Con& c, PlanAcc& acc) {
  switch (c.kind) {
    case Con::Zero:
    case Con::One: return;
    case Con::Give: plan_con(*c.a, acc); return;
    case Con::And:
    case Con::Or:
      plan_con(*c.a, acc);
      plan_con(*c.b, acc);
      return;
    case Con::Scale:
      plan_obs(*c.obs, acc);
      plan_con(*c.a, acc);
      return;
    default:
      plan_cond(*c.cond, acc);
      if (condition_is_price_dependent(*c.cond)) acc.fine = true;
      plan_con(*c.a, acc);
  }
}
}  // namespace

Plan compile_plan(const Con& c) {
  PlanAcc acc;
  plan_con(c, acc);
  Plan p;
  std::sort(acc.assets.begin(), acc.assets.end());
  p.assets = acc.assets;
  double h = 0.0;
  for (double d : acc.dates)
    if (d > h) h = d;
  p.horizon = h;
  p.fixed_dates = sort_dedup(acc.dates);
  p.needs_fine_grid = acc.fine;
  return p;
}

std::vector<double> time_grid_forced(const Plan& p, uint64_t spy, bool force_fine) {
  std::vector<double> pts;
  pts.push_back(0.0);
  pts.insert(pts.end(), p.fixed_dates.begin(), p.fixed_dates.end());
  pts.push_back(p.horizon);
  if ((p.needs_fine_grid || force_fine) && p.horizon > 0.0 && spy > 0) {
    uint64_t n = (uint64_t)std::ceil(p.horizon * (double)spy);
    for (uint64_t k = 1; k <= n; ++k) pts.push_back(p.horizon * (double)k / (double)n);
  }
  return sort_dedup(pts);
}

void validate_grid(const std::vector<double>& g) {
  if (g.empty()) raise(KT_ERR_INCONSISTENT_PATH, "empty time grid");
  double prev = 0.0;
  for (double t : g) {
    if (t < prev) {
      char buf[96];
      snprintf(buf, sizeof buf, "time grid not increasing near %.17g", t);
      raise(KT_ERR_INCONSISTENT_PATH, buf);
    }
    prev = t;
  }
}

bool flows_at_acquisition(const Con& c) {
  switch (c.kind) {
    case Con::Zero:
    case Con::One: return true;
    case Con::Give:
    case Con::Scale: return flows_at_acquisition(*c.a);
    case Con::And: return flows_at_acquisition(*c.a) && flows_at_acquisition(*c.b);
    default: return false;
  }
}

// ===========================================================================
// Lowering
//
// The reference evaluates a contract per path recursively
// (cashflows / first_activation, src/pricer.rs:374-433), with random access to
// the whole path.  The engine instead streams each path once, grid point by
// grid point, and keeps per path only:
//   * an acquisition flag (and "acquired at this step" flag) for every
//     when/anytime whose activation depends on the price path,
//   * a knock-out flag for every until,
//   * accumulators for RunningMax / RunningMin / Average,
//   * one present value per output.
// Acquisitions that depend on time only (at(t) under a deterministic parent)
// are resolved here, at compile time, to a grid index.  The op list of grid
// point k evaluates, in pre-order of the contract tree, the conditions whose
// node may be scanning at k and the flows that may be paid at k; the results
// are identical to the recursive definition:
//   when(c, x)   acquires x at the first step >= its own acquisition where c holds,
//   until(c, x)  keeps the flows of x strictly before the first such step,
//   scale(o, x)  multiplies each flow of x by o evaluated at the flow's step.
// ===========================================================================
namespace {

struct Acq {
  enum K { Static, Dynamic, Never } k = Never;
  int step = 0;      // Static: acquisition step
  int earliest = 0;  // Dynamic: earliest possible acquisition step
  int slot = -1;     // Dynamic: acquired flag; slot+1 = acquired-this-step flag
};

struct KoRef {
  int slot;
  int start;  // first step at which the until may be scanning
};

struct Event {
  enum K { When, Until, Emit, OrErr } k;
  int lo = 0, hi = -1;
  const Cond* cond = nullptr;
  int slot = -1;         // When: acq slot of the child; Until: ko slot
  int parent_slot = -1;  // acquisition slot of the parent; -1 = acquired (ONE)
  // Emit
  std::vector<const Obs*> scales;  // outermost first
  double sign = 1.0;
  std::vector<KoRef> kos;
  int out = 0;
};

struct Temporal {
  const Obs* node;
  int depth;          // number of temporal ancestors
  int slot = -1;      // RunningMax/Min: acc; Average: sum, sum+1 = compensation
  int last_read = -1; // last step the node is read at (-1: never)
  int static_err = 0; // error raised whenever the node is read
};

class Lowerer {
 public:
  Lowerer(const std::vector<double>& grid, const std::vector<std::string>& assets)
      : grid_(grid), assets_(assets), n_((int)grid.size()) {
    next_slot_ = 2;
  }

  void add_contract(const Con& c, int out) {
    collect_temporals(c, 0);
    Acq root;
    root.k = Acq::Static;
    root.step = 0;
    Event ctx;
    ctx.k = Event::Emit;
    ctx.out = out;
    visit(c, root, ctx);
    if (c.kind != Con::Zero && contains_price_condition(c)) time_only_ = false;
  }

  // LSM exercise payoff: the contract acquired at every grid point k >= 1
  // (value_at, src/lsm.rs:238-251).  With every flow at its acquisition step,
  // each point's list sums the flows paid there; the caller runs it with unit
  // discount factors and reads the per-point value.
  void add_exercise_contract(const Con& c) {
    collect_temporals(c, 0);
    for (int k = 1; k < n_; ++k) {
      Acq root;
      root.k = Acq::Static;
      root.step = k;
      Event ctx;
      ctx.k = Event::Emit;
      ctx.out = 0;
      visit(c, root, ctx);
    }
  }

  Program finish() {
    // Spot slots for every model asset referenced anywhere.
    for (size_t a = 0; a < assets_.size(); ++a) spot_slot_.push_back(-1);
    temp_base_ = next_slot_ + (int)assets_.size();
    int spot_next = next_slot_;
    for (size_t a = 0; a < assets_.size(); ++a) spot_slot_[a] = spot_next++;

    // Event op lists per step.
    std::vector<std::vector<KOp>> ev_ops(n_);
    std::vector<uint32_t> spot_used(n_, 0);
    for (int k = 0; k < n_; ++k) {
      for (const Event& e : events_) {
        if (e.hi < e.lo || k < e.lo || k > e.hi) continue;
        Group g{ev_ops[k], k, temp_base_, {}, spot_used[k]};
        compile_event(e, g);
      }
    }
    // Temporal updates, outermost first (they register reads of the inner ones).
    std::vector<size_t> order(temporals_.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t x, size_t y) { return temporals_[x].depth < temporals_[y].depth; });
    std::vector<std::vector<std::vector<KOp>>> upd(temporals_.size(), std::vector<std::vector<KOp>>(n_));
    for (size_t oi : order) {
      Temporal& T = temporals_[oi];
      if (T.last_read < 0) continue;
      for (int k = 0; k <= T.last_read && k < n_; ++k) {
        if (T.node->kind == Obs::Average && !in_window(*T.node, k)) continue;
        if (T.node->kind == Obs::AverageAt && !is_fixing(*T.node, k)) continue;
        Group g{upd[oi][k], k, temp_base_, {}, spot_used[k]};
        int v = comp_obs(*T.node->a, g);
        for (int code : g.pending)
          if (!T.static_err) T.static_err = code;
        KOp op{};
        op.code = T.node->kind == Obs::RunningMax ? OP_RMAX : T.node->kind == Obs::RunningMin ? OP_RMIN : OP_AVG;
        op.d = (uint8_t)T.slot;
        op.a = (uint8_t)v;
        g.out.push_back(op);
      }
    }
    // Assemble: spots, temporal updates (innermost first), events.
    Program prog;
    prog.steps.resize(n_);
    std::map<std::vector<uint64_t>, int> dedup;
    for (int k = 0; k < n_; ++k) {
      std::vector<KOp> list;
      for (size_t a = 0; a < assets_.size(); ++a)
        if (spot_used[k] & (1u << a)) list.push_back(KOp{OP_SPOT, (uint8_t)spot_slot_[a], (uint8_t)a, 0, 0.0f});
      for (auto it = order.rbegin(); it != order.rend(); ++it)
        list.insert(list.end(), upd[*it][k].begin(), upd[*it][k].end());
      list.insert(list.end(), ev_ops[k].begin(), ev_ops[k].end());
      prog.max_ops_per_step = std::max(prog.max_ops_per_step, (int)list.size());
      if (list.empty()) {
        prog.steps[k] = KStep{0, 0};
        continue;
      }
      std::vector<uint64_t> key(list.size());
      for (size_t i = 0; i < list.size(); ++i) memcpy(&key[i], &list[i], 8);
      auto f = dedup.find(key);
      if (f != dedup.end()) {
        prog.steps[k] = KStep{f->second, (int)list.size()};
      } else {
        int begin = (int)prog.ops.size();
        prog.ops.insert(prog.ops.end(), list.begin(), list.end());
        dedup[key] = begin;
        prog.steps[k] = KStep{begin, (int)list.size()};
      }
    }
    prog.init.push_back(KOp{OP_CONST, (uint8_t)kSlotOne, 0, 0, 1.0f});
    for (Temporal& T : temporals_) {
      if (T.last_read < 0) continue;
      if (T.node->kind == Obs::RunningMax) prog.init.push_back(KOp{OP_CONST, (uint8_t)T.slot, 0, 0, -INFINITY});
      if (T.node->kind == Obs::RunningMin) prog.init.push_back(KOp{OP_CONST, (uint8_t)T.slot, 0, 0, INFINITY});
    }
    prog.n_slots = std::max(temp_base_, max_temp_);
    if (prog.n_slots > kMaxSlots)
      raise(KT_ERR_UNSUPPORTED, "contract needs " + std::to_string(prog.n_slots) +
                                    " registers; the path kernel holds " + std::to_string(kMaxSlots));
    prog.time_only = time_only_;
    prog.time_payload = time_payload_;
    return prog;
  }

  int n_outputs_hint = 1;

 private:
  const std::vector<double>& grid_;
  const std::vector<std::string>& assets_;
  int n_;
  int next_slot_;
  int temp_base_ = 0;
  int max_temp_ = 0;
  bool time_only_ = true;
  double time_payload_ = __builtin_nan("");
  std::vector<Event> events_;
  std::vector<Temporal> temporals_;
  std::unordered_map<const Obs*, int> temporal_index_;
  std::vector<int> spot_slot_;

  struct Group {
    std::vector<KOp>& out;
    int k;
    int temp;
    std::vector<int> pending;  // error codes raised by the expressions of this group
    uint32_t& spot_used;
  };

  int alloc_persistent(int n) {
    int s = next_slot_;
    next_slot_ += n;
    if (next_slot_ > 250) raise(KT_ERR_UNSUPPORTED, "contract too large for the path kernel");
    return s;
  }

  static bool contains_price_condition(const Con& c) {
    switch (c.kind) {
      case Con::Zero:
      case Con::One: return false;
      case Con::Give:
      case Con::Scale: return contains_price_condition(*c.a);
      case Con::And:
      case Con::Or: return contains_price_condition(*c.a) || contains_price_condition(*c.b);
      default: return condition_is_price_dependent(*c.cond) || contains_price_condition(*c.a);
    }
  }

  // --- temporal nodes -------------------------------------------------------
  void collect_obs(const Obs& o, int depth) {
    bool temporal = o.kind == Obs::Average || o.kind == Obs::RunningMax || o.kind == Obs::RunningMin ||
                    o.kind == Obs::AverageAt;
    if (temporal && !temporal_index_.count(&o)) {
      Temporal T{&o, depth};
      T.slot = alloc_persistent(o.kind == Obs::Average || o.kind == Obs::AverageAt ? 2 : 1);
      temporal_index_[&o] = (int)temporals_.size();
      temporals_.push_back(T);
    }
    if (o.a) collect_obs(*o.a, depth + (temporal ? 1 : 0));
    if (o.b) collect_obs(*o.b, depth + (temporal ? 1 : 0));
  }
  void collect_cond(const Cond& c) {
    if (c.lhs) collect_obs(*c.lhs, 0);
    if (c.rhs) collect_obs(*c.rhs, 0);
    if (c.a) collect_cond(*c.a);
    if (c.b) collect_cond(*c.b);
  }
  void collect_temporals(const Con& c, int) {
    if (c.obs) collect_obs(*c.obs, 0);
    if (c.cond) collect_cond(*c.cond);
    if (c.a) collect_temporals(*c.a, 0);
    if (c.b) collect_temporals(*c.b, 0);
  }

  // AverageAt: grid point i is one of the fixing dates.
  bool is_fixing(const Obs& o, int i) const {
    for (double d : o.dates)
      if (std::fabs(grid_[i] - d) <= 1e-12) return true;
    return false;
  }

  bool in_window(const Obs& o, int i) const {
    double lo = (o.has_from ? o.from : 0.0) - 1e-12;
    double ti = grid_[i];
    if (ti < lo) return false;
    if (o.has_to && ti > o.to + 1e-12) return false;
    return true;
  }

  // --- static (time-only) conditions ---------------------------------------
  bool static_true(const Cond& c, int k) const {
    switch (c.kind) {
      case Cond::Bool: return c.flag;
      case Cond::At: return grid_[k] + 1e-12 >= c.t;
      case Cond::And: return static_true(*c.a, k) && static_true(*c.b, k);
      case Cond::Or: return static_true(*c.a, k) || static_true(*c.b, k);
      case Cond::Not: return !static_true(*c.a, k);
      default: return false;
    }
  }
  int first_static(const Cond& c, int from) const {
    for (int k = from; k < n_; ++k)
      if (static_true(c, k)) return k;
    return -1;
  }

  // --- pass A: acquisitions and events, in pre-order ------------------------
  int visit(const Con& c, const Acq& acq, Event ctx) {
    if (acq.k == Acq::Never) return -1;
    int start = acq.k == Acq::Static ? acq.step : acq.earliest;
    if (start >= n_) return -1;
    switch (c.kind) {
      case Con::Zero: return -1;
      case Con::One: {
        Event e = ctx;
        e.k = Event::Emit;
        e.parent_slot = acq.k == Acq::Dynamic ? acq.slot + 1 : -1;
        e.lo = start;
        e.hi = acq.k == Acq::Static ? start : n_ - 1;
        events_.push_back(e);
        return e.hi;
      }
      case Con::Give: {
        ctx.sign = -ctx.sign;
        return visit(*c.a, acq, ctx);
      }
      case Con::And: {
        int l1 = visit(*c.a, acq, ctx);
        int l2 = visit(*c.b, acq, ctx);
        return std::max(l1, l2);
      }
      case Con::Scale: {
        ctx.scales.push_back(c.obs.get());
        return visit(*c.a, acq, ctx);
      }
      case Con::Or: {
        if (acq.k == Acq::Static) raise(KT_ERR_UNSUPPORTED, "or (jalon J17, LSM)");
        Event e;
        e.k = Event::OrErr;
        e.parent_slot = acq.slot + 1;
        e.lo = start;
        e.hi = n_ - 1;
        events_.push_back(e);
        return e.hi;
      }
      case Con::When:
      case Con::Anytime: {
        const Cond& cond = *c.cond;
        bool dynamic_cond = condition_is_price_dependent(cond);
        if (!dynamic_cond && acq.k == Acq::Static) {
          Acq child;
          int k = first_static(cond, start);
          if (k >= 0) {
            child.k = Acq::Static;
            child.step = k;
          }
          return visit(*c.a, child, ctx);
        }
        int earliest = start;
        if (!dynamic_cond) {
          earliest = first_static(cond, start);
          if (earliest < 0) return -1;
        }
        size_t idx = events_.size();
        Event e;
        e.k = Event::When;
        e.cond = &cond;
        e.slot = alloc_persistent(2);
        e.parent_slot = acq.k == Acq::Dynamic ? acq.slot : -1;
        e.lo = earliest;
        events_.push_back(e);
        Acq child;
        child.k = Acq::Dynamic;
        child.earliest = earliest;
        child.slot = e.slot;
        int last = visit(*c.a, child, ctx);
        events_[idx].hi = last;  // last < lo disables the event
        return last;
      }
      case Con::Until: {
        size_t idx = events_.size();
        Event e;
        e.k = Event::Until;
        e.cond = c.cond.get();
        e.slot = alloc_persistent(1);
        e.parent_slot = acq.k == Acq::Dynamic ? acq.slot : -1;
        e.lo = start;
        events_.push_back(e);
        ctx.kos.push_back(KoRef{e.slot, start});
        int last = visit(*c.a, acq, ctx);
        events_[idx].hi = last;
        return last;
      }
    }
    return -1;
  }

  // --- expression compilation into one step's op list -----------------------
  int temp(Group& g) {
    int s = g.temp++;
    if (g.temp > max_temp_) max_temp_ = g.temp;
    return s;
  }
  void op(Group& g, uint8_t code, int d, int a, int b, float imm = 0.0f) {
    g.out.push_back(KOp{code, (uint8_t)d, (uint8_t)a, (uint8_t)b, imm});
  }

  static bool is_const(const Obs& o) { return o.kind == Obs::Const; }

  int comp_obs(const Obs& o, Group& g) {
    switch (o.kind) {
      case Obs::Const: {
        int d = temp(g);
        op(g, OP_CONST, d, 0, 0, (float)o.value);
        return d;
      }
      case Obs::Spot: {
        for (size_t a = 0; a < assets_.size(); ++a)
          if (assets_[a] == o.name) {
            g.spot_used |= 1u << a;
            return spot_slot_[a];
          }
        g.pending.push_back(KT_ERR_UNKNOWN_ASSET);
        return kSlotZero;
      }
      case Obs::Neg: {
        int a = comp_obs(*o.a, g);
        int d = temp(g);
        op(g, OP_NEG, d, a, 0);
        return d;
      }
      case Obs::Add:
      case Obs::Sub:
      case Obs::Mul:
      case Obs::Div:
      case Obs::Max:
      case Obs::Min: return comp_binary(o, g);
      case Obs::RunningMax:
      case Obs::RunningMin:
      case Obs::Average:
      case Obs::AverageAt: return comp_temporal(o, g);
    }
    return kSlotZero;
  }

  int comp_binary(const Obs& o, Group& g) {
    static const uint8_t rr[] = {0, 0, 0, OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_MAX, OP_MIN};
    static const uint8_t ri[] = {0, 0, 0, OP_ADDI, OP_SUBI, OP_MULI, OP_DIVI, OP_MAXI, OP_MINI};
    // imm op r[a] forms: commutative ops reuse ri, else the reversed opcode.
    static const uint8_t ir[] = {0, 0, 0, OP_ADDI, OP_RSUBI, OP_MULI, OP_RDIVI, OP_MAXI, OP_MINI};
    int kind = (int)o.kind;
    if (is_const(*o.b) && !is_const(*o.a)) {
      int a = comp_obs(*o.a, g);
      int d = temp(g);
      op(g, ri[kind], d, a, 0, (float)o.b->value);
      return d;
    }
    if (is_const(*o.a) && !is_const(*o.b)) {
      int b = comp_obs(*o.b, g);
      int d = temp(g);
      op(g, ir[kind], d, b, 0, (float)o.a->value);
      return d;
    }
    int a = comp_obs(*o.a, g);
    int b = comp_obs(*o.b, g);
    int d = temp(g);
    op(g, rr[kind], d, a, b);
    return d;
  }

  int comp_temporal(const Obs& o, Group& g) {
    Temporal& T = temporals_[temporal_index_.at(&o)];
    if (g.k > T.last_read) T.last_read = g.k;
    if (T.static_err) g.pending.push_back(T.static_err);
    // Errors of the inner expression are only known once its update ops are
    // compiled; check the inner expression for unknown assets right away.
    if (has_unknown_asset(*o.a)) g.pending.push_back(KT_ERR_UNKNOWN_ASSET);
    if (o.kind == Obs::AverageAt) {
      // mean over the fixings on or before now; each must be a grid point
      if (o.dates.empty()) {
        g.pending.push_back(KT_ERR_MALFORMED_CONTRACT);
        return kSlotZero;
      }
      for (size_t i = 1; i < o.dates.size(); ++i)
        if (!(o.dates[i] > o.dates[i - 1])) {
          g.pending.push_back(KT_ERR_MALFORMED_CONTRACT);
          return kSlotZero;
        }
      const double t_now = grid_[g.k];
      int count = 0;
      for (double d : o.dates) {
        if (d > t_now + 1e-12) break;
        bool on_grid = false;
        for (int i = 0; i < n_ && !on_grid; ++i) on_grid = std::fabs(grid_[i] - d) <= 1e-12;
        if (!on_grid) {
          g.pending.push_back(KT_ERR_TIME_OUT_OF_RANGE);
          if (std::isnan(time_payload_)) time_payload_ = d;
          return kSlotZero;
        }
        ++count;
      }
      if (count == 0) {
        g.pending.push_back(KT_ERR_MALFORMED_CONTRACT);
        return kSlotZero;
      }
      int d = temp(g);
      op(g, OP_SUB, d, T.slot, T.slot + 1);
      op(g, OP_MULI, d, d, 0, (float)(1.0 / (double)count));
      return d;
    }
    if (o.kind != Obs::Average) return T.slot;
    // Average over the window (src/observable.rs:103-133).
    double t_now = grid_[g.k];
    double t_from = o.has_from ? o.from : 0.0;
    double t_to = o.has_to ? o.to : t_now;
    if (t_from > t_to + 1e-12) {
      g.pending.push_back(KT_ERR_MALFORMED_CONTRACT);
      return kSlotZero;
    }
    int count = 0, last = -1;
    for (int i = 0; i < n_; ++i) {
      double ti = grid_[i];
      if (ti >= t_from - 1e-12 && ti <= t_to + 1e-12) {
        ++count;
        last = i;
      }
    }
    if (count == 0) {
      g.pending.push_back(KT_ERR_MALFORMED_CONTRACT);
      return kSlotZero;
    }
    if (last > g.k) {
      // The window reaches grid points after the evaluation step: a streaming
      // pass cannot see them.
      g.pending.push_back(KT_ERR_UNSUPPORTED);
      return kSlotZero;
    }
    int d = temp(g);
    op(g, OP_SUB, d, T.slot, T.slot + 1);
    op(g, OP_MULI, d, d, 0, (float)(1.0 / (double)count));
    return d;
  }

  bool has_unknown_asset(const Obs& o) const {
    if (o.kind == Obs::Spot) return std::find(assets_.begin(), assets_.end(), o.name) == assets_.end();
    return (o.a && has_unknown_asset(*o.a)) || (o.b && has_unknown_asset(*o.b));
  }

  int comp_cond(const Cond& c, Group& g) {
    switch (c.kind) {
      case Cond::Bool: return c.flag ? kSlotOne : kSlotZero;
      case Cond::At: return static_true(c, g.k) ? kSlotOne : kSlotZero;
      case Cond::Ge:
      case Cond::Gt:
      case Cond::Le:
      case Cond::Lt: {
        static const uint8_t rr[] = {0, 0, OP_GE, OP_GT, OP_LE, OP_LT};
        static const uint8_t ri[] = {0, 0, OP_GEI, OP_GTI, OP_LEI, OP_LTI};
        static const uint8_t mirrored[] = {0, 0, OP_LEI, OP_LTI, OP_GEI, OP_GTI};
        int kind = (int)c.kind;
        if (is_const(*c.rhs) && !is_const(*c.lhs)) {
          int a = comp_obs(*c.lhs, g);
          int d = temp(g);
          op(g, ri[kind], d, a, 0, (float)c.rhs->value);
          return d;
        }
        if (is_const(*c.lhs) && !is_const(*c.rhs)) {
          int b = comp_obs(*c.rhs, g);
          int d = temp(g);
          op(g, mirrored[kind], d, b, 0, (float)c.lhs->value);
          return d;
        }
        int a = comp_obs(*c.lhs, g);
        int b = comp_obs(*c.rhs, g);
        int d = temp(g);
        op(g, rr[kind], d, a, b);
        return d;
      }
      case Cond::And:
      case Cond::Or: {
        int a = comp_cond(*c.a, g);
        int b = comp_cond(*c.b, g);
        int d = temp(g);
        op(g, c.kind == Cond::And ? OP_AND : OP_OR, d, a, b);
        return d;
      }
      case Cond::Not: {
        int a = comp_cond(*c.a, g);
        int d = temp(g);
        op(g, OP_NOT, d, a, 0);
        return d;
      }
    }
    return kSlotZero;
  }

  void flag_pending(Group& g, int mask) {
    std::sort(g.pending.begin(), g.pending.end());
    g.pending.erase(std::unique(g.pending.begin(), g.pending.end()), g.pending.end());
    for (int code : g.pending) op(g, OP_FLAGERR, 0, mask, 0, (float)code);
  }

  void compile_event(const Event& e, Group& g) {
    int parent = e.parent_slot < 0 ? kSlotOne : e.parent_slot;
    switch (e.k) {
      case Event::When: {
        int c = comp_cond(*e.cond, g);
        if (!g.pending.empty()) {  // lanes that evaluate the condition: acquired parent, not yet fired
          int m = temp(g);
          op(g, OP_NOT, m, e.slot, 0);
          op(g, OP_AND, m, m, parent);
          flag_pending(g, m);
        }
        op(g, OP_WHEN, e.slot, parent, c);
        break;
      }
      case Event::Until: {
        int c = comp_cond(*e.cond, g);
        if (!g.pending.empty()) {
          int m = temp(g);
          op(g, OP_NOT, m, e.slot, 0);
          op(g, OP_AND, m, m, parent);
          flag_pending(g, m);
        }
        op(g, OP_UNTIL, e.slot, parent, c);
        break;
      }
      case Event::OrErr: {
        op(g, OP_FLAGERR, 0, parent, 0, (float)KT_ERR_UNSUPPORTED);
        break;
      }
      case Event::Emit: {
        // mask = acquired-this-step * prod(1 - ko) over the enclosing untils
        int mask = e.parent_slot < 0 ? kSlotOne : e.parent_slot;
        for (const KoRef& ko : e.kos) {
          if (g.k < ko.start) continue;
          int m = temp(g);
          op(g, OP_NOT, m, ko.slot, 0);
          if (mask != kSlotOne) op(g, OP_AND, m, m, mask);
          mask = m;
        }
        // amount = ((1 * o_inner) * ...) * o_outer, each sampled at this step
        int amt = kSlotOne;
        for (auto it = e.scales.rbegin(); it != e.scales.rend(); ++it) {
          int v = comp_obs(**it, g);
          if (amt == kSlotOne) {
            amt = v;
          } else {
            int d = temp(g);
            op(g, OP_MUL, d, amt, v);
            amt = d;
          }
        }
        if (!g.pending.empty()) flag_pending(g, mask);
        op(g, OP_EMIT, e.out, amt, mask, (float)e.sign);
        break;
      }
    }
  }
};

const char* op_name(int c) {
  static const char* names[] = {"nop",  "const", "spot", "neg",  "add",   "sub",   "mul",   "div",  "max",
                                "min",  "addi",  "subi", "rsubi", "muli", "divi",  "rdivi", "maxi", "mini",
                                "ge",   "gt",    "le",   "lt",   "gei",   "gti",   "lei",   "lti",  "and",
                                "or",   "not",   "when", "until", "emit", "rmax",  "rmin",  "avg",  "flagerr"};
  return c < (int)(sizeof(names) / sizeof(names[0])) ? names[c] : "?";
}

}  // namespace

Program lower(const std::vector<const Con*>& contracts, const std::vector<double>& grid,
              const std::vector<std::string>& model_assets) {
  if (model_assets.size() > (size_t)kMaxAssets) raise(KT_ERR_INVALID_ARGUMENT, "too many assets");
  if (contracts.size() > (size_t)kMaxOutputs) raise(KT_ERR_INVALID_ARGUMENT, "too many contracts in one program");
  Lowerer L(grid, model_assets);
  for (size_t i = 0; i < contracts.size(); ++i) L.add_contract(*contracts[i], (int)i);
  Program p = L.finish();
  p.n_outputs = (int)contracts.size();
  classify_groups(p);
  return p;
}

// A point's op list is a single monitor when it is exactly
//   spot s <- asset a ; t <- s cmp imm ; until(ko, ONE, t)      (knock-out)
//   spot s <- asset a ; avg|rmax|rmin(acc, s)                    (accumulator)
// The fast path does not write the spot slot: nothing else reads it at that point.
static bool single_monitor(const Program& p, const KStep& st, KGroup& g) {
  if (st.count == 0) {
    g = KGroup{G_NONE, 0, 0, 0.0f};
    return true;
  }
  const KOp* o = &p.ops[st.begin];
  if (st.count == 3 && o[0].code == OP_SPOT && o[2].code == OP_UNTIL && o[2].a == kSlotOne &&
      o[1].a == o[0].d && o[2].b == o[1].d && o[1].d != o[2].d) {
    int kind = -1;
    switch (o[1].code) {
      case OP_GEI: kind = G_KO_GE; break;
      case OP_GTI: kind = G_KO_GT; break;
      case OP_LEI: kind = G_KO_LE; break;
      case OP_LTI: kind = G_KO_LT; break;
      default: break;
    }
    if (kind < 0) return false;
    g = KGroup{kind, o[2].d, o[0].a, o[1].imm};
    return true;
  }
  if (st.count == 2 && o[0].code == OP_SPOT && o[1].a == o[0].d &&
      (o[1].code == OP_AVG || o[1].code == OP_RMAX || o[1].code == OP_RMIN)) {
    int kind = o[1].code == OP_AVG ? G_AVG : o[1].code == OP_RMAX ? G_RMAX : G_RMIN;
    g = KGroup{kind, o[1].d, o[0].a, 0.0f};
    return true;
  }
  return false;
}

void classify_groups(Program& p) {
  const int n = (int)p.steps.size();
  p.groups.clear();
  for (int g0 = 0; g0 < n; g0 += 4) {
    KGroup grp{G_GENERIC, 0, 0, 0.0f};
    if (g0 + 4 <= n) {
      KGroup first;
      bool ok = single_monitor(p, p.steps[g0], first);
      for (int s = 1; ok && s < 4; ++s) {
        const KStep& a = p.steps[g0];
        const KStep& b = p.steps[g0 + s];
        ok = a.count == b.count && (a.count == 0 || a.begin == b.begin);
      }
      if (ok) grp = first;
      // fixings of a running sum at some of the four points, nothing at the
      // others (an average_at between its dates): one fast group with the
      // fixing points' mask
      if (!ok) {
        uint32_t mask = 0;
        KGroup fix{G_GENERIC, 0, 0, 0.0f};
        bool fixes = true;
        for (int s = 0; fixes && s < 4; ++s) {
          const KStep& st = p.steps[g0 + s];
          if (st.count == 0) continue;
          KGroup m;
          fixes = single_monitor(p, st, m) && m.kind == G_AVG && (mask == 0 || (m.slot == fix.slot && m.asset == fix.asset));
          fix = m;
          mask |= 1u << s;
        }
        if (fixes && mask != 0) {
          grp = KGroup{G_AVG_AT, fix.slot, fix.asset, 0.0f};
          std::memcpy(&grp.imm, &mask, sizeof mask);
        }
      }
    }
    p.groups.push_back(grp);
  }
  // each op-free group names where its run ends, so that a kernel finds the
  // run's length with one read instead of reading every group's kind
  int next = (int)p.groups.size();
  for (int g = (int)p.groups.size() - 1; g >= 0; --g) {
    if (p.groups[g].kind == G_NONE)
      p.groups[g].slot = next;
    else
      next = g;
  }
}

Program lower_exercise(const Con& contract, const std::vector<double>& grid,
                       const std::vector<std::string>& model_assets) {
  if (!flows_at_acquisition(contract))
    raise(KT_ERR_UNSUPPORTED,
          "LSM: the exercised payoff must pay at the exercise date (no when/anytime/until/or inside)");
  if (model_assets.size() > (size_t)kMaxAssets) raise(KT_ERR_INVALID_ARGUMENT, "too many assets");
  Lowerer L(grid, model_assets);
  L.add_exercise_contract(contract);
  Program p = L.finish();
  p.n_outputs = 1;
  p.groups.assign((p.steps.size() + 3) / 4, KGroup{G_GENERIC, 0, 0, 0.0f});
  return p;
}

AdjPlan classify_adjoint(const Program& prog) {
  AdjPlan plan;
  const size_t n_pts = prog.steps.size();
  if (n_pts == 0) raise(KT_ERR_UNSUPPORTED, "pathwise Greeks: empty grid");
  const size_t last = n_pts - 1;
  plan.feed.assign(n_pts, 0u);
  auto unsupported = [](const std::string& why) {
    raise(KT_ERR_UNSUPPORTED, "pathwise Greeks need a Lipschitz payoff paid on one date: " + why +
                                  " (use bump Greeks)");
  };
  std::vector<bool> is_init(kMaxSlots, false);
  for (const KOp& op : prog.init) is_init[op.d] = true;
  is_init[kSlotZero] = is_init[kSlotOne] = true;
  std::vector<int> acc_of(kMaxSlots, -1);  // slot -> accumulator (AVG also owns slot d + 1)
  auto reads = [](const KOp& op, int* r) {
    switch (op.code) {
      case OP_NOP: case OP_CONST: case OP_SPOT: return 0;
      case OP_NEG: case OP_ADDI: case OP_SUBI: case OP_RSUBI: case OP_MULI: case OP_DIVI: case OP_RDIVI:
      case OP_MAXI: case OP_MINI: case OP_AVG: case OP_RMAX: case OP_RMIN:
        r[0] = op.a;
        return 1;
      default:
        r[0] = op.a;
        r[1] = op.b;
        return 2;
    }
  };
  for (size_t k = 0; k < n_pts; ++k) {
    const KStep& st = prog.steps[k];
    std::vector<int> spot_of(kMaxSlots, -1);  // slot -> asset, if last written by SPOT at this point
    std::vector<bool> written(kMaxSlots, false);
    for (int32_t i = 0; i < st.count; ++i) {
      const KOp& op = prog.ops[(size_t)st.begin + i];
      int rd[2];
      const int nr = reads(op, rd);
      switch (op.code) {
        case OP_NOP: case OP_CONST: case OP_SPOT: case OP_NEG: case OP_ADD: case OP_SUB: case OP_MUL:
        case OP_DIV: case OP_MAX: case OP_MIN: case OP_ADDI: case OP_SUBI: case OP_RSUBI: case OP_MULI:
        case OP_DIVI: case OP_RDIVI: case OP_MAXI: case OP_MINI:
          for (int j = 0; j < nr; ++j) {
            if (acc_of[rd[j]] >= 0 && k != last) unsupported("a running average/extreme read before the payment date");
            if (k == last && !written[rd[j]] && !is_init[rd[j]] && acc_of[rd[j]] < 0)
              unsupported("a value carried between dates");
          }
          break;
        case OP_AVG: case OP_RMAX: case OP_RMIN: {
          if (spot_of[op.a] < 0) unsupported("a running average/extreme of something other than a spot");
          int q = acc_of[op.d];
          if (q < 0) {
            if (plan.n_acc == 4) unsupported("more than 4 running averages/extremes");
            q = plan.n_acc++;
            plan.slot[q] = op.d;
            plan.kind[q] = op.code;
            plan.asset[q] = spot_of[op.a];
            acc_of[op.d] = q;
            if (op.code == OP_AVG) acc_of[(op.d + 1) % kMaxSlots] = q;
          } else if (plan.kind[q] != op.code || plan.asset[q] != spot_of[op.a] || plan.slot[q] != op.d) {
            unsupported("a register shared by different running averages/extremes");
          }
          plan.feed[k] |= 1u << q;
          break;
        }
        case OP_EMIT:
          if (k != last) unsupported("flows on more than one date");
          if (op.b != kSlotOne) unsupported("a conditional flow");
          if (!written[rd[0]] && !is_init[rd[0]] && acc_of[rd[0]] < 0) unsupported("a value carried between dates");
          break;
        default:
          unsupported("a price-dependent condition, digital or knock-out");
      }
      if (op.code != OP_EMIT && op.code != OP_NOP) {
        written[op.d] = true;
        spot_of[op.d] = op.code == OP_SPOT ? (int)op.a : -1;
        if (op.code == OP_AVG) written[(op.d + 1) % kMaxSlots] = true;
      }
    }
  }
  return plan;
}

std::string Program::describe() const {
  std::ostringstream os;
  os << "slots " << n_slots << ", outputs " << n_outputs << ", ops " << ops.size() << ", max/step "
     << max_ops_per_step << (time_only ? ", time-only" : ", price-dependent") << "\n";
  int ng[10] = {0};
  for (const KGroup& g : groups) ng[g.kind]++;
  os << "groups: generic " << ng[G_GENERIC] << ", none " << ng[G_NONE] << ", knock-out "
     << ng[G_KO_GE] + ng[G_KO_GT] + ng[G_KO_LE] + ng[G_KO_LT] << ", avg " << ng[G_AVG] << ", avg at points "
     << ng[G_AVG_AT] << ", rmax/rmin " << ng[G_RMAX] + ng[G_RMIN] << "\n";
  os << "init:";
  for (const KOp& o : init) os << " " << op_name(o.code) << "(" << (int)o.d << "," << o.imm << ")";
  os << "\n";
  int prev_begin = -2, prev_count = -1, run_start = 0;
  auto flush = [&](int upto) {
    if (prev_count < 0) return;
    os << "steps " << run_start << ".." << upto << ":";
    if (prev_count == 0) os << " -";
    for (int i = 0; i < prev_count; ++i) {
      const KOp& o = ops[prev_begin + i];
      os << " " << op_name(o.code) << "(" << (int)o.d << "," << (int)o.a << "," << (int)o.b << "," << o.imm << ")";
    }
    os << "\n";
  };
  for (size_t k = 0; k < steps.size(); ++k) {
    int b = steps[k].count ? steps[k].begin : -1;
    if (b != prev_begin || steps[k].count != prev_count) {
      flush((int)k - 1);
      prev_begin = b;
      prev_count = steps[k].count;
      run_start = (int)k;
    }
  }
  flush((int)steps.size() - 1);
  return os.str();
}

}  // namespace kt
