// oracle.cpp — CPU FP64 restatement of kontract's pricing hot path.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker: only the tests under
// tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load
// it.  The product (kontracts_amd) never links, loads or calls it, and nothing
// here is shipped as part of the pricing engine.
//
// It restates, in plain C++ on f64, the reference crate's algorithm:
//   * serde externally-tagged JSON of Contract/Observable/Condition  (src/ast.rs:34-218)
//   * compile() / Plan / time_grid_forced                          (src/compiler.rs:25-189)
//   * Gbm::simulate exact log step, validate_grid                  (src/simulator.rs:215-253, 413-428)
//   * HestonSimulator::simulate_one_path (floor after each step)  (src/simulator.rs:506-541, 556-580)
//   * CorrelatedGbmN with cholesky_lower                           (src/multi_asset.rs:62-179, src/numerics.rs:158-183)
//   * Observable::eval / Condition::eval                           (src/observable.rs:92-168)
//   * cashflows / first_activation / present_value                 (src/pricer.rs:344-433)
//   * summarize (two-pass f64)                                     (src/pricer.rs:303-328)
//   * price_batch_gbm (merged grid, shared simulation)             (src/pricer.rs:250-294)
//   * greeks_gbm CRN bump-and-reprice                              (src/greeks.rs:60-98)
//   * price_american_lsm / backward_induction / poly_fit / solve_linear (src/lsm.rs:80-295, src/numerics.rs:190-236)
//   * Black-Scholes closed form (exact erfc instead of A&S 7.1.26)  (src/numerics.rs:63-91)
// The one deliberate difference from the reference is the random source: the
// reference draws ChaCha8 + Ziggurat normals (rand_chacha 0.3, rand_distr 0.4),
// which cannot be reproduced here; the oracle draws from the same Philox4x32-10
// counter stream as the GPU engine (or from normals handed in by the caller),
// widened to f64, so per-path values can be compared one to one.
//
// Closed forms used only as pins for tests: Heston (1993) characteristic
// function in the Albrecher et al. (2007) "little trap" form, the
// Reiner-Rubinstein up-and-out call (Haug, "Complete Guide to Option Pricing
// Formulas", 2nd ed., 4.17.1) and the Cox-Ross-Rubinstein tree.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <thread>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------
// Errors: codes follow KontractError's variants (src/error.rs:7-28).
// ---------------------------------------------------------------------------
enum ErrCode {
  E_OK = 0,
  E_UNKNOWN_ASSET = 1,
  E_TIME_OUT_OF_RANGE = 2,
  E_STEP_OUT_OF_RANGE = 3,
  E_INCONSISTENT_PATH = 4,
  E_SERDE = 5,
  E_MALFORMED = 6,
  E_UNSUPPORTED = 7,
  E_ARG = 101,
};

struct KErr {
  int code;
  std::string msg;
};

[[noreturn]] static void fail(int code, const std::string& msg) { throw KErr{code, msg}; }

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
// ---------------------------------------------------------------------------
static inline void philox4x32_10(const uint32_t in[4], const uint32_t k[2], uint32_t out[4]) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
  uint32_t k0 = k[0], k1 = k[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    uint64_t p0 = (uint64_t)M0 * c0;
    uint64_t p1 = (uint64_t)M1 * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

// Uniform on the 2^24-cell midpoint grid: u = (2*floor(w/256)+1) / 2^25.
static inline double uniform_from_word(uint32_t w) {
  return (double)((w >> 7) | 1u) * (1.0 / 33554432.0);
}

// Inverse standard normal CDF in f64: Acklam's rational approximation followed
// by one Halley step on erfc, which brings it to full double accuracy.  The
// upper half is computed by symmetry from 1-p (exact for the dyadic uniforms
// the engine uses) so the Halley residual never subtracts numbers near 1.
static double inv_norm_lower(double p);
static double inv_norm(double p) { return p > 0.5 ? -inv_norm_lower(1.0 - p) : inv_norm_lower(p); }
static double inv_norm_lower(double p) {
  static const double a[6] = {-3.969683028665376e+01, 2.209460984245205e+02, -2.759285104469687e+02,
                              1.383577518672690e+02,  -3.066479806614716e+01, 2.506628277459239e+00};
  static const double b[5] = {-5.447609879822406e+01, 1.615858368580409e+02, -1.556989798598866e+02,
                              6.680131188771972e+01,  -1.328068155288572e+01};
  static const double c[6] = {-7.784894002430293e-03, -3.223964580411365e-01, -2.400758277161838e+00,
                              -2.549732539343734e+00, 4.374664141464968e+00,  2.938163982698783e+00};
  static const double d[4] = {7.784695709041462e-03, 3.224671290700398e-01, 2.445134137142996e+00,
                              3.754408661907416e+00};
  const double plow = 0.02425, phigh = 1.0 - plow;
  double x;
  if (p < plow) {
    double q = std::sqrt(-2.0 * std::log(p));
    x = (((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
        ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1.0);
  } else if (p <= phigh) {
    double q = p - 0.5, r = q * q;
    x = (((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * q /
        (((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1.0);
  } else {
    double q = std::sqrt(-2.0 * std::log(1.0 - p));
    x = -(((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
        ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1.0);
  }
  // Halley refinement.
  double e = 0.5 * std::erfc(-x / std::sqrt(2.0)) - p;
  double u = e * std::sqrt(2.0 * M_PI) * std::exp(x * x / 2.0);
  x = x - u / (1.0 + x * u / 2.0);
  return x;
}

// Normal number j of a path: Philox block j/4, word j%4.  Counter layout:
// (block, path_lo, path_hi, stream); key = (seed_lo, seed_hi).
static inline uint32_t philox_word(uint64_t seed, uint64_t path, uint64_t j, uint32_t stream) {
  uint32_t ctr[4] = {(uint32_t)(j >> 2), (uint32_t)path, (uint32_t)(path >> 32), stream};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t out[4];
  philox4x32_10(ctr, key, out);
  return out[j & 3];
}

// ---------------------------------------------------------------------------
// AST (src/ast.rs:34-206) and a small JSON reader for serde's externally
// tagged representation.
// ---------------------------------------------------------------------------
struct Obs {
  enum K { Const, Spot, Neg, Add, Sub, Mul, Div, Max, Min, Average, RunningMax, RunningMin, AverageAt } k;
  double c = 0.0;
  std::string name;
  int asset = -1;  // resolved index into the path's assets; -1 = unknown
  std::unique_ptr<Obs> a, b;
  bool has_from = false, has_to = false;
  double from = 0.0, to = 0.0;
  std::vector<double> dates;  // AverageAt fixing dates
};
struct Cond {
  enum K { Bool, At, Ge, Gt, Le, Lt, And, Or, Not } k;
  bool bv = false;
  double t = 0.0;
  std::unique_ptr<Obs> oa, ob;
  std::unique_ptr<Cond> a, b;
};
struct Con {
  enum K { Zero, One, Give, And, Or, Scale, When, Anytime, Until } k;
  std::string ccy;
  std::unique_ptr<Obs> obs;
  std::unique_ptr<Cond> cond;
  std::unique_ptr<Con> a, b;
};

struct JVal {
  enum T { Null, Bool, Num, Str, Arr, Obj } t = Null;
  bool b = false;
  double n = 0.0;
  std::string s;
  std::vector<JVal> arr;
  std::vector<std::pair<std::string, JVal>> obj;
};

struct JParser {
  const char* p;
  const char* end;
  void ws() {
    while (p < end && (*p == ' ' || *p == '\n' || *p == '\t' || *p == '\r')) ++p;
  }
  [[noreturn]] void bad(const char* what) { fail(E_SERDE, std::string("json: ") + what); }
  JVal parse() {
    ws();
    if (p >= end) bad("unexpected end");
    JVal v;
    char ch = *p;
    if (ch == '{') {
      ++p;
      v.t = JVal::Obj;
      ws();
      if (p < end && *p == '}') {
        ++p;
        return v;
      }
      for (;;) {
        ws();
        if (p >= end || *p != '"') bad("expected key");
        std::string key = str();
        ws();
        if (p >= end || *p != ':') bad("expected ':'");
        ++p;
        JVal val = parse();
        v.obj.emplace_back(std::move(key), std::move(val));
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        if (p < end && *p == '}') {
          ++p;
          break;
        }
        bad("expected ',' or '}'");
      }
    } else if (ch == '[') {
      ++p;
      v.t = JVal::Arr;
      ws();
      if (p < end && *p == ']') {
        ++p;
        return v;
      }
      for (;;) {
        v.arr.push_back(parse());
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        if (p < end && *p == ']') {
          ++p;
          break;
        }
        bad("expected ',' or ']'");
      }
    } else if (ch == '"') {
      v.t = JVal::Str;
      v.s = str();
    } else if (!strncmp(p, "true", 4) && end - p >= 4) {
      v.t = JVal::Bool;
      v.b = true;
      p += 4;
    } else if (!strncmp(p, "false", 5) && end - p >= 5) {
      v.t = JVal::Bool;
      p += 5;
    } else if (!strncmp(p, "null", 4) && end - p >= 4) {
      v.t = JVal::Null;
      p += 4;
    } else {
      char* e = nullptr;
      v.n = strtod(p, &e);
      if (e == p) bad("bad token");
      v.t = JVal::Num;
      p = e;
    }
    return v;
  }
  std::string str() {
    ++p;  // opening quote
    std::string out;
    while (p < end && *p != '"') {
      if (*p == '\\') {
        ++p;
        if (p >= end) bad("bad escape");
        char e = *p++;
        switch (e) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u': {
            if (end - p < 4) bad("bad \\u");
            unsigned cp = (unsigned)strtoul(std::string(p, p + 4).c_str(), nullptr, 16);
            p += 4;
            if (cp < 0x80) out += (char)cp;
            else if (cp < 0x800) {
              out += (char)(0xC0 | (cp >> 6));
              out += (char)(0x80 | (cp & 0x3F));
            } else {
              out += (char)(0xE0 | (cp >> 12));
              out += (char)(0x80 | ((cp >> 6) & 0x3F));
              out += (char)(0x80 | (cp & 0x3F));
            }
            break;
          }
          default: out += e;
        }
      } else {
        out += *p++;
      }
    }
    if (p >= end) bad("unterminated string");
    ++p;
    return out;
  }
};

static const JVal& tagged(const JVal& v, std::string& tag) {
  if (v.t != JVal::Obj || v.obj.size() != 1) fail(E_SERDE, "json: expected single-key object");
  tag = v.obj[0].first;
  return v.obj[0].second;
}
static double num(const JVal& v) {
  if (v.t != JVal::Num) fail(E_SERDE, "json: expected number");
  return v.n;
}
static const JVal& arr2(const JVal& v, int i) {
  if (v.t != JVal::Arr || v.arr.size() != 2) fail(E_SERDE, "json: expected 2-element array");
  return v.arr[i];
}

static std::unique_ptr<Obs> to_obs(const JVal& v) {
  auto o = std::make_unique<Obs>();
  std::string tag;
  const JVal& body = tagged(v, tag);
  auto bin = [&](Obs::K k) {
    o->k = k;
    o->a = to_obs(arr2(body, 0));
    o->b = to_obs(arr2(body, 1));
  };
  if (tag == "Const") {
    o->k = Obs::Const;
    o->c = num(body);
  } else if (tag == "Spot") {
    if (body.t != JVal::Str) fail(E_SERDE, "json: Spot expects string");
    o->k = Obs::Spot;
    o->name = body.s;
  } else if (tag == "Neg") {
    o->k = Obs::Neg;
    o->a = to_obs(body);
  } else if (tag == "Add") bin(Obs::Add);
  else if (tag == "Sub") bin(Obs::Sub);
  else if (tag == "Mul") bin(Obs::Mul);
  else if (tag == "Div") bin(Obs::Div);
  else if (tag == "Max") bin(Obs::Max);
  else if (tag == "Min") bin(Obs::Min);
  else if (tag == "Average") {
    if (body.t != JVal::Obj) fail(E_SERDE, "json: Average expects object");
    o->k = Obs::Average;
    bool seen_obs = false;
    for (auto& kv : body.obj) {
      if (kv.first == "obs") {
        o->a = to_obs(kv.second);
        seen_obs = true;
      } else if (kv.first == "from_year") {
        if (kv.second.t != JVal::Null) {
          o->has_from = true;
          o->from = num(kv.second);
        }
      } else if (kv.first == "to_year") {
        if (kv.second.t != JVal::Null) {
          o->has_to = true;
          o->to = num(kv.second);
        }
      } else
        fail(E_SERDE, "json: unknown Average field");
    }
    if (!seen_obs) fail(E_SERDE, "json: Average missing obs");
  } else if (tag == "AverageAt") {
    // Extension (no reference counterpart): arithmetic mean of obs over fixing
    // dates, {"AverageAt":{"obs":...,"dates":[t1,...]}}.
    if (body.t != JVal::Obj) fail(E_SERDE, "json: AverageAt expects object");
    o->k = Obs::AverageAt;
    bool seen_obs = false, seen_dates = false;
    for (auto& kv : body.obj) {
      if (kv.first == "obs") {
        o->a = to_obs(kv.second);
        seen_obs = true;
      } else if (kv.first == "dates") {
        if (kv.second.t != JVal::Arr) fail(E_SERDE, "json: AverageAt dates expects array");
        for (auto& d : kv.second.arr) o->dates.push_back(num(d));
        seen_dates = true;
      } else
        fail(E_SERDE, "json: unknown AverageAt field");
    }
    if (!seen_obs || !seen_dates) fail(E_SERDE, "json: AverageAt needs obs and dates");
  } else if (tag == "RunningMax") {
    o->k = Obs::RunningMax;
    o->a = to_obs(body);
  } else if (tag == "RunningMin") {
    o->k = Obs::RunningMin;
    o->a = to_obs(body);
  } else
    fail(E_SERDE, "json: unknown Observable variant " + tag);
  return o;
}

static std::unique_ptr<Cond> to_cond(const JVal& v) {
  auto c = std::make_unique<Cond>();
  std::string tag;
  const JVal& body = tagged(v, tag);
  auto cmp = [&](Cond::K k) {
    c->k = k;
    c->oa = to_obs(arr2(body, 0));
    c->ob = to_obs(arr2(body, 1));
  };
  if (tag == "Bool") {
    if (body.t != JVal::Bool) fail(E_SERDE, "json: Bool expects bool");
    c->k = Cond::Bool;
    c->bv = body.b;
  } else if (tag == "At") {
    c->k = Cond::At;
    c->t = num(body);
  } else if (tag == "Ge") cmp(Cond::Ge);
  else if (tag == "Gt") cmp(Cond::Gt);
  else if (tag == "Le") cmp(Cond::Le);
  else if (tag == "Lt") cmp(Cond::Lt);
  else if (tag == "And" || tag == "Or") {
    c->k = tag == "And" ? Cond::And : Cond::Or;
    c->a = to_cond(arr2(body, 0));
    c->b = to_cond(arr2(body, 1));
  } else if (tag == "Not") {
    c->k = Cond::Not;
    c->a = to_cond(body);
  } else
    fail(E_SERDE, "json: unknown Condition variant " + tag);
  return c;
}

static std::unique_ptr<Con> to_con(const JVal& v) {
  auto c = std::make_unique<Con>();
  if (v.t == JVal::Str) {
    if (v.s == "Zero") {
      c->k = Con::Zero;
      return c;
    }
    fail(E_SERDE, "json: unknown unit variant " + v.s);
  }
  std::string tag;
  const JVal& body = tagged(v, tag);
  if (tag == "One") {
    if (body.t != JVal::Str) fail(E_SERDE, "json: One expects currency string");
    c->k = Con::One;
    c->ccy = body.s;
  } else if (tag == "Give") {
    c->k = Con::Give;
    c->a = to_con(body);
  } else if (tag == "And" || tag == "Or") {
    c->k = tag == "And" ? Con::And : Con::Or;
    c->a = to_con(arr2(body, 0));
    c->b = to_con(arr2(body, 1));
  } else if (tag == "Scale") {
    c->k = Con::Scale;
    c->obs = to_obs(arr2(body, 0));
    c->a = to_con(arr2(body, 1));
  } else if (tag == "When" || tag == "Anytime" || tag == "Until") {
    c->k = tag == "When" ? Con::When : (tag == "Anytime" ? Con::Anytime : Con::Until);
    c->cond = to_cond(arr2(body, 0));
    c->a = to_con(arr2(body, 1));
  } else
    fail(E_SERDE, "json: unknown Contract variant " + tag);
  return c;
}

static std::unique_ptr<Con> parse_contract(const char* json) {
  JParser jp{json, json + strlen(json)};
  JVal v = jp.parse();
  jp.ws();
  if (jp.p != jp.end) fail(E_SERDE, "json: trailing characters");
  return to_con(v);
}

// ---------------------------------------------------------------------------
// compile() / Plan / time_grid_forced (src/compiler.rs:25-189)
// ---------------------------------------------------------------------------
struct Plan {
  std::vector<std::string> assets;
  std::vector<double> fixed_dates;
  double horizon = 0.0;
  bool needs_fine_grid = false;
};

static std::vector<double> sort_dedup(std::vector<double> v) {
  std::stable_sort(v.begin(), v.end(), [](double a, double b) { return a < b; });
  std::vector<double> out;
  for (double x : v) {
    if (!out.empty() && std::fabs(out.back() - x) < 1e-12) continue;
    out.push_back(x);
  }
  return out;
}

struct Acc {
  std::vector<std::string> assets;  // kept sorted, unique
  std::vector<double> dates;
  bool fine = false;
  void add_asset(const std::string& s) {
    auto it = std::lower_bound(assets.begin(), assets.end(), s);
    if (it == assets.end() || *it != s) assets.insert(it, s);
  }
};

static void walk_obs(const Obs& o, Acc& acc) {
  switch (o.k) {
    case Obs::Const: break;
    case Obs::Spot: acc.add_asset(o.name); break;
    case Obs::Neg: walk_obs(*o.a, acc); break;
    case Obs::Average:
    case Obs::RunningMax:
    case Obs::RunningMin:
      acc.fine = true;
      walk_obs(*o.a, acc);
      break;
    case Obs::AverageAt:  // the fixings join the grid as fixed dates
      acc.dates.insert(acc.dates.end(), o.dates.begin(), o.dates.end());
      walk_obs(*o.a, acc);
      break;
    default:
      walk_obs(*o.a, acc);
      walk_obs(*o.b, acc);
  }
}
static void walk_cond(const Cond& c, Acc& acc) {
  switch (c.k) {
    case Cond::Bool: break;
    case Cond::At:
      if (!std::isfinite(c.t) || c.t < 0.0) fail(E_MALFORMED, "invalid acquisition date");
      acc.dates.push_back(c.t);
      break;
    case Cond::Ge:
    case Cond::Gt:
    case Cond::Le:
    case Cond::Lt:
      walk_obs(*c.oa, acc);
      walk_obs(*c.ob, acc);
      break;
    case Cond::And:
    case Cond::Or:
      walk_cond(*c.a, acc);
      walk_cond(*c.b, acc);
      break;
    case Cond::Not: walk_cond(*c.a, acc); break;
  }
}
static bool price_dependent(const Cond& c) {
  switch (c.k) {
    case Cond::Bool:
    case Cond::At: return false;
    case Cond::And:
    case Cond::Or: return price_dependent(*c.a) || price_dependent(*c.b);
    case Cond::Not: return price_dependent(*c.a);
    default: return true;
  }
}
static void walk_con(const Con& c, Acc& acc) {
  switch (c.k) {
    case Con::Zero:
    case Con::One: break;
    case Con::Give: walk_con(*c.a, acc); break;
    case Con::And:
    case Con::Or:
      walk_con(*c.a, acc);
      walk_con(*c.b, acc);
      break;
    case Con::Scale:
      walk_obs(*c.obs, acc);
      walk_con(*c.a, acc);
      break;
    default:
      walk_cond(*c.cond, acc);
      if (price_dependent(*c.cond)) acc.fine = true;
      walk_con(*c.a, acc);
  }
}
static Plan compile(const Con& c) {
  Acc acc;
  walk_con(c, acc);
  Plan p;
  double h = 0.0;
  for (double d : acc.dates)
    if (d > h) h = d;
  p.assets = acc.assets;
  p.fixed_dates = sort_dedup(acc.dates);
  p.horizon = h;
  p.needs_fine_grid = acc.fine;
  return p;
}
static std::vector<double> time_grid_forced(const Plan& p, uint64_t spy, bool force) {
  std::vector<double> pts{0.0};
  pts.insert(pts.end(), p.fixed_dates.begin(), p.fixed_dates.end());
  pts.push_back(p.horizon);
  if ((p.needs_fine_grid || force) && p.horizon > 0.0 && spy > 0) {
    uint64_t n = (uint64_t)std::ceil(p.horizon * (double)spy);
    for (uint64_t k = 1; k <= n; ++k) pts.push_back(p.horizon * (double)k / (double)n);
  }
  return sort_dedup(pts);
}

// ---------------------------------------------------------------------------
// Path evaluation (src/observable.rs:92-168, src/pricer.rs:344-433)
// ---------------------------------------------------------------------------
struct PathView {
  const double* times;
  size_t n;
  const double* spots;  // [n_assets][n]
};

static void resolve_obs(Obs& o, const std::vector<std::string>& names) {
  if (o.k == Obs::Spot) {
    o.asset = -1;
    for (size_t i = 0; i < names.size(); ++i)
      if (names[i] == o.name) o.asset = (int)i;
  }
  if (o.a) resolve_obs(*o.a, names);
  if (o.b) resolve_obs(*o.b, names);
}
static void resolve_cond(Cond& c, const std::vector<std::string>& names) {
  if (c.oa) resolve_obs(*c.oa, names);
  if (c.ob) resolve_obs(*c.ob, names);
  if (c.a) resolve_cond(*c.a, names);
  if (c.b) resolve_cond(*c.b, names);
}
static void resolve_con(Con& c, const std::vector<std::string>& names) {
  if (c.obs) resolve_obs(*c.obs, names);
  if (c.cond) resolve_cond(*c.cond, names);
  if (c.a) resolve_con(*c.a, names);
  if (c.b) resolve_con(*c.b, names);
}

static double eval_obs(const Obs& o, const PathView& p, size_t t) {
  switch (o.k) {
    case Obs::Const: return o.c;
    case Obs::Spot:
      if (o.asset < 0) fail(E_UNKNOWN_ASSET, o.name);
      if (t >= p.n) fail(E_STEP_OUT_OF_RANGE, "step");
      return p.spots[(size_t)o.asset * p.n + t];
    case Obs::Neg: return -eval_obs(*o.a, p, t);
    case Obs::Add: {
      double x = eval_obs(*o.a, p, t);
      return x + eval_obs(*o.b, p, t);
    }
    case Obs::Sub: {
      double x = eval_obs(*o.a, p, t);
      return x - eval_obs(*o.b, p, t);
    }
    case Obs::Mul: {
      double x = eval_obs(*o.a, p, t);
      return x * eval_obs(*o.b, p, t);
    }
    case Obs::Div: {
      double x = eval_obs(*o.a, p, t);
      return x / eval_obs(*o.b, p, t);
    }
    case Obs::Max: {  // f64::max: a NaN operand yields the other operand
      double x = eval_obs(*o.a, p, t);
      return std::fmax(x, eval_obs(*o.b, p, t));
    }
    case Obs::Min: {
      double x = eval_obs(*o.a, p, t);
      return std::fmin(x, eval_obs(*o.b, p, t));
    }
    case Obs::Average: {
      if (t >= p.n) fail(E_STEP_OUT_OF_RANGE, "step");
      double t_now = p.times[t];
      double t_from = o.has_from ? o.from : 0.0;
      double t_to = o.has_to ? o.to : t_now;
      if (t_from > t_to + 1e-12) fail(E_MALFORMED, "average window: from > to");
      double sum = 0.0;
      size_t cnt = 0;
      for (size_t i = 0; i < p.n; ++i) {
        double ti = p.times[i];
        if (ti >= t_from - 1e-12 && ti <= t_to + 1e-12) {
          sum += eval_obs(*o.a, p, i);
          ++cnt;
        }
      }
      if (cnt == 0) fail(E_MALFORMED, "average window empty");
      return sum / (double)cnt;
    }
    case Obs::AverageAt: {
      // mean of obs at the fixings on or before now; each fixing must be a
      // grid point; fixings strictly increasing
      if (t >= p.n) fail(E_STEP_OUT_OF_RANGE, "step");
      if (o.dates.empty()) fail(E_MALFORMED, "AverageAt: no fixing dates");
      for (size_t i = 1; i < o.dates.size(); ++i)
        if (!(o.dates[i] > o.dates[i - 1])) fail(E_MALFORMED, "AverageAt: fixing dates must increase");
      double t_now = p.times[t];
      double sum = 0.0;
      size_t cnt = 0;
      for (double d : o.dates) {
        if (d > t_now + 1e-12) break;
        size_t hit = p.n;
        for (size_t i = 0; i < p.n; ++i)
          if (std::fabs(p.times[i] - d) <= 1e-12) {
            hit = i;
            break;
          }
        if (hit == p.n) fail(E_TIME_OUT_OF_RANGE, "AverageAt: fixing date not on the grid");
        sum += eval_obs(*o.a, p, hit);
        ++cnt;
      }
      if (cnt == 0) fail(E_MALFORMED, "AverageAt: no fixing on or before the evaluation date");
      return sum / (double)cnt;
    }
    case Obs::RunningMax: {
      double acc = -INFINITY;
      for (size_t i = 0; i <= t; ++i) acc = std::fmax(acc, eval_obs(*o.a, p, i));
      return acc;
    }
    case Obs::RunningMin: {
      double acc = INFINITY;
      for (size_t i = 0; i <= t; ++i) acc = std::fmin(acc, eval_obs(*o.a, p, i));
      return acc;
    }
  }
  return 0.0;
}

static bool eval_cond(const Cond& c, const PathView& p, size_t t) {
  switch (c.k) {
    case Cond::Bool: return c.bv;
    case Cond::At:
      if (t >= p.n) fail(E_STEP_OUT_OF_RANGE, "step");
      return p.times[t] + 1e-12 >= c.t;
    case Cond::Ge: {
      double x = eval_obs(*c.oa, p, t);
      return x >= eval_obs(*c.ob, p, t);
    }
    case Cond::Gt: {
      double x = eval_obs(*c.oa, p, t);
      return x > eval_obs(*c.ob, p, t);
    }
    case Cond::Le: {
      double x = eval_obs(*c.oa, p, t);
      return x <= eval_obs(*c.ob, p, t);
    }
    case Cond::Lt: {
      double x = eval_obs(*c.oa, p, t);
      return x < eval_obs(*c.ob, p, t);
    }
    case Cond::And: return eval_cond(*c.a, p, t) && eval_cond(*c.b, p, t);
    case Cond::Or: return eval_cond(*c.a, p, t) || eval_cond(*c.b, p, t);
    case Cond::Not: return !eval_cond(*c.a, p, t);
  }
  return false;
}

static bool first_activation(const Cond& c, const PathView& p, size_t start, size_t& out) {
  for (size_t t = start; t < p.n; ++t)
    if (eval_cond(c, p, t)) {
      out = t;
      return true;
    }
  return false;
}

typedef std::vector<std::pair<double, size_t>> Flows;

static void cashflows(const Con& c, size_t t_idx, const PathView& p, Flows& out) {
  switch (c.k) {
    case Con::Zero: return;
    case Con::One: out.emplace_back(1.0, t_idx); return;
    case Con::Give: {
      size_t s = out.size();
      cashflows(*c.a, t_idx, p, out);
      for (size_t i = s; i < out.size(); ++i) out[i].first = -out[i].first;
      return;
    }
    case Con::And:
      cashflows(*c.a, t_idx, p, out);
      cashflows(*c.b, t_idx, p, out);
      return;
    case Con::Scale: {
      size_t s = out.size();
      cashflows(*c.a, t_idx, p, out);
      for (size_t i = s; i < out.size(); ++i) out[i].first *= eval_obs(*c.obs, p, out[i].second);
      return;
    }
    case Con::When:
    case Con::Anytime: {
      size_t k;
      if (first_activation(*c.cond, p, t_idx, k)) cashflows(*c.a, k, p, out);
      return;
    }
    case Con::Until: {
      size_t s = out.size();
      cashflows(*c.a, t_idx, p, out);
      size_t k;
      if (first_activation(*c.cond, p, t_idx, k)) {
        size_t w = s;
        for (size_t i = s; i < out.size(); ++i)
          if (out[i].second < k) out[w++] = out[i];
        out.resize(w);
      }
      return;
    }
    case Con::Or: fail(E_UNSUPPORTED, "or (LSM only)");
  }
}

static double present_value(const Con& c, const PathView& p, double rate, Flows& scratch) {
  scratch.clear();
  cashflows(c, 0, p, scratch);
  double pv = 0.0;
  for (auto& f : scratch) pv += f.first * std::exp(-rate * p.times[f.second]);
  return pv;
}

// ---------------------------------------------------------------------------
// Models (src/simulator.rs, src/multi_asset.rs)
// ---------------------------------------------------------------------------
}  // namespace orc

extern "C" {
typedef struct {
  int kind;  // 0 GBM, 1 Heston, 2 correlated GBM
  int n_assets;
  const char* const* names;
  const double* s0;
  const double* mu;
  const double* sigma;
  const double* corr;  // row-major n_assets^2 (kind 2)
  double v0, kappa, theta, sigma_v, rho;
  int heston_scheme;  // 0 floor after each step (reference), 1 full truncation
} orc_model;

typedef struct {
  uint64_t n_paths;
  uint64_t seed;
  uint64_t steps_per_year;
  double rate;
  const double* grid;  // optional explicit grid
  uint64_t n_grid;
  int threads;  // 0 = hardware concurrency
} orc_cfg;

typedef struct {
  double price, sample_std, std_error, ci95_low, ci95_high;
  uint64_t n_paths;
} orc_result;
}

namespace orc {

static const double Z95 = 1.959963984540054;

static orc_result summarize(const std::vector<double>& pvs) {
  size_t n = pvs.size();
  double s = 0.0;
  for (double x : pvs) s += x;
  double price = s / (double)(n > 0 ? n : 1);
  double sd = 0.0;
  if (n > 1) {
    double v = 0.0;
    for (double x : pvs) v += (x - price) * (x - price);
    sd = std::sqrt(v / ((double)n - 1.0));
  }
  double se = n > 0 ? sd / std::sqrt((double)n) : 0.0;
  return orc_result{price, sd, se, price - Z95 * se, price + Z95 * se, (uint64_t)n};
}

static std::vector<std::vector<double>> cholesky_lower(const std::vector<std::vector<double>>& a) {
  size_t n = a.size();
  std::vector<std::vector<double>> l(n, std::vector<double>(n, 0.0));
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j <= i; ++j) {
      double sum = a[i][j];
      for (size_t k = 0; k < j; ++k) sum -= l[i][k] * l[j][k];
      if (i == j)
        l[i][j] = std::sqrt(std::max(sum, 0.0));
      else {
        double piv = l[j][j];
        l[i][j] = piv > 0.0 ? sum / piv : 0.0;
      }
    }
  return l;
}

static void validate_grid(const std::vector<double>& g) {
  if (g.empty()) fail(E_INCONSISTENT_PATH, "empty grid");
  double prev = 0.0;
  for (double t : g) {
    if (t < prev) fail(E_INCONSISTENT_PATH, "grid not increasing");
    prev = t;
  }
}

struct Model {
  int kind;
  std::vector<std::string> names;
  std::vector<double> s0, mu, sigma;
  std::vector<std::vector<double>> chol;
  double v0 = 0, kappa = 0, theta = 0, sigma_v = 0, rho = 0;
  int scheme = 0;
  int m = 1;  // normals per grid point
  bool requires_fine() const { return kind == 1; }
};

static Model make_model(const orc_model* m) {
  Model md;
  md.kind = m->kind;
  if (m->n_assets < 1 || m->n_assets > 64) fail(E_ARG, "bad n_assets");
  for (int i = 0; i < m->n_assets; ++i) {
    md.names.push_back(m->names[i]);
    md.s0.push_back(m->s0[i]);
    md.mu.push_back(m->mu[i]);
    md.sigma.push_back(m->sigma ? m->sigma[i] : 0.0);
  }
  if (md.kind == 0) {
    if (m->n_assets != 1) fail(E_ARG, "GBM has one asset");
    md.m = 1;
  } else if (md.kind == 1) {
    if (m->n_assets != 1) fail(E_ARG, "Heston has one asset");
    md.v0 = m->v0;
    md.kappa = m->kappa;
    md.theta = m->theta;
    md.sigma_v = m->sigma_v;
    md.rho = m->rho;
    md.scheme = m->heston_scheme;
    md.m = 2;
  } else if (md.kind == 2) {
    int k = m->n_assets;
    if (k < 2) fail(E_MALFORMED, "CorrelatedGbmN: at least 2 assets");
    for (int i = 0; i < k; ++i)
      if (md.sigma[i] < 0.0) fail(E_MALFORMED, "CorrelatedGbmN: negative vol");
    std::vector<std::vector<double>> c(k, std::vector<double>(k));
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) c[i][j] = m->corr[i * k + j];
    md.chol = cholesky_lower(c);
    md.m = k;
  } else
    fail(E_ARG, "unknown model kind");
  return md;
}

static void validate_model(const Model& md) {
  if (md.kind == 0 && md.sigma[0] < 0.0) fail(E_MALFORMED, "GBM sigma must be non-negative");
  if (md.kind == 1) {
    if (!(md.rho >= -1.0 && md.rho <= 1.0)) fail(E_MALFORMED, "Heston rho must be in [-1, 1]");
    if (md.v0 < 0.0) fail(E_MALFORMED, "Heston v0 must be non-negative");
    if (md.sigma_v < 0.0) fail(E_MALFORMED, "Heston sigma_v must be non-negative");
    if (md.theta < 0.0) fail(E_MALFORMED, "Heston theta must be non-negative");
    if (md.kappa < 0.0) fail(E_MALFORMED, "Heston kappa must be non-negative");
  }
}

// Normal source: either the Philox stream or a caller-provided float array
// laid out [path][grid point][component].
// flip = 0xffffffff draws the antithetic path: every Philox word complemented,
// so u -> 1 - u on the dyadic cell grid and each normal exactly negated.
struct Normals {
  uint64_t seed;
  const float* fed;
  size_t n_grid;
  int m;
  uint32_t flip = 0;
  double get(uint64_t path, size_t k, int c) const {
    uint64_t j = (uint64_t)k * (uint64_t)m + (uint64_t)c;
    if (fed) {
      double z = (double)fed[((size_t)path * n_grid + k) * (size_t)m + (size_t)c];
      return flip ? -z : z;
    }
    return inv_norm(uniform_from_word(philox_word(seed, path, j, 0u) ^ flip));
  }
};

// Simulate one path into row[a*n + k].
static void simulate_path(const Model& md, const std::vector<double>& times, uint64_t path, const Normals& nz,
                          double* row) {
  size_t n = times.size();
  if (md.kind == 0) {
    double s = md.s0[0], prev = 0.0, mu = md.mu[0], sg = md.sigma[0];
    for (size_t k = 0; k < n; ++k) {
      double dt = times[k] - prev;
      if (dt > 0.0) {
        double z = nz.get(path, k, 0);
        double drift = (mu - 0.5 * sg * sg) * dt;
        double diff = sg * std::sqrt(dt) * z;
        s *= std::exp(drift + diff);
      }
      row[k] = s;
      prev = times[k];
    }
  } else if (md.kind == 1) {
    double rho_perp = std::sqrt(std::max(1.0 - md.rho * md.rho, 0.0));
    double s = md.s0[0], v = std::max(md.v0, 0.0), prev = 0.0, r = md.mu[0];
    for (size_t k = 0; k < n; ++k) {
      double dt = times[k] - prev;
      if (dt > 0.0) {
        double sdt = std::sqrt(dt);
        double z1 = nz.get(path, k, 0), z2 = nz.get(path, k, 1);
        if (md.scheme == 0) {
          double sv = std::sqrt(v);
          double dws = z1 * sdt, dwv = (md.rho * z1 + rho_perp * z2) * sdt;
          s *= std::exp(r * dt - 0.5 * v * dt + sv * dws);
          v = std::max(v + md.kappa * (md.theta - v) * dt + md.sigma_v * sv * dwv, 0.0);
        } else {
          double vp = std::max(v, 0.0), sv = std::sqrt(vp);
          double dws = z1 * sdt, dwv = (md.rho * z1 + rho_perp * z2) * sdt;
          s *= std::exp(r * dt - 0.5 * vp * dt + sv * dws);
          v = v + md.kappa * (md.theta - vp) * dt + md.sigma_v * sv * dwv;
        }
      }
      row[k] = s;
      prev = times[k];
    }
  } else {
    size_t K = md.names.size();
    std::vector<double> z(K), w(K), s(md.s0);
    double prev = 0.0;
    for (size_t k = 0; k < n; ++k) {
      for (size_t c = 0; c < K; ++c) z[c] = nz.get(path, k, (int)c);
      for (size_t i = 0; i < K; ++i) {
        double acc = 0.0;
        for (size_t c = 0; c <= i; ++c) acc += md.chol[i][c] * z[c];
        w[i] = acc;
      }
      double dt = times[k] - prev;
      for (size_t i = 0; i < K; ++i) {
        if (dt > 0.0)
          s[i] *= std::exp((md.mu[i] - 0.5 * md.sigma[i] * md.sigma[i]) * dt + md.sigma[i] * std::sqrt(dt) * w[i]);
        row[i * n + k] = s[i];
      }
      prev = times[k];
    }
  }
}

static int nthreads(int req) {
  if (req > 0) return req;
  unsigned h = std::thread::hardware_concurrency();
  return h ? (int)h : 1;
}

// Run fn(path_begin, path_end, thread_id) over [0, n) split in chunks; errors
// propagate from the lowest failing chunk (deterministic error choice).
static void parallel_paths(uint64_t n, int threads, const std::function<void(uint64_t, uint64_t)>& fn) {
  if (n == 0) return;
  const uint64_t chunk = 4096;
  uint64_t nchunks = (n + chunk - 1) / chunk;
  std::atomic<uint64_t> next{0};
  std::vector<KErr> errs(nchunks);
  std::vector<char> bad(nchunks, 0);
  auto worker = [&]() {
    for (;;) {
      uint64_t c = next.fetch_add(1);
      if (c >= nchunks) return;
      try {
        fn(c * chunk, std::min(n, (c + 1) * chunk));
      } catch (KErr& e) {
        errs[c] = e;
        bad[c] = 1;
      }
    }
  };
  int T = (int)std::min<uint64_t>((uint64_t)nthreads(threads), nchunks);
  std::vector<std::thread> th;
  for (int i = 1; i < T; ++i) th.emplace_back(worker);
  worker();
  for (auto& t : th) t.join();
  for (uint64_t c = 0; c < nchunks; ++c)
    if (bad[c]) throw errs[c];
}

// price_on_paths for a list of contracts over one shared simulation.
static void price_contracts(std::vector<std::unique_ptr<Con>>& cons, const Model& md,
                            const std::vector<double>& grid, uint64_t n_paths, uint64_t seed, double rate,
                            const float* fed, int threads, std::vector<std::vector<double>>& pvs,
                            uint32_t flip = 0) {
  validate_model(md);
  validate_grid(grid);
  for (auto& c : cons) resolve_con(*c, md.names);
  size_t n = grid.size(), K = md.names.size();
  pvs.assign(cons.size(), std::vector<double>(n_paths));
  Normals nz{seed, fed, n, md.m, flip};
  parallel_paths(n_paths, threads, [&](uint64_t b, uint64_t e) {
    std::vector<double> row(K * n);
    Flows scratch;
    for (uint64_t i = b; i < e; ++i) {
      simulate_path(md, grid, i, nz, row.data());
      PathView pv{grid.data(), n, row.data()};
      for (size_t c = 0; c < cons.size(); ++c) pvs[c][i] = present_value(*cons[c], pv, rate, scratch);
    }
  });
}

static std::vector<double> grid_for(const Con& c, const Model& md, const orc_cfg* cfg) {
  if (cfg->grid && cfg->n_grid > 0) return std::vector<double>(cfg->grid, cfg->grid + cfg->n_grid);
  Plan p = compile(c);
  return time_grid_forced(p, cfg->steps_per_year, md.requires_fine());
}

static int report(const KErr& e, char* err, int cap) {
  if (err && cap > 0) snprintf(err, (size_t)cap, "%s", e.msg.c_str());
  return e.code;
}

// LSM (src/lsm.rs) -----------------------------------------------------------
static std::vector<double> solve_linear(std::vector<std::vector<double>> a, std::vector<double> b) {
  size_t n = b.size();
  for (size_t col = 0; col < n; ++col) {
    size_t piv = col;
    double best = std::fabs(a[col][col]);
    for (size_t r = col + 1; r < n; ++r) {
      double v = std::fabs(a[r][col]);
      if (v > best) {
        best = v;
        piv = r;
      }
    }
    if (best < 1e-12) return std::vector<double>(n, 0.0);
    std::swap(a[col], a[piv]);
    std::swap(b[col], b[piv]);
    for (size_t r = col + 1; r < n; ++r) {
      double f = a[r][col] / a[col][col];
      for (size_t c = col; c < n; ++c) a[r][c] -= f * a[col][c];
      b[r] -= f * b[col];
    }
  }
  std::vector<double> x(n, 0.0);
  for (size_t r = n; r-- > 0;) {
    double s = b[r];
    for (size_t c = r + 1; c < n; ++c) s -= a[r][c] * x[c];
    x[r] = s / a[r][r];
  }
  return x;
}

}  // namespace orc

using namespace orc;

extern "C" {

int orc_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  philox4x32_10(ctr, key, out);
  return 0;
}

double orc_uniform_from_word(uint32_t w) { return uniform_from_word(w); }
double orc_inv_norm(double u) { return inv_norm(u); }

// Normals j = first .. first+count-1 of one path (f64).
void orc_normals(uint64_t seed, uint64_t path, uint64_t first, uint64_t count, double* out) {
  for (uint64_t i = 0; i < count; ++i) out[i] = inv_norm(uniform_from_word(philox_word(seed, path, first + i, 0u)));
}

// compile(): grid (capacity cap), fine flag, horizon, fixed dates (cap), assets
// as '\n'-joined names.  Returns the error code.
int orc_compile(const char* json, uint64_t spy, int force_fine, double* grid, uint64_t* n_grid, uint64_t cap,
                int* fine, double* horizon, double* dates, uint64_t* n_dates, char* assets, int assets_cap,
                char* err, int errcap) {
  try {
    auto c = parse_contract(json);
    Plan p = compile(*c);
    auto g = time_grid_forced(p, spy, force_fine != 0);
    *n_grid = g.size();
    for (size_t i = 0; i < g.size() && i < cap; ++i) grid[i] = g[i];
    *fine = p.needs_fine_grid;
    *horizon = p.horizon;
    *n_dates = p.fixed_dates.size();
    for (size_t i = 0; i < p.fixed_dates.size() && i < cap; ++i) dates[i] = p.fixed_dates[i];
    std::string a;
    for (size_t i = 0; i < p.assets.size(); ++i) {
      if (i) a += "\n";
      a += p.assets[i];
    }
    snprintf(assets, (size_t)assets_cap, "%s", a.c_str());
    return 0;
  } catch (KErr& e) {
    return report(e, err, errcap);
  }
}

// Evaluate one contract on one hand-built path (tests of src/observable.rs and
// src/pricer.rs semantics).  spots is [n_assets][n]; returns the pv and the
// flows (amount, step) up to cap.
int orc_eval_path(const char* json, const double* times, uint64_t n, const char* const* names, int n_assets,
                  const double* spots, double rate, double* pv, double* flow_amt, uint64_t* flow_idx,
                  uint64_t* n_flows, uint64_t cap, char* err, int errcap) {
  try {
    auto c = parse_contract(json);
    std::vector<std::string> nm(names, names + n_assets);
    resolve_con(*c, nm);
    PathView p{times, (size_t)n, spots};
    Flows f;
    cashflows(*c, 0, p, f);
    double s = 0.0;
    for (auto& x : f) s += x.first * std::exp(-rate * times[x.second]);
    *pv = s;
    *n_flows = f.size();
    for (size_t i = 0; i < f.size() && i < cap; ++i) {
      flow_amt[i] = f[i].first;
      flow_idx[i] = f[i].second;
    }
    return 0;
  } catch (KErr& e) {
    return report(e, err, errcap);
  }
}

// Evaluate an observable given as JSON at step t of a hand-built path.
int orc_eval_observable(const char* json, const double* times, uint64_t n, const char* const* names, int n_assets,
                        const double* spots, uint64_t t, double* out, char* err, int errcap) {
  try {
    JParser jp{json, json + strlen(json)};
    auto o = to_obs(jp.parse());
    std::vector<std::string> nm(names, names + n_assets);
    resolve_obs(*o, nm);
    PathView p{times, (size_t)n, spots};
    *out = eval_obs(*o, p, (size_t)t);
    return 0;
  } catch (KErr& e) {
    return report(e, err, errcap);
  }
}

// price_gbm (src/pricer.rs:97-115) under any of the three models.
// normals (nullable): [n_paths][n_grid][m] floats fed instead of Philox.
// pv_out (nullable): per-path present values.  grid_out/n_grid_out (nullable):
// the grid actually used (capacity grid_cap).
int orc_price(const char* json, const orc_model* m, const orc_cfg* cfg, const float* normals, double* pv_out,
              double* grid_out, uint64_t* n_grid_out, uint64_t grid_cap, orc_result* res, char* err, int errcap) {
  try {
    Model md = make_model(m);
    std::vector<std::unique_ptr<Con>> cons;
    cons.push_back(parse_contract(json));
    auto grid = grid_for(*cons[0], md, cfg);
    if (n_grid_out) *n_grid_out = grid.size();
    if (grid_out)
      for (size_t i = 0; i < grid.size() && i < grid_cap; ++i) grid_out[i] = grid[i];
    std::vector<std::vector<double>> pvs;
    price_contracts(cons, md, grid, cfg->n_paths, cfg->seed, cfg->rate, normals, cfg->threads, pvs);
    if (pv_out) memcpy(pv_out, pvs[0].data(), pvs[0].size() * sizeof(double));
    *res = summarize(pvs[0]);
    return 0;
  } catch (KErr& e) {
    return report(e, err, errcap);
  }
}

// Simulator::simulate_paths (src/simulator.rs:70-178; GBM :215-253, Heston
// :506-541, CorrelatedGbmN src/multi_asset.rs:130-179) on an explicit grid:
// out[(path * n_assets + asset) * n_grid + k] = spot of `asset` at grid point
// k, drawn from the same Philox stream as orc_price.
int orc_simulate(const orc_model* m, const double* grid, uint64_t n_grid, uint64_t n_paths, uint64_t seed,
                 int threads, double* out, char* err, int errcap) {
  try {
    Model md = make_model(m);
    validate_model(md);
    std::vector<double> g(grid, grid + n_grid);
    validate_grid(g);
    const size_t K = md.names.size(), n = g.size();
    Normals nz{seed, nullptr, n, md.m, 0u};
    parallel_paths(n_paths, threads, [&](uint64_t b, uint64_t e) {
      for (uint64_t i = b; i < e; ++i) simulate_path(md, g, i, nz, out + i * K * n);
    });
    return 0;
  } catch (KErr& e) {
    return report(e, err, errcap);
  }
}

// price_batch_gbm (src/pricer.rs:250-294): merged plan, one simulation.
int orc_price_batch(const char* const* jsons, uint64_t n_contracts, const orc_model* m, const orc_cfg* cfg,
                    orc_result* out, char* err, int errcap) {
  try {
    if (n_contracts == 0) return 0;
    Model md = make_model(m);
    std::vector<std::unique_ptr<Con>> cons;
    Plan merged;
    for (uint64_t i = 0; i < n_contracts; ++i) {
      cons.push_back(parse_contract(jsons[i]));
      Plan p = compile(*cons.back());
      merged.fixed_dates.insert(merged.fixed_dates.end(), p.fixed_dates.begin(), p.fixed_dates.end());
      merged.horizon = std::max(merged.horizon, p.horizon);
      merged.needs_fine_grid = merged.needs_fine_grid || p.needs_fine_grid;
    }
    auto grid = time_grid_forced(merged, cfg->steps_per_year, md.requires_fine());
    std::vector<std::vector<double>> pvs;
    price_contracts(cons, md, grid, cfg->n_paths, cfg->seed, cfg->rate, nullptr, cfg->threads, pvs);
    for (uint64_t i = 0; i < n_contracts; ++i) out[i] = summarize(pvs[i]);
    return 0;
  } catch (KErr& e) {
    return report(e, err, errcap);
  }
}

// Black-Scholes call with the Abramowitz-Stegun 7.1.26 normal CDF
// (src/numerics.rs:22-41, 63-71): the control variate's correction.
static double ncdf_as(double x) {
  double z = x / std::sqrt(2.0);
  double sign = z < 0.0 ? -1.0 : 1.0, a = std::fabs(z);
  double t = 1.0 / (1.0 + 0.3275911 * a);
  double poly = ((((1.061405429 * t - 1.453152027) * t) + 1.421413741) * t - 0.284496736) * t + 0.254829592;
  return 0.5 * (1.0 + sign * (1.0 - poly * t * std::exp(-a * a)));
}
double orc_bs_call_as(double s, double k, double t, double r, double sigma) {
  if (t <= 0.0 || sigma <= 0.0) return std::max(s - k * std::exp(-r * t), 0.0);
  double vst = sigma * std::sqrt(t);
  double d1 = (std::log(s / k) + (r + 0.5 * sigma * sigma) * t) / vst, d2 = d1 - vst;
  return s * ncdf_as(d1) - k * std::exp(-r * t) * ncdf_as(d2);
}
double orc_norm_cdf_as(double x) { return ncdf_as(x); }

// price_gbm with cfg.variance_reduction (src/pricer.rs:117-228,
// src/variance_reduction.rs:76-235).  Antithetic pairs exist for the one-asset
// GBM only (src/simulator.rs:132-140, 297-314): pair i is path i of the
// stream and its negation.  The control is the ATM call on the model's first
// asset at the grid's last date with beta = 1.
int orc_price_vr(const char* json, const orc_model* m, const orc_cfg* cfg, int antithetic, int control_variate,
                 orc_result* res, char* err, int errcap) {
  try {
    Model md = make_model(m);
    std::vector<std::unique_ptr<Con>> cons;
    cons.push_back(parse_contract(json));
    auto grid = grid_for(*cons[0], md, cfg);
    const bool gbm = md.kind == 0;
    const bool anti = antithetic && gbm;
    const double s0 = gbm ? md.s0[0] : 100.0, sg = gbm ? md.sigma[0] : 0.2;
    const double t_mat = grid.empty() ? 1.0 : grid.back();
    const double bs = orc_bs_call_as(s0, s0, t_mat, cfg->rate, sg);
    if (control_variate) {
      char buf[512];
      snprintf(buf, sizeof buf,
               "{\"When\":[{\"At\":%.17g},{\"Scale\":[{\"Max\":[{\"Sub\":[{\"Spot\":\"%s\"},{\"Const\":%.17g}]},"
               "{\"Const\":0.0}]},{\"One\":\"USD\"}]}]}",
               t_mat, md.names[0].c_str(), s0);
      cons.push_back(parse_contract(buf));
    }
    if (anti) {
      const uint64_t n_half = cfg->n_paths / 2;
      std::vector<std::vector<double>> base, neg;
      price_contracts(cons, md, grid, n_half, cfg->seed, cfg->rate, nullptr, cfg->threads, base, 0u);
      price_contracts(cons, md, grid, n_half, cfg->seed, cfg->rate, nullptr, cfg->threads, neg, 0xffffffffu);
      std::vector<std::vector<double>> pair(cons.size(), std::vector<double>(n_half));
      for (size_t c = 0; c < cons.size(); ++c)
        for (uint64_t i = 0; i < n_half; ++i) pair[c][i] = 0.5 * (base[c][i] + neg[c][i]);
      orc_result rt = summarize(pair[0]);
      if (!control_variate) {
        *res = rt;
        return 0;
      }
      orc_result rc = summarize(pair[1]);
      double price = rt.price - 1.0 * (rc.price - bs);
      uint64_t n = 2 * n_half;
      double se = rt.sample_std / std::sqrt((double)n);
      *res = orc_result{price, rt.sample_std, se, price - Z95 * se, price + Z95 * se, n};
      return 0;
    }
    std::vector<std::vector<double>> pvs;
    price_contracts(cons, md, grid, cfg->n_paths, cfg->seed, cfg->rate, nullptr, cfg->threads, pvs);
    if (!control_variate) {
      *res = summarize(pvs[0]);
      return 0;
    }
    std::vector<double> y(cfg->n_paths);
    for (uint64_t i = 0; i < cfg->n_paths; ++i) y[i] = pvs[0][i] - 1.0 * (pvs[1][i] - bs);
    *res = summarize(y);
    return 0;
  } catch (KErr& e) {
    return report(e, err, errcap);
  }
}

// greeks_gbm (src/greeks.rs:60-98): 7 CRN repricings of a one-asset GBM with
// drift = rate.  out = {price, delta, gamma, vega, rho}.
int orc_greeks_gbm(const char* json, const char* asset, double s0, double sigma, const orc_cfg* cfg, double h_spot,
                   double h_vol, double h_rate, double* out, char* err, int errcap) {
  try {
    auto reprice = [&](double s, double sg, double r) {
      const char* names[1] = {asset};
      double s0a[1] = {s}, mua[1] = {r}, sga[1] = {sg};
      orc_model m{0, 1, names, s0a, mua, sga, nullptr, 0, 0, 0, 0, 0, 0};
      orc_cfg c = *cfg;
      c.rate = r;
      orc_result res;
      char e2[256];
      int rc = orc_price(json, &m, &c, nullptr, nullptr, nullptr, nullptr, 0, &res, e2, sizeof e2);
      if (rc) fail(rc, e2);
      return res.price;
    };
    double r = cfg->rate;
    double p = reprice(s0, sigma, r);
    double su = reprice(s0 + h_spot, sigma, r), sd = reprice(s0 - h_spot, sigma, r);
    out[0] = p;
    out[1] = (su - sd) / (2.0 * h_spot);
    out[2] = (su - 2.0 * p + sd) / (h_spot * h_spot);
    out[3] = (reprice(s0, sigma + h_vol, r) - reprice(s0, sigma - h_vol, r)) / (2.0 * h_vol);
    out[4] = (reprice(s0, sigma, r + h_rate) - reprice(s0, sigma, r - h_rate)) / (2.0 * h_rate);
    return 0;
  } catch (KErr& e) {
    return report(e, err, errcap);
  }
}

// greek_surface (src/surface.rs:104-138): greeks_gbm at every (spots[i],
// vols[j]), out[(i * n_vols + j) * 5 + q]; an empty axis is
// InconsistentPath; the lowest failing cell's error is returned.
int orc_greek_surface(const char* json, const char* asset, const double* spots, uint64_t n_spots, const double* vols,
                      uint64_t n_vols, const orc_cfg* cfg, double h_spot, double h_vol, double h_rate, double* out,
                      char* err, int errcap) {
  if (n_spots == 0 || n_vols == 0) return report(KErr{4, "empty surface grid"}, err, errcap);
  for (uint64_t c = 0; c < n_spots * n_vols; ++c) {
    int rc = orc_greeks_gbm(json, asset, spots[c / n_vols], vols[c % n_vols], cfg, h_spot, h_vol, h_rate,
                            out + 5 * c, err, errcap);
    if (rc) return rc;
  }
  return 0;
}

// Heston CRN bump-and-reprice, the same central-difference pattern as
// greeks_gbm applied to (spot, v0, kappa, theta, sigma_v, rho).
// h[6] absolute bumps; out[7] = {price, dP/dspot, dP/dv0, dP/dkappa,
// dP/dtheta, dP/dsigma_v, dP/drho}.
int orc_greeks_heston_bump(const char* json, const orc_model* m, const orc_cfg* cfg, const double* h, double* out,
                           char* err, int errcap) {
  try {
    auto reprice = [&](int which, double delta) {
      orc_model mm = *m;
      double s0a[1] = {m->s0[0]};
      mm.s0 = s0a;
      if (which == 0) s0a[0] += delta;
      if (which == 1) mm.v0 += delta;
      if (which == 2) mm.kappa += delta;
      if (which == 3) mm.theta += delta;
      if (which == 4) mm.sigma_v += delta;
      if (which == 5) mm.rho += delta;
      orc_result res;
      char e2[256];
      int rc = orc_price(json, &mm, cfg, nullptr, nullptr, nullptr, nullptr, 0, &res, e2, sizeof e2);
      if (rc) fail(rc, e2);
      return res.price;
    };
    out[0] = reprice(-1, 0.0);
    for (int j = 0; j < 6; ++j) out[1 + j] = (reprice(j, h[j]) - reprice(j, -h[j])) / (2.0 * h[j]);
    return 0;
  } catch (KErr& e) {
    return report(e, err, errcap);
  }
}

// price_american_lsm (src/lsm.rs:80-124) and backward_induction (:129-229).
int orc_price_american(const char* json, const double* dates, uint64_t n_dates, const orc_model* m,
                       const orc_cfg* cfg, int n_basis, double* pv_out, orc_result* res, char* err, int errcap) {
  try {
    if (n_dates == 0) fail(E_MALFORMED, "LSM: at least one exercise date");
    for (uint64_t i = 1; i < n_dates; ++i)
      if (dates[i] <= dates[i - 1]) fail(E_MALFORMED, "LSM: dates strictly increasing and > 0");
    if (dates[0] <= 0.0) fail(E_MALFORMED, "LSM: dates strictly increasing and > 0");
    if (n_basis < 2) fail(E_MALFORMED, "LSM: n_basis >= 2");
    Model md = make_model(m);
    validate_model(md);
    std::vector<double> grid{0.0};
    grid.insert(grid.end(), dates, dates + n_dates);
    validate_grid(grid);
    auto con = parse_contract(json);
    resolve_con(*con, md.names);
    size_t n = grid.size(), K = md.names.size(), D = (size_t)n_dates;
    uint64_t N = cfg->n_paths;
    double rate = cfg->rate;
    std::vector<std::vector<double>> payoff(D, std::vector<double>(N)), spots(D, std::vector<double>(N));
    Normals nz{cfg->seed, nullptr, n, md.m};
    parallel_paths(N, cfg->threads, [&](uint64_t b, uint64_t e) {
      std::vector<double> row(K * n);
      Flows f;
      for (uint64_t i = b; i < e; ++i) {
        simulate_path(md, grid, i, nz, row.data());
        PathView pv{grid.data(), n, row.data()};
        for (size_t d = 0; d < D; ++d) {
          size_t ti = d + 1;
          f.clear();
          cashflows(*con, ti, pv, f);
          double v = 0.0;
          for (auto& fl : f) v += fl.first * std::exp(-rate * (grid[fl.second] - grid[ti]));
          payoff[d][i] = v;
          spots[d][i] = row[ti];  // the model's first asset
        }
      }
    });
    std::vector<double> value = payoff[D - 1];
    for (size_t d = D - 1; d-- > 0;) {
      double disc = std::exp(-rate * (grid[d + 2] - grid[d + 1]));
      std::vector<double> y(N);
      for (uint64_t i = 0; i < N; ++i) y[i] = value[i] * disc;
      size_t mm = (size_t)n_basis;
      std::vector<std::vector<double>> ata(mm, std::vector<double>(mm, 0.0));
      std::vector<double> aty(mm, 0.0), pw(mm, 1.0);
      size_t n_itm = 0;
      for (uint64_t i = 0; i < N; ++i) {
        if (!(payoff[d][i] > 0.0)) continue;
        ++n_itm;
        pw[0] = 1.0;
        for (size_t j = 1; j < mm; ++j) pw[j] = pw[j - 1] * spots[d][i];
        for (size_t r = 0; r < mm; ++r) {
          aty[r] += pw[r] * y[i];
          for (size_t c = 0; c < mm; ++c) ata[r][c] += pw[r] * pw[c];
        }
      }
      bool have = n_itm >= mm;
      std::vector<double> coef;
      if (have) coef = solve_linear(ata, aty);
      for (uint64_t i = 0; i < N; ++i) {
        if (payoff[d][i] > 0.0) {
          double cont;
          if (have) {
            cont = 0.0;
            for (size_t j = mm; j-- > 0;) cont = cont * spots[d][i] + coef[j];
          } else
            cont = y[i];
          value[i] = payoff[d][i] > cont ? payoff[d][i] : y[i];
        } else
          value[i] = y[i];
      }
    }
    double disc0 = std::exp(-rate * grid[1]);
    for (auto& v : value) v *= disc0;
    if (pv_out) memcpy(pv_out, value.data(), N * sizeof(double));
    *res = summarize(value);
    return 0;
  } catch (KErr& e) {
    return report(e, err, errcap);
  }
}

// One LSM regression solve with the reference's normal equations + Gauss.
int orc_solve_linear(const double* a, const double* b, int n, double* x) {
  std::vector<std::vector<double>> A(n, std::vector<double>(n));
  std::vector<double> B(b, b + n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) A[i][j] = a[i * n + j];
  auto r = solve_linear(A, B);
  for (int i = 0; i < n; ++i) x[i] = r[i];
  return 0;
}

// Summarize an array of per-path values (two-pass, src/pricer.rs:303-328).
void orc_summarize(const double* pvs, uint64_t n, orc_result* res) {
  *res = summarize(std::vector<double>(pvs, pvs + n));
}

// ---------------------------------------------------------------------------
// Closed forms (test pins)
// ---------------------------------------------------------------------------
static double ncdf(double x) { return 0.5 * std::erfc(-x / std::sqrt(2.0)); }

double orc_bs_call(double s, double k, double t, double r, double sigma) {
  if (t <= 0.0 || sigma <= 0.0) return std::max(s - k * std::exp(-r * t), 0.0);
  double vst = sigma * std::sqrt(t);
  double d1 = (std::log(s / k) + (r + 0.5 * sigma * sigma) * t) / vst, d2 = d1 - vst;
  return s * ncdf(d1) - k * std::exp(-r * t) * ncdf(d2);
}
double orc_bs_put(double s, double k, double t, double r, double sigma) {
  if (t <= 0.0 || sigma <= 0.0) return std::max(k * std::exp(-r * t) - s, 0.0);
  double vst = sigma * std::sqrt(t);
  double d1 = (std::log(s / k) + (r + 0.5 * sigma * sigma) * t) / vst, d2 = d1 - vst;
  return k * std::exp(-r * t) * ncdf(-d2) - s * ncdf(-d1);
}
// out = {delta, gamma, vega, rho} of a call.
void orc_bs_call_greeks(double s, double k, double t, double r, double sigma, double* out) {
  double st = sigma * std::sqrt(t);
  double d1 = (std::log(s / k) + (r + 0.5 * sigma * sigma) * t) / st, d2 = d1 - st;
  double pdf = std::exp(-0.5 * d1 * d1) / std::sqrt(2.0 * M_PI);
  out[0] = ncdf(d1);
  out[1] = pdf / (s * st);
  out[2] = s * pdf * std::sqrt(t);
  out[3] = k * t * std::exp(-r * t) * ncdf(d2);
}

// Up-and-out call, K < H, continuous monitoring (Reiner-Rubinstein via Haug:
// A - B + C - D with phi = 1, eta = -1, cost of carry b = r).
double orc_uo_call(double s, double k, double h, double t, double r, double sigma) {
  if (s >= h) return 0.0;
  double b = r, st = sigma * std::sqrt(t);
  double mu = (b - 0.5 * sigma * sigma) / (sigma * sigma);
  double x1 = std::log(s / k) / st + (1 + mu) * st;
  double x2 = std::log(s / h) / st + (1 + mu) * st;
  double y1 = std::log(h * h / (s * k)) / st + (1 + mu) * st;
  double y2 = std::log(h / s) / st + (1 + mu) * st;
  double cf = std::exp((b - r) * t), df = std::exp(-r * t);
  double hs2 = std::pow(h / s, 2 * (mu + 1)), hs = std::pow(h / s, 2 * mu);
  double A = s * cf * ncdf(x1) - k * df * ncdf(x1 - st);
  double B = s * cf * ncdf(x2) - k * df * ncdf(x2 - st);
  double C = s * cf * hs2 * ncdf(-y1) - k * df * hs * ncdf(-y1 + st);
  double D = s * cf * hs2 * ncdf(-y2) - k * df * hs * ncdf(-y2 + st);
  return A - B + C - D;
}

// Heston (1993) call via the Albrecher et al. (2007) "little trap" form:
// C = S0 P1 - K e^{-rT} P2, P_j = 1/2 + 1/pi int_0^inf Re[e^{-iu lnK} f_j(u)/(iu)] du,
// integrated by composite Gauss-Legendre over [0, umax].
static std::complex<double> heston_cf(std::complex<double> u, double s0, double t, double r, double v0,
                                      double kappa, double theta, double sv, double rho) {
  using C = std::complex<double>;
  const C i(0.0, 1.0);
  C a = kappa - rho * sv * i * u;
  C d = std::sqrt(a * a + sv * sv * (i * u + u * u));
  C g = (a - d) / (a + d);
  C e = std::exp(-d * t);
  C A = i * u * (std::log(s0) + r * t);
  C B = (kappa * theta / (sv * sv)) * ((a - d) * t - 2.0 * std::log((1.0 - g * e) / (1.0 - g)));
  C Cc = (v0 / (sv * sv)) * (a - d) * (1.0 - e) / (1.0 - g * e);
  return std::exp(A + B + Cc);
}

double orc_heston_call(double s0, double k, double t, double r, double v0, double kappa, double theta, double sv,
                       double rho) {
  using C = std::complex<double>;
  const C i(0.0, 1.0);
  // 8-point Gauss-Legendre nodes/weights on [-1, 1].
  static const double xg[8] = {-0.9602898564975363, -0.7966664774136267, -0.5255324099163290,
                               -0.1834346424956498, 0.1834346424956498,  0.5255324099163290,
                               0.7966664774136267,  0.9602898564975363};
  static const double wg[8] = {0.1012285362903763, 0.2223810344533745, 0.3137066458778873, 0.3626837833783620,
                               0.3626837833783620, 0.3137066458778873, 0.2223810344533745, 0.1012285362903763};
  const double umax = 250.0;
  const int panels = 2500;
  const double hw = umax / panels;
  double lk = std::log(k);
  C phi_mi = heston_cf(C(0.0, -1.0), s0, t, r, v0, kappa, theta, sv, rho);
  double p1 = 0.0, p2 = 0.0;
  for (int pnl = 0; pnl < panels; ++pnl) {
    double a = pnl * hw, b = a + hw;
    for (int q = 0; q < 8; ++q) {
      double u = 0.5 * (a + b) + 0.5 * hw * xg[q];
      double w = 0.5 * hw * wg[q];
      C e = std::exp(-i * u * lk);
      C f2 = heston_cf(C(u, 0.0), s0, t, r, v0, kappa, theta, sv, rho);
      C f1 = heston_cf(C(u, -1.0), s0, t, r, v0, kappa, theta, sv, rho) / phi_mi;
      p2 += w * std::real(e * f2 / (i * u));
      p1 += w * std::real(e * f1 / (i * u));
    }
  }
  p1 = 0.5 + p1 / M_PI;
  p2 = 0.5 + p2 / M_PI;
  return s0 * p1 - k * std::exp(-r * t) * p2;
}

// Cox-Ross-Rubinstein American (is_call) / Bermudan on a binomial tree.
double orc_crr_american(double s0, double k, double t, double r, double sigma, int n, int is_call) {
  double dt = t / n, u = std::exp(sigma * std::sqrt(dt)), d = 1.0 / u, disc = std::exp(-r * dt);
  double p = (std::exp(r * dt) - d) / (u - d);
  auto pay = [&](double s) { return is_call ? std::max(s - k, 0.0) : std::max(k - s, 0.0); };
  std::vector<double> v(n + 1);
  for (int j = 0; j <= n; ++j) v[j] = pay(s0 * std::pow(u, j) * std::pow(d, n - j));
  for (int st = n - 1; st >= 0; --st)
    for (int j = 0; j <= st; ++j) {
      double s = s0 * std::pow(u, j) * std::pow(d, st - j);
      v[j] = std::max(disc * (p * v[j + 1] + (1 - p) * v[j]), pay(s));
    }
  return v[0];
}

}  // extern "C"
