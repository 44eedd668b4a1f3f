// objective.hpp -- scalar objective expressions with reverse-mode
// gradients, for C++ design loops on the engine.
//
// The reference builds its objectives from an expression-template AutoDiff
// library (AutoDiff.h:400-850: + - * /, pow, log, exp, sgm, abs on scalar
// nodes with value and diff) over the homogenized tensor entries
// (heat_tensor_t, homoExpression.h:68-150), e.g. runCustom's
// sum (C_ij - t_ij)^2 - 1e-4 (Framework.cu:168-170). Here an objective is a
// small graph of shared nodes; backward() accumulates d objective / d node
// into every node, and tensor_grad() gives the dC that otm_backward takes.
//
//   auto C = homo::ad::heat_tensor(Cval);              // 3x3 leaves
//   auto f = (C[0][0] - 0.3).pow(2) + (C[1][1] - 0.2).pow(2);
//   f = f.log() + 0.1 * C[0][1].abs();
//   f.backward();
//   double dC[9]; homo::ad::tensor_grad(C, dC);        // for otm_backward
#pragma once

#include <math.h>

#include <array>
#include <memory>
#include <unordered_set>
#include <utility>
#include <vector>

namespace homo {
namespace ad {

struct Node {
  double value = 0.0, grad = 0.0;
  std::vector<std::pair<std::shared_ptr<Node>, double>> parents;  // (operand, d this / d operand)
};

class Var {
 public:
  Var(double v = 0.0) : n_(std::make_shared<Node>()) { n_->value = v; }
  double value() const { return n_->value; }
  double diff() const { return n_->grad; }

  friend Var operator+(const Var& a, const Var& b) { return make(a.value() + b.value(), {{a, 1.0}, {b, 1.0}}); }
  friend Var operator-(const Var& a, const Var& b) { return make(a.value() - b.value(), {{a, 1.0}, {b, -1.0}}); }
  friend Var operator*(const Var& a, const Var& b) {
    return make(a.value() * b.value(), {{a, b.value()}, {b, a.value()}});
  }
  // div_exp_t (AutoDiff.h:717-727)
  friend Var operator/(const Var& a, const Var& b) {
    const double bv = b.value();
    return make(a.value() / bv, {{a, 1.0 / bv}, {b, -a.value() / (bv * bv)}});
  }
  Var operator-() const { return make(-value(), {{*this, -1.0}}); }
  // pow_exp_t (AutoDiff.h:738-748)
  Var pow(double e) const { return make(::pow(value(), e), {{*this, e * ::pow(value(), e - 1.0)}}); }
  // ln_exp_t (AutoDiff.h:782-793)
  Var log() const { return make(::log(value()), {{*this, 1.0 / value()}}); }
  // exp_exp_t (AutoDiff.h:771-780)
  Var exp() const {
    const double v = ::exp(value());
    return make(v, {{*this, v}});
  }
  // sgm_exp_t (AutoDiff.h:833-849): 1 / (1 + exp(-s (x - c)))
  Var sgm(double s = 1.0, double c = 0.0) const {
    const double v = 1.0 / (1.0 + ::exp(-s * (value() - c)));
    return make(v, {{*this, s * v * (1.0 - v)}});
  }
  // abs_exp_t (AutoDiff.h:751-768): derivative 0 at 0
  Var abs() const {
    const double x = value();
    return make(::fabs(x), {{*this, x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0)}});
  }

  // Reverse sweep over every node reachable from this one.
  double backward(double seed = 1.0) const {
    std::vector<Node*> order;
    std::unordered_set<Node*> seen;
    std::vector<std::pair<Node*, bool>> stack{{n_.get(), false}};
    while (!stack.empty()) {
      auto [node, done] = stack.back();
      stack.pop_back();
      if (done) {
        order.push_back(node);
        continue;
      }
      if (!seen.insert(node).second) continue;
      stack.push_back({node, true});
      for (auto& p : node->parents)
        if (!seen.count(p.first.get())) stack.push_back({p.first.get(), false});
    }
    for (Node* node : order) node->grad = 0.0;
    n_->grad = seed;
    for (auto it = order.rbegin(); it != order.rend(); ++it)
      for (auto& p : (*it)->parents) p.first->grad += (*it)->grad * p.second;
    return value();
  }

 private:
  std::shared_ptr<Node> n_;
  static Var make(double v, std::initializer_list<std::pair<Var, double>> ops) {
    Var r(v);
    for (auto& o : ops) r.n_->parents.push_back({o.first.n_, o.second});
    return r;
  }
};

inline Var operator+(const Var& a, double b) { return a + Var(b); }
inline Var operator+(double a, const Var& b) { return Var(a) + b; }
inline Var operator-(const Var& a, double b) { return a - Var(b); }
inline Var operator-(double a, const Var& b) { return Var(a) - b; }
inline Var operator*(const Var& a, double b) { return a * Var(b); }
inline Var operator*(double a, const Var& b) { return Var(a) * b; }
inline Var operator/(const Var& a, double b) { return a / Var(b); }
inline Var operator/(double a, const Var& b) { return Var(a) / b; }

using Tensor3 = std::array<std::array<Var, 3>, 3>;

// The homogenized tensor (row-major 9 doubles) as 3x3 leaf variables.
inline Tensor3 heat_tensor(const double C[9]) {
  Tensor3 t;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) t[i][j] = Var(C[3 * i + j]);
  return t;
}

// d objective / d C after backward(): the dC of otm_backward / Sensitivity.
inline void tensor_grad(const Tensor3& t, double dC[9]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) dC[3 * i + j] = t[i][j].diff();
}

}  // namespace ad
}  // namespace homo
