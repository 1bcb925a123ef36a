// Expression evaluation for boundary-condition strings such as "0.01*t".
//
// Stands in for goal::eval (src/goal_control.cpp:107-123), which hands the
// string to a run-time C expression compiler with the variables x, y, z, t.
// Supported: numbers, x y z t, + - * / ^ (power, right associative), unary
// +/-, parentheses, and sin cos tan asin acos atan sinh cosh tanh exp log
// log10 sqrt abs fabs floor ceil, pow(a,b), atan2(a,b), min(a,b), max(a,b).
#pragma once

#include <memory>
#include <string>
#include <vector>

namespace goal {

class Expr {
 public:
  // Parses `text`; throws std::invalid_argument with a message on error.
  explicit Expr(const std::string& text);
  ~Expr();
  double operator()(double x, double y, double z, double t) const;

  struct Node;

 private:
  std::unique_ptr<Node> root_;
};

double eval(const std::string& v, double x, double y, double z, double t);

}  // namespace goal
