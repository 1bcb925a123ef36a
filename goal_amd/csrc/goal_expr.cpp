#include "goal_expr.hpp"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <stdexcept>

namespace goal {

struct Expr::Node {
  enum Kind { NUM, VAR, NEG, ADD, SUB, MUL, DIV, POW, CALL1, CALL2 } kind;
  double value = 0.0;
  int var = 0;
  double (*f1)(double) = nullptr;
  double (*f2)(double, double) = nullptr;
  std::unique_ptr<Node> a, b;
  double eval(const double* v) const {
    switch (kind) {
      case NUM: return value;
      case VAR: return v[var];
      case NEG: return -a->eval(v);
      case ADD: return a->eval(v) + b->eval(v);
      case SUB: return a->eval(v) - b->eval(v);
      case MUL: return a->eval(v) * b->eval(v);
      case DIV: return a->eval(v) / b->eval(v);
      case POW: return std::pow(a->eval(v), b->eval(v));
      case CALL1: return f1(a->eval(v));
      case CALL2: return f2(a->eval(v), b->eval(v));
    }
    return 0.0;
  }
};

namespace {

using NodeP = std::unique_ptr<Expr::Node>;

double f_abs(double x) { return std::fabs(x); }
double f_min(double a, double b) { return a < b ? a : b; }
double f_max(double a, double b) { return a > b ? a : b; }

struct Parser {
  const std::string& s;
  size_t i = 0;
  explicit Parser(const std::string& str) : s(str) {}

  [[noreturn]] void fail(const std::string& why) {
    throw std::invalid_argument("expression \"" + s + "\": " + why + " at offset " + std::to_string(i));
  }
  void skip() {
    while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
  }
  bool accept(char c) {
    skip();
    if (i < s.size() && s[i] == c) {
      ++i;
      return true;
    }
    return false;
  }
  static NodeP make(Expr::Node::Kind k, NodeP a = nullptr, NodeP b = nullptr) {
    NodeP n(new Expr::Node);
    n->kind = k;
    n->a = std::move(a);
    n->b = std::move(b);
    return n;
  }
  NodeP expr() {  // sum
    NodeP l = term();
    for (;;) {
      if (accept('+')) l = make(Expr::Node::ADD, std::move(l), term());
      else if (accept('-')) l = make(Expr::Node::SUB, std::move(l), term());
      else return l;
    }
  }
  NodeP term() {
    NodeP l = unary();
    for (;;) {
      if (accept('*')) l = make(Expr::Node::MUL, std::move(l), unary());
      else if (accept('/')) l = make(Expr::Node::DIV, std::move(l), unary());
      else return l;
    }
  }
  NodeP unary() {
    if (accept('-')) return make(Expr::Node::NEG, unary());
    if (accept('+')) return unary();
    return power();
  }
  NodeP power() {
    NodeP base = primary();
    if (accept('^')) return make(Expr::Node::POW, std::move(base), unary());
    return base;
  }
  NodeP primary() {
    skip();
    if (i >= s.size()) fail("unexpected end");
    char c = s[i];
    if (c == '(') {
      ++i;
      NodeP e = expr();
      if (!accept(')')) fail("expected ')'");
      return e;
    }
    if (std::isdigit((unsigned char)c) || c == '.') {
      const char* begin = s.c_str() + i;
      char* end = nullptr;
      double v = std::strtod(begin, &end);
      if (end == begin) fail("bad number");
      i += (size_t)(end - begin);
      NodeP n = make(Expr::Node::NUM);
      n->value = v;
      return n;
    }
    if (std::isalpha((unsigned char)c) || c == '_') {
      size_t b = i;
      while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '_')) ++i;
      std::string id = s.substr(b, i - b);
      if (id == "x" || id == "y" || id == "z" || id == "t") {
        NodeP n = make(Expr::Node::VAR);
        n->var = id == "x" ? 0 : id == "y" ? 1 : id == "z" ? 2 : 3;
        return n;
      }
      if (id == "pi") {
        NodeP n = make(Expr::Node::NUM);
        n->value = M_PI;
        return n;
      }
      double (*f1)(double) = nullptr;
      double (*f2)(double, double) = nullptr;
      if (id == "sin") f1 = ::sin;
      else if (id == "cos") f1 = ::cos;
      else if (id == "tan") f1 = ::tan;
      else if (id == "asin") f1 = ::asin;
      else if (id == "acos") f1 = ::acos;
      else if (id == "atan") f1 = ::atan;
      else if (id == "sinh") f1 = ::sinh;
      else if (id == "cosh") f1 = ::cosh;
      else if (id == "tanh") f1 = ::tanh;
      else if (id == "exp") f1 = ::exp;
      else if (id == "log") f1 = ::log;
      else if (id == "log10") f1 = ::log10;
      else if (id == "sqrt") f1 = ::sqrt;
      else if (id == "abs" || id == "fabs") f1 = f_abs;
      else if (id == "floor") f1 = ::floor;
      else if (id == "ceil") f1 = ::ceil;
      else if (id == "pow") f2 = ::pow;
      else if (id == "atan2") f2 = ::atan2;
      else if (id == "min") f2 = f_min;
      else if (id == "max") f2 = f_max;
      else fail("unknown identifier '" + id + "'");
      if (!accept('(')) fail("expected '(' after function name");
      NodeP a = expr();
      NodeP out;
      if (f2) {
        if (!accept(',')) fail("expected ','");
        NodeP b = expr();
        out = make(Expr::Node::CALL2, std::move(a), std::move(b));
        out->f2 = f2;
      } else {
        out = make(Expr::Node::CALL1, std::move(a));
        out->f1 = f1;
      }
      if (!accept(')')) fail("expected ')'");
      return out;
    }
    fail(std::string("unexpected character '") + c + "'");
  }
};

}  // namespace

Expr::Expr(const std::string& text) {
  Parser p(text);
  root_ = p.expr();
  p.skip();
  if (p.i != text.size()) p.fail("trailing characters");
}

Expr::~Expr() = default;

double Expr::operator()(double x, double y, double z, double t) const {
  const double v[4] = {x, y, z, t};
  return root_->eval(v);
}

double eval(const std::string& v, double x, double y, double z, double t) {
  return Expr(v)(x, y, z, t);
}

}  // namespace goal
