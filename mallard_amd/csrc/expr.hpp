// Expression evaluator for the analytical initial state
// (solver_initialize.cpp:101-186). It covers the exprtk subset Mallard
// inputs use: ';'-separated statements, `var name := expr`, arithmetic
// + - * / ^, comparisons returning 1/0, and/or/not (&&, ||, !), unary minus,
// parentheses, the functions sin cos tan exp log sqrt abs tanh floor ceil
// min max pow, the variables x and y, and the constants pi and e. The value
// of the last statement is the result.
#pragma once

#include <memory>
#include <string>
#include <vector>

namespace mlr {

class Expression {
public:
    explicit Expression(const std::string& src);
    ~Expression();
    Expression(Expression&&) noexcept;
    Expression& operator=(Expression&&) noexcept;
    double eval(double x, double y) const;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace mlr
