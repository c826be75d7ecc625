#include "expr.hpp"

#include <cctype>
#include <cmath>
#include <cstdlib>

#include "common.hpp"

namespace mlr {

namespace {

struct Node {
    enum Op { Const, Var, Neg, Not, Add, Sub, Mul, Div, Pow, Lt, Le, Gt, Ge, Eq, Ne, And, Or, Assign, Fn1, Fn2 };
    Op op;
    double c = 0.0;
    int slot = -1;  // variable slot, or the function id of Fn1 / Fn2
    std::unique_ptr<Node> a, b;
};

// exprtk's common functions of one and two arguments
const char* const kFn1[] = {"sin", "cos", "tan", "exp", "log", "sqrt", "abs", "tanh", "floor", "ceil"};
const char* const kFn2[] = {"min", "max", "pow"};

using P = std::unique_ptr<Node>;

P mk(Node::Op op, P a = nullptr, P b = nullptr) {
    P n(new Node{op});
    n->a = std::move(a);
    n->b = std::move(b);
    return n;
}

}  // namespace

struct Expression::Impl {
    std::vector<std::string> names{"x", "y"};
    std::vector<P> stmts;

    // ---- lexer / recursive descent parser
    std::string s;
    size_t p = 0;
    explicit Impl(const std::string& src) : s(src) {}

    [[noreturn]] void err(const std::string& m) {
        fail("expression error at offset " + std::to_string(p) + ": " + m + " in \"" + s + "\"");
    }
    void ws() {
        while (p < s.size() && std::isspace((unsigned char)s[p])) ++p;
    }
    bool eat(const char* tok) {
        ws();
        size_t n = 0;
        while (tok[n]) ++n;
        if (s.compare(p, n, tok) == 0) {
            // keywords must not run into an identifier
            if (std::isalpha((unsigned char)tok[0]) && p + n < s.size() &&
                (std::isalnum((unsigned char)s[p + n]) || s[p + n] == '_'))
                return false;
            p += n;
            return true;
        }
        return false;
    }
    std::string ident() {
        ws();
        size_t q = p;
        if (q < s.size() && (std::isalpha((unsigned char)s[q]) || s[q] == '_')) {
            ++q;
            while (q < s.size() && (std::isalnum((unsigned char)s[q]) || s[q] == '_')) ++q;
        }
        std::string id = s.substr(p, q - p);
        p = q;
        return id;
    }
    int slot_of(const std::string& n, bool create) {
        for (size_t k = 0; k < names.size(); ++k)
            if (names[k] == n) return (int)k;
        if (!create) err("unknown variable '" + n + "'");
        names.push_back(n);
        return (int)names.size() - 1;
    }

    P primary() {
        ws();
        if (p >= s.size()) err("unexpected end");
        const char c = s[p];
        if (c == '(') {
            ++p;
            P e = logical_or();
            if (!eat(")")) err("expected ')'");
            return e;
        }
        if (std::isdigit((unsigned char)c) || c == '.') {
            const char* b = s.c_str() + p;
            char* e = nullptr;
            const double v = std::strtod(b, &e);  // correctly rounded decimal parse
            if (e == b) err("bad number");
            p += (size_t)(e - b);
            P n = mk(Node::Const);
            n->c = v;
            return n;
        }
        if (std::isalpha((unsigned char)c) || c == '_') {
            std::string id = ident();
            ws();
            if (p < s.size() && s[p] == '(') {
                for (int f = 0; f < (int)(sizeof(kFn1) / sizeof(kFn1[0])); ++f)
                    if (id == kFn1[f]) {
                        ++p;
                        P n = mk(Node::Fn1, logical_or());
                        n->slot = f;
                        if (!eat(")")) err("expected ')'");
                        return n;
                    }
                for (int f = 0; f < (int)(sizeof(kFn2) / sizeof(kFn2[0])); ++f)
                    if (id == kFn2[f]) {
                        ++p;
                        P a = logical_or();
                        if (!eat(",")) err("expected ','");
                        P n = mk(Node::Fn2, std::move(a), logical_or());
                        n->slot = f;
                        if (!eat(")")) err("expected ')'");
                        return n;
                    }
                err("unknown function '" + id + "'");
            }
            if (id == "pi") { P n = mk(Node::Const); n->c = M_PI; return n; }
            if (id == "e" && slot_of_opt(id) < 0) { P n = mk(Node::Const); n->c = M_E; return n; }
            if (id == "true") { P n = mk(Node::Const); n->c = 1.0; return n; }
            if (id == "false") { P n = mk(Node::Const); n->c = 0.0; return n; }
            P n = mk(Node::Var);
            n->slot = slot_of(id, false);
            return n;
        }
        err(std::string("unexpected '") + c + "'");
    }
    int slot_of_opt(const std::string& n) {
        for (size_t k = 0; k < names.size(); ++k)
            if (names[k] == n) return (int)k;
        return -1;
    }
    P unary() {
        if (eat("-")) return mk(Node::Neg, unary());
        if (eat("+")) return unary();
        if (eat("!") || eat("not")) return mk(Node::Not, unary());
        return power();
    }
    P power() {
        P a = primary();
        if (eat("^")) return mk(Node::Pow, std::move(a), unary());
        return a;
    }
    P mul() {
        P a = unary();
        for (;;) {
            if (eat("*")) a = mk(Node::Mul, std::move(a), unary());
            else if (eat("/")) a = mk(Node::Div, std::move(a), unary());
            else return a;
        }
    }
    P add() {
        P a = mul();
        for (;;) {
            if (eat("+")) a = mk(Node::Add, std::move(a), mul());
            else if (eat("-")) a = mk(Node::Sub, std::move(a), mul());
            else return a;
        }
    }
    P cmp() {
        P a = add();
        for (;;) {
            if (eat("<=")) a = mk(Node::Le, std::move(a), add());
            else if (eat(">=")) a = mk(Node::Ge, std::move(a), add());
            else if (eat("==")) a = mk(Node::Eq, std::move(a), add());
            else if (eat("!=") || eat("<>")) a = mk(Node::Ne, std::move(a), add());
            else if (eat("<")) a = mk(Node::Lt, std::move(a), add());
            else if (eat(">")) a = mk(Node::Gt, std::move(a), add());
            else if (eat("=")) a = mk(Node::Eq, std::move(a), add());
            else return a;
        }
    }
    P logical_and() {
        P a = cmp();
        while (eat("&&") || eat("and")) a = mk(Node::And, std::move(a), cmp());
        return a;
    }
    P logical_or() {
        P a = logical_and();
        while (eat("||") || eat("or")) a = mk(Node::Or, std::move(a), logical_and());
        return a;
    }
    void parse() {
        for (;;) {
            ws();
            if (p >= s.size()) break;
            if (eat(";")) continue;
            if (eat("var")) {
                std::string n = ident();
                if (n.empty()) err("expected a variable name");
                if (!eat(":=")) err("expected ':='");
                P e = logical_or();
                P st = mk(Node::Assign, std::move(e));
                st->slot = slot_of(n, true);
                stmts.push_back(std::move(st));
            } else {
                stmts.push_back(logical_or());
            }
            ws();
            if (p < s.size() && !eat(";")) err("expected ';'");
        }
        if (stmts.empty()) err("empty expression");
    }

    static double ev(const Node* n, double* vars) {
        switch (n->op) {
        case Node::Const: return n->c;
        case Node::Var: return vars[n->slot];
        case Node::Neg: return -ev(n->a.get(), vars);
        case Node::Not: return ev(n->a.get(), vars) == 0.0 ? 1.0 : 0.0;
        case Node::Add: return ev(n->a.get(), vars) + ev(n->b.get(), vars);
        case Node::Sub: return ev(n->a.get(), vars) - ev(n->b.get(), vars);
        case Node::Mul: return ev(n->a.get(), vars) * ev(n->b.get(), vars);
        case Node::Div: return ev(n->a.get(), vars) / ev(n->b.get(), vars);
        case Node::Pow: return std::pow(ev(n->a.get(), vars), ev(n->b.get(), vars));
        case Node::Lt: return ev(n->a.get(), vars) < ev(n->b.get(), vars) ? 1.0 : 0.0;
        case Node::Le: return ev(n->a.get(), vars) <= ev(n->b.get(), vars) ? 1.0 : 0.0;
        case Node::Gt: return ev(n->a.get(), vars) > ev(n->b.get(), vars) ? 1.0 : 0.0;
        case Node::Ge: return ev(n->a.get(), vars) >= ev(n->b.get(), vars) ? 1.0 : 0.0;
        case Node::Eq: return ev(n->a.get(), vars) == ev(n->b.get(), vars) ? 1.0 : 0.0;
        case Node::Ne: return ev(n->a.get(), vars) != ev(n->b.get(), vars) ? 1.0 : 0.0;
        case Node::And: return (ev(n->a.get(), vars) != 0.0 && ev(n->b.get(), vars) != 0.0) ? 1.0 : 0.0;
        case Node::Or: return (ev(n->a.get(), vars) != 0.0 || ev(n->b.get(), vars) != 0.0) ? 1.0 : 0.0;
        case Node::Assign: return vars[n->slot] = ev(n->a.get(), vars);
        case Node::Fn1: {
            const double v = ev(n->a.get(), vars);
            switch (n->slot) {
            case 0: return std::sin(v);
            case 1: return std::cos(v);
            case 2: return std::tan(v);
            case 3: return std::exp(v);
            case 4: return std::log(v);
            case 5: return std::sqrt(v);
            case 6: return std::fabs(v);
            case 7: return std::tanh(v);
            case 8: return std::floor(v);
            default: return std::ceil(v);
            }
        }
        case Node::Fn2: {
            const double u = ev(n->a.get(), vars), v = ev(n->b.get(), vars);
            return n->slot == 0 ? std::fmin(u, v) : (n->slot == 1 ? std::fmax(u, v) : std::pow(u, v));
        }
        }
        return 0.0;
    }
};

Expression::Expression(const std::string& src) : impl_(new Impl(src)) {
    impl_->parse();
}
Expression::~Expression() = default;
Expression::Expression(Expression&&) noexcept = default;
Expression& Expression::operator=(Expression&&) noexcept = default;

double Expression::eval(double x, double y) const {
    std::vector<double> vars(impl_->names.size(), 0.0);
    vars[0] = x;
    vars[1] = y;
    double r = 0.0;
    for (auto& st : impl_->stmts) r = Impl::ev(st.get(), vars.data());
    return r;
}

}  // namespace mlr
