// FunctionInit.hpp -- initial condition given as one expression string per
// component in the variables x, y, z (init/FunctionInit.hpp:34-101, which
// hands the strings to amrex::Parser). Initial data are evaluated on the host
// at cell centres, so the expressions are parsed here into a small tree and
// evaluated in double precision.
//
// Grammar (precedence low to high; AMReX parser conventions):
//   expr   := term (('+' | '-') term)*
//   term   := unary (('*' | '/') unary)*
//   unary  := ('-' | '+') unary | power
//   power  := atom (('^' | '**') unary)?          right-associative, binds tighter than unary minus
//   atom   := number | x | y | z | pi | name '(' expr (',' expr)* ')' | '(' expr ')'
// Functions: sqrt exp log log10 sin cos tan asin acos atan atan2 sinh cosh tanh
// abs fabs floor ceil pow min max heaviside. The functor itself only holds
// pointers into a process-wide registry of parsed expressions, so it stays
// trivially copyable as InitConcept requires.
//
// The integrators evaluate the initial data on the device: program() compiles
// every component's expression into a postfix program (tsp_expr_op, codes of
// treesap_hip.h; children in order, then the node), which tsp_init_expr runs
// at every cell centre. + - * / sqrt floor ceil abs min max heaviside round as
// on the host; the other functions use the device's libm, which may differ
// from the host's in the last bits. operator()<c>(x, y, z) evaluates on the host.
#pragma once

#include "../pde-model/PDEModelConcepts.hpp"
#include "treesap_hip.h"

#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <deque>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace treesap {

namespace detail {

class Expression {
  public:
    explicit Expression(const std::string &text) : s_(text) {
        root_ = parse_expr();
        skip();
        if (pos_ != s_.size()) fail("unexpected '" + std::string(1, s_[pos_]) + "'");
    }
    double eval(double x, double y, double z) const { return eval(root_, x, y, z); }
    // postfix program of the expression, appended to `out`
    void emit(std::vector<tsp_expr_op> &out) const { emit(root_, out); }

  private:
    enum Op { NUM, VX, VY, VZ, ADD, SUB, MUL, DIV, NEG, POW, FN1, FN2 };
    struct Node {
        Op op;
        double value = 0.0;
        int a = -1, b = -1;
        double (*f1)(double) = nullptr;
        double (*f2)(double, double) = nullptr;
        int code = -1; // tsp_expr_code of a function node
    };

    int add(Node n) {
        nodes_.push_back(n);
        return (int)nodes_.size() - 1;
    }
    [[noreturn]] void fail(const std::string &why) const {
        throw std::invalid_argument("treesap::FunctionInit: " + why + " in \"" + s_ + "\" at " + std::to_string(pos_));
    }
    void skip() {
        while (pos_ < s_.size() && std::isspace((unsigned char)s_[pos_])) ++pos_;
    }
    bool eat(char c) {
        skip();
        if (pos_ < s_.size() && s_[pos_] == c) {
            ++pos_;
            return true;
        }
        return false;
    }

    int parse_expr() {
        int l = parse_term();
        for (;;) {
            if (eat('+')) l = add({ADD, 0.0, l, parse_term()});
            else if (eat('-')) l = add({SUB, 0.0, l, parse_term()});
            else return l;
        }
    }
    int parse_term() {
        int l = parse_unary();
        for (;;) {
            skip();
            if (pos_ + 1 < s_.size() && s_[pos_] == '*' && s_[pos_ + 1] == '*') return l; // power handled below
            if (eat('*')) l = add({MUL, 0.0, l, parse_unary()});
            else if (eat('/')) l = add({DIV, 0.0, l, parse_unary()});
            else return l;
        }
    }
    int parse_unary() {
        if (eat('-')) return add({NEG, 0.0, parse_unary()});
        if (eat('+')) return parse_unary();
        return parse_power();
    }
    int parse_power() {
        const int base = parse_atom();
        skip();
        if (pos_ + 1 < s_.size() && s_[pos_] == '*' && s_[pos_ + 1] == '*') {
            pos_ += 2;
            return add({POW, 0.0, base, parse_unary()});
        }
        if (eat('^')) return add({POW, 0.0, base, parse_unary()});
        return base;
    }
    int parse_atom() {
        skip();
        if (pos_ >= s_.size()) fail("unexpected end");
        const char c = s_[pos_];
        if (std::isdigit((unsigned char)c) || c == '.') {
            const char *begin = s_.c_str() + pos_;
            char *end = nullptr;
            const double v = std::strtod(begin, &end);
            if (end == begin) fail("bad number");
            pos_ += (std::size_t)(end - begin);
            return add({NUM, v});
        }
        if (std::isalpha((unsigned char)c) || c == '_') {
            std::size_t st = pos_;
            while (pos_ < s_.size() && (std::isalnum((unsigned char)s_[pos_]) || s_[pos_] == '_')) ++pos_;
            const std::string name = s_.substr(st, pos_ - st);
            if (!eat('(')) {
                if (name == "x") return add({VX});
                if (name == "y") return add({VY});
                if (name == "z") return add({VZ});
                if (name == "pi") return add({NUM, 3.14159265358979323846});
                fail("unknown variable '" + name + "'");
            }
            std::vector<int> args{parse_expr()};
            while (eat(',')) args.push_back(parse_expr());
            if (!eat(')')) fail("missing ')'");
            return call(name, args);
        }
        if (eat('(')) {
            const int e = parse_expr();
            if (!eat(')')) fail("missing ')'");
            return e;
        }
        fail("unexpected '" + std::string(1, c) + "'");
    }
    int call(const std::string &name, const std::vector<int> &args) {
        struct F1 {
            const char *n;
            double (*f)(double);
            int code;
        };
        static const F1 one[] = {{"sqrt", [](double v) { return std::sqrt(v); }, TSP_EXPR_SQRT},
                                 {"exp", [](double v) { return std::exp(v); }, TSP_EXPR_EXP},
                                 {"log", [](double v) { return std::log(v); }, TSP_EXPR_LOG},
                                 {"log10", [](double v) { return std::log10(v); }, TSP_EXPR_LOG10},
                                 {"sin", [](double v) { return std::sin(v); }, TSP_EXPR_SIN},
                                 {"cos", [](double v) { return std::cos(v); }, TSP_EXPR_COS},
                                 {"tan", [](double v) { return std::tan(v); }, TSP_EXPR_TAN},
                                 {"asin", [](double v) { return std::asin(v); }, TSP_EXPR_ASIN},
                                 {"acos", [](double v) { return std::acos(v); }, TSP_EXPR_ACOS},
                                 {"atan", [](double v) { return std::atan(v); }, TSP_EXPR_ATAN},
                                 {"sinh", [](double v) { return std::sinh(v); }, TSP_EXPR_SINH},
                                 {"cosh", [](double v) { return std::cosh(v); }, TSP_EXPR_COSH},
                                 {"tanh", [](double v) { return std::tanh(v); }, TSP_EXPR_TANH},
                                 {"abs", [](double v) { return std::fabs(v); }, TSP_EXPR_ABS},
                                 {"fabs", [](double v) { return std::fabs(v); }, TSP_EXPR_ABS},
                                 {"floor", [](double v) { return std::floor(v); }, TSP_EXPR_FLOOR},
                                 {"ceil", [](double v) { return std::ceil(v); }, TSP_EXPR_CEIL},
                                 {"heaviside", [](double v) { return v > 0.0 ? 1.0 : 0.0; }, TSP_EXPR_HEAVISIDE}};
        struct F2 {
            const char *n;
            double (*f)(double, double);
            int code;
        };
        static const F2 two[] = {{"pow", [](double a, double b) { return std::pow(a, b); }, TSP_EXPR_POW},
                                 {"atan2", [](double a, double b) { return std::atan2(a, b); }, TSP_EXPR_ATAN2},
                                 {"min", [](double a, double b) { return std::fmin(a, b); }, TSP_EXPR_MIN},
                                 {"max", [](double a, double b) { return std::fmax(a, b); }, TSP_EXPR_MAX}};
        for (const auto &f : one)
            if (name == f.n) {
                if (args.size() != 1) fail(name + " takes one argument");
                Node n{FN1, 0.0, args[0]};
                n.f1 = f.f;
                n.code = f.code;
                return add(n);
            }
        for (const auto &f : two)
            if (name == f.n) {
                if (args.size() != 2) fail(name + " takes two arguments");
                Node n{FN2, 0.0, args[0], args[1]};
                n.f2 = f.f;
                n.code = f.code;
                return add(n);
            }
        fail("unknown function '" + name + "'");
    }

    double eval(int i, double x, double y, double z) const {
        const Node &n = nodes_[i];
        switch (n.op) {
        case NUM: return n.value;
        case VX: return x;
        case VY: return y;
        case VZ: return z;
        case ADD: return eval(n.a, x, y, z) + eval(n.b, x, y, z);
        case SUB: return eval(n.a, x, y, z) - eval(n.b, x, y, z);
        case MUL: return eval(n.a, x, y, z) * eval(n.b, x, y, z);
        case DIV: return eval(n.a, x, y, z) / eval(n.b, x, y, z);
        case NEG: return -eval(n.a, x, y, z);
        case POW: return std::pow(eval(n.a, x, y, z), eval(n.b, x, y, z));
        case FN1: return n.f1(eval(n.a, x, y, z));
        case FN2: return n.f2(eval(n.a, x, y, z), eval(n.b, x, y, z));
        }
        return 0.0;
    }

    void emit(int i, std::vector<tsp_expr_op> &out) const {
        const Node &n = nodes_[i];
        static constexpr int leaf[] = {TSP_EXPR_NUM, TSP_EXPR_X, TSP_EXPR_Y, TSP_EXPR_Z};
        static constexpr int binary[] = {TSP_EXPR_ADD, TSP_EXPR_SUB, TSP_EXPR_MUL, TSP_EXPR_DIV};
        switch (n.op) {
        case NUM:
        case VX:
        case VY:
        case VZ: out.push_back({leaf[n.op - NUM], 0, n.op == NUM ? n.value : 0.0}); return;
        case ADD:
        case SUB:
        case MUL:
        case DIV:
            emit(n.a, out);
            emit(n.b, out);
            out.push_back({binary[n.op - ADD], 0, 0.0});
            return;
        case NEG:
            emit(n.a, out);
            out.push_back({TSP_EXPR_NEG, 0, 0.0});
            return;
        case POW:
            emit(n.a, out);
            emit(n.b, out);
            out.push_back({TSP_EXPR_POW, 0, 0.0});
            return;
        case FN1:
            emit(n.a, out);
            out.push_back({n.code, 0, 0.0});
            return;
        case FN2:
            emit(n.a, out);
            emit(n.b, out);
            out.push_back({n.code, 0, 0.0});
            return;
        }
    }

    std::string s_;
    std::size_t pos_ = 0;
    std::vector<Node> nodes_;
    int root_ = -1;
};

// Parsed expressions live here for the rest of the process, like the
// reference's parser registry (init/FunctionInit.hpp:95-98); not thread-safe.
inline std::deque<Expression> &expression_registry() {
    static std::deque<Expression> r;
    return r;
}

} // namespace detail

template <int ncomp_> struct FunctionInit {
    static_assert(ncomp_ > 0, "FunctionInit needs at least one component");
    static constexpr int ncomp = ncomp_;

    FunctionInit() = delete;
    explicit FunctionInit(const amrex::Array<std::string, ncomp> &exprs) {
        for (int c = 0; c < ncomp; ++c) {
            detail::expression_registry().emplace_back(exprs[c]);
            expr_[c] = &detail::expression_registry().back();
        }
    }

    template <int comp> amrex::Real operator()(amrex::Real x, amrex::Real y, amrex::Real z) const {
        return expr_[comp]->eval(x, y, z);
    }

    // every component's postfix program, concatenated; component c is
    // ops[start[c] .. start[c + 1]) (the layout tsp_init_expr reads)
    void program(std::vector<tsp_expr_op> &ops, std::vector<std::int32_t> &start) const {
        ops.clear();
        start.assign(1, 0);
        for (int c = 0; c < ncomp; ++c) {
            expr_[c]->emit(ops);
            start.push_back((std::int32_t)ops.size());
        }
    }

  private:
    const detail::Expression *expr_[ncomp]{};
};

static_assert(InitConcept<FunctionInit<1>>);

template <class T> struct is_function_init : std::false_type {};
template <int n> struct is_function_init<FunctionInit<n>> : std::true_type {};

} // namespace treesap
