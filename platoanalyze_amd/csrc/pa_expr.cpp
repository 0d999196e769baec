// Scalar expressions of x, y, z and t: the 'Function' strings of the input
// decks.  The reference hands them to Omega_h's expression reader (body loads
// at the cubature points, src/PlatoMeshExpr.hpp:16-50) and to Teuchos'
// MathExpr (Dirichlet values as functions of time, src/PlatoMathExpr.hpp,
// src/EssentialBC.hpp:52-100).  This reader takes the common ground of the
// two: numbers, the variables x y z t, pi, + - * / ^ (right associative,
// binding tighter than a leading sign), parentheses, the functions sin cos
// tan asin acos atan sinh cosh tanh exp log log10 sqrt abs fabs (one
// argument), pow atan2 min max (two), and leading assignments
// 'name = expr;' whose names later parts may use.  A string is compiled once
// to a postfix program, which evaluates without allocation.
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "pa_internal.h"

namespace pa {

namespace {

enum Op : int {
    kConst, kVar, kSlot, kAdd, kSub, kMul, kDiv, kPow, kNeg,
    kSin, kCos, kTan, kAsin, kAcos, kAtan, kSinh, kCosh, kTanh, kExp, kLog, kLog10, kSqrt, kAbs,
    kPow2, kAtan2, kMin, kMax
};

struct Fn {
    const char* name;
    int op, args;
};
constexpr Fn kFns[] = {{"sin", kSin, 1},    {"cos", kCos, 1},     {"tan", kTan, 1},     {"asin", kAsin, 1},
                       {"acos", kAcos, 1},  {"atan", kAtan, 1},   {"sinh", kSinh, 1},   {"cosh", kCosh, 1},
                       {"tanh", kTanh, 1},  {"exp", kExp, 1},     {"log", kLog, 1},     {"log10", kLog10, 1},
                       {"sqrt", kSqrt, 1},  {"abs", kAbs, 1},     {"fabs", kAbs, 1},    {"pow", kPow2, 2},
                       {"atan2", kAtan2, 2}, {"min", kMin, 2},    {"max", kMax, 2}};

class Parser {
  public:
    Parser(const std::string& s, std::map<std::string, int>& names, std::vector<Expr::Inst>& out)
        : s_(s), names_(names), out_(out) {}

    void expr() {
        term();
        for (;;) {
            skip();
            if (peek('+') || peek('-')) {
                const char c = s_[p_++];
                term();
                emit(c == '+' ? kAdd : kSub);
            } else {
                return;
            }
        }
    }
    size_t pos() const { return p_; }
    void set_pos(size_t p) { p_ = p; }
    void skip() {
        while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) ++p_;
    }
    bool peek(char c) { return p_ < s_.size() && s_[p_] == c; }
    bool ident(std::string& id) {
        skip();
        if (p_ >= s_.size() || !(std::isalpha((unsigned char)s_[p_]) || s_[p_] == '_')) return false;
        const size_t b = p_;
        while (p_ < s_.size() && (std::isalnum((unsigned char)s_[p_]) || s_[p_] == '_')) ++p_;
        id = s_.substr(b, p_ - b);
        return true;
    }
    [[noreturn]] void fail(const std::string& what) const {
        throw Error(PA_ERR_ARG, "function '" + s_ + "': " + what + " at position " + std::to_string(p_));
    }

  private:
    void emit(int op, double v = 0.0, int i = 0) { out_.push_back(Expr::Inst{op, i, v}); }

    void term() {
        unary();
        for (;;) {
            skip();
            if (peek('*') || peek('/')) {
                const char c = s_[p_++];
                unary();
                emit(c == '*' ? kMul : kDiv);
            } else {
                return;
            }
        }
    }
    void unary() {
        skip();
        if (peek('-')) {
            ++p_;
            unary();
            emit(kNeg);
        } else if (peek('+')) {
            ++p_;
            unary();
        } else {
            power();
        }
    }
    void power() {
        primary();
        skip();
        if (peek('^')) {
            ++p_;
            unary();  // right associative: a^b^c = a^(b^c); a^-b allowed
            emit(kPow);
        }
    }
    void primary() {
        skip();
        if (p_ >= s_.size()) fail("unexpected end");
        const char c = s_[p_];
        if (std::isdigit((unsigned char)c) || c == '.') {
            const char* b = s_.c_str() + p_;
            char* e = nullptr;
            const double v = std::strtod(b, &e);
            if (e == b) fail("bad number");
            p_ += (size_t)(e - b);
            emit(kConst, v);
            return;
        }
        if (c == '(') {
            ++p_;
            expr();
            skip();
            if (!peek(')')) fail("')' expected");
            ++p_;
            return;
        }
        std::string id;
        if (!ident(id)) fail(std::string("unexpected '") + c + "'");
        skip();
        if (peek('(')) {
            ++p_;
            const Fn* f = nullptr;
            for (const Fn& k : kFns)
                if (id == k.name) f = &k;
            if (!f) fail("unknown function '" + id + "'");
            expr();
            for (int a = 1; a < f->args; ++a) {
                skip();
                if (!peek(',')) fail("',' expected in " + id);
                ++p_;
                expr();
            }
            skip();
            if (!peek(')')) fail("')' expected after the arguments of " + id);
            ++p_;
            emit(f->op);
            return;
        }
        if (id == "x" || id == "y" || id == "z" || id == "t") {
            emit(kVar, 0.0, id == "x" ? 0 : id == "y" ? 1 : id == "z" ? 2 : 3);
        } else if (names_.count(id)) {
            emit(kSlot, 0.0, names_[id]);
        } else if (id == "pi") {
            emit(kConst, 3.14159265358979323846);
        } else {
            fail("unknown name '" + id + "'");
        }
    }

    const std::string& s_;
    std::map<std::string, int>& names_;
    std::vector<Expr::Inst>& out_;
    size_t p_ = 0;
};

}  // namespace

Expr::Expr(const std::string& text) : text_(text) {
    std::map<std::string, int> names;
    // statements separated by ';': leading ones are 'name = expr'
    std::vector<std::string> parts;
    size_t b = 0;
    for (size_t i = 0; i <= text.size(); ++i)
        if (i == text.size() || text[i] == ';') {
            parts.push_back(text.substr(b, i - b));
            b = i + 1;
        }
    while (!parts.empty() && parts.back().find_first_not_of(" \t\r\n") == std::string::npos) parts.pop_back();
    if (parts.empty()) throw Error(PA_ERR_ARG, "function '" + text + "': empty");
    for (size_t k = 0; k < parts.size(); ++k) {
        const std::string& s = parts[k];
        std::vector<Inst> code;
        Parser P(s, names, code);
        std::string id;
        int slot = -1;
        const bool last = k + 1 == parts.size();
        P.skip();
        const size_t p0 = P.pos();
        if (P.ident(id)) {
            P.skip();
            if (P.peek('=')) {
                if (id == "x" || id == "y" || id == "z" || id == "t" || id == "pi")
                    P.fail("cannot assign to '" + id + "'");
                P.set_pos(P.pos() + 1);
                slot = (int)names.size();
                if (names.count(id)) slot = names[id];
            } else {
                P.set_pos(p0);
            }
        } else {
            P.set_pos(p0);
        }
        if (slot < 0 && !last) P.fail("only the last part may be an expression without 'name ='");
        P.expr();
        P.skip();
        if (P.pos() != s.size()) P.fail("unexpected text");
        if (slot >= 0) {
            code.push_back(Inst{kSlot, -1 - slot, 0.0});  // store marker
            names[(std::string)id] = slot;
        }
        for (const Inst& i : code) prog_.push_back(i);
        if (slot >= 0 && last) prog_.push_back(Inst{kSlot, slot, 0.0});  // the value of the last assignment
    }
    nslots_ = (int)names.size();
    // stack depth
    int d = 0, dmax = 0;
    for (const Inst& i : prog_) {
        switch (i.op) {
            case kConst: case kVar: ++d; break;
            case kSlot: d += i.i >= 0 ? 1 : -1; break;
            case kAdd: case kSub: case kMul: case kDiv: case kPow: case kPow2: case kAtan2: case kMin: case kMax: --d; break;
            default: break;
        }
        dmax = std::max(dmax, d);
    }
    depth_ = dmax;
}

double Expr::eval(double x, double y, double z, double t) const {
    double st[64], slots[64];
    if (depth_ > 64 || nslots_ > 64) throw Error(PA_ERR_LIMIT, "function '" + text_ + "': too deep");
    const double var[4] = {x, y, z, t};
    int sp = 0;
    for (const Inst& in : prog_) {
        switch (in.op) {
            case kConst: st[sp++] = in.v; break;
            case kVar: st[sp++] = var[in.i]; break;
            case kSlot:
                if (in.i >= 0) st[sp++] = slots[in.i];
                else slots[-1 - in.i] = st[--sp];
                break;
            case kAdd: --sp; st[sp - 1] += st[sp]; break;
            case kSub: --sp; st[sp - 1] -= st[sp]; break;
            case kMul: --sp; st[sp - 1] *= st[sp]; break;
            case kDiv: --sp; st[sp - 1] /= st[sp]; break;
            case kPow: case kPow2: --sp; st[sp - 1] = std::pow(st[sp - 1], st[sp]); break;
            case kAtan2: --sp; st[sp - 1] = std::atan2(st[sp - 1], st[sp]); break;
            case kMin: --sp; st[sp - 1] = std::fmin(st[sp - 1], st[sp]); break;
            case kMax: --sp; st[sp - 1] = std::fmax(st[sp - 1], st[sp]); break;
            case kNeg: st[sp - 1] = -st[sp - 1]; break;
            case kSin: st[sp - 1] = std::sin(st[sp - 1]); break;
            case kCos: st[sp - 1] = std::cos(st[sp - 1]); break;
            case kTan: st[sp - 1] = std::tan(st[sp - 1]); break;
            case kAsin: st[sp - 1] = std::asin(st[sp - 1]); break;
            case kAcos: st[sp - 1] = std::acos(st[sp - 1]); break;
            case kAtan: st[sp - 1] = std::atan(st[sp - 1]); break;
            case kSinh: st[sp - 1] = std::sinh(st[sp - 1]); break;
            case kCosh: st[sp - 1] = std::cosh(st[sp - 1]); break;
            case kTanh: st[sp - 1] = std::tanh(st[sp - 1]); break;
            case kExp: st[sp - 1] = std::exp(st[sp - 1]); break;
            case kLog: st[sp - 1] = std::log(st[sp - 1]); break;
            case kLog10: st[sp - 1] = std::log10(st[sp - 1]); break;
            case kSqrt: st[sp - 1] = std::sqrt(st[sp - 1]); break;
            case kAbs: st[sp - 1] = std::fabs(st[sp - 1]); break;
            default: throw Error(PA_ERR_STATE, "function '" + text_ + "': bad program");
        }
    }
    return st[sp - 1];
}

}  // namespace pa

using pa::guarded;

extern "C" {

int pa_expression_eval(const char* function, int n, const double* x, const double* y, const double* z, double t,
                       double* out) {
    return guarded([&] {
        PA_REQUIRE(function && n >= 0 && (n == 0 || out), PA_ERR_ARG, "pa_expression_eval: bad argument");
        const pa::Expr f(function);
        for (int i = 0; i < n; ++i) out[i] = f.eval(x ? x[i] : 0.0, y ? y[i] : 0.0, z ? z[i] : 0.0, t);
    });
}

}  // extern "C"
