// Reverse-mode automatic differentiation by operator overloading, for the
// CPU adjoint of the oracle (semo_oracle.cpp built with -DORC_AD).
//
// TEST INFRASTRUCTURE ONLY: the checker of the device adjoint.  Nothing in
// the product links it.
//
// An adouble carries its value and, while the tape records, the index of a
// tape variable.  Every arithmetic operation on recorded operands appends one
// statement res <- f(a, b) with the partial derivatives; the reverse sweep
// walks the statements backwards, adj[a] += da adj[res], adj[b] += db adj[res].
// A statement may instead be a custom reverse step (a linear solve: its
// adjoint is a transposed solve), run at its place in the sweep.  Values are
// computed with the same IEEE operations as plain double code, so a function
// evaluated in adouble returns bit-identical values.
#pragma once

#include <cmath>
#include <functional>
#include <vector>

namespace ad {

struct Stmt {
    int res, a, b;      // a, b < 0: no operand; a == -2: custom step number b
    double da, db;
};

struct Tape {
    bool on = false;
    int next = 0;                       // variables allocated
    std::vector<Stmt> st;
    std::vector<std::function<void(std::vector<double>&)>> custom;
    int fresh() { return next++; }
    void clear() { st.clear(); custom.clear(); next = 0; }
    void push(int r, int a, double da, int b, double db) { st.push_back({r, a, b, da, db}); }
    void push_custom(std::function<void(std::vector<double>&)> f) {
        st.push_back({-1, -2, (int)custom.size(), 0.0, 0.0});
        custom.push_back(std::move(f));
    }
    // adj must hold `next` entries, seeded at the outputs
    void reverse(std::vector<double>& adj) const {
        for (size_t k = st.size(); k-- > 0;) {
            const Stmt& s = st[k];
            if (s.a == -2) { custom[s.b](adj); continue; }
            double g = adj[s.res];
            if (g == 0.0) continue;
            if (s.a >= 0) adj[s.a] += s.da * g;
            if (s.b >= 0) adj[s.b] += s.db * g;
        }
    }
};

inline Tape& tape() {
    static Tape t;
    return t;
}

struct adouble {
    double v = 0.0;
    int i = -1;                         // tape variable, -1: a constant
    adouble() = default;
    adouble(double x) : v(x), i(-1) {}
    // a new independent variable (recorded)
    static adouble var(double x) {
        adouble r(x);
        if (tape().on) r.i = tape().fresh();
        return r;
    }
    adouble& operator+=(const adouble& o);
    adouble& operator-=(const adouble& o);
    adouble& operator*=(const adouble& o);
    adouble& operator/=(const adouble& o);
};

inline double val(const adouble& a) { return a.v; }

inline adouble rec1(double v, const adouble& a, double da) {
    adouble r(v);
    if (tape().on && a.i >= 0) { r.i = tape().fresh(); tape().push(r.i, a.i, da, -1, 0.0); }
    return r;
}
inline adouble rec2(double v, const adouble& a, double da, const adouble& b, double db) {
    adouble r(v);
    if (tape().on && (a.i >= 0 || b.i >= 0)) { r.i = tape().fresh(); tape().push(r.i, a.i, da, b.i, db); }
    return r;
}

inline adouble operator+(const adouble& a, const adouble& b) { return rec2(a.v + b.v, a, 1.0, b, 1.0); }
inline adouble operator-(const adouble& a, const adouble& b) { return rec2(a.v - b.v, a, 1.0, b, -1.0); }
inline adouble operator*(const adouble& a, const adouble& b) { return rec2(a.v * b.v, a, b.v, b, a.v); }
inline adouble operator/(const adouble& a, const adouble& b) {
    double q = a.v / b.v;
    return rec2(q, a, 1.0 / b.v, b, -q / b.v);
}
inline adouble operator-(const adouble& a) { return rec1(-a.v, a, -1.0); }
inline adouble operator+(const adouble& a) { return a; }

inline adouble operator+(const adouble& a, double b) { return rec1(a.v + b, a, 1.0); }
inline adouble operator+(double a, const adouble& b) { return rec1(a + b.v, b, 1.0); }
inline adouble operator-(const adouble& a, double b) { return rec1(a.v - b, a, 1.0); }
inline adouble operator-(double a, const adouble& b) { return rec1(a - b.v, b, -1.0); }
inline adouble operator*(const adouble& a, double b) { return rec1(a.v * b, a, b); }
inline adouble operator*(double a, const adouble& b) { return rec1(a * b.v, b, a); }
inline adouble operator/(const adouble& a, double b) { return rec1(a.v / b, a, 1.0 / b); }
inline adouble operator/(double a, const adouble& b) {
    double q = a / b.v;
    return rec1(q, b, -q / b.v);
}

inline adouble& adouble::operator+=(const adouble& o) { return *this = *this + o; }
inline adouble& adouble::operator-=(const adouble& o) { return *this = *this - o; }
inline adouble& adouble::operator*=(const adouble& o) { return *this = *this * o; }
inline adouble& adouble::operator/=(const adouble& o) { return *this = *this / o; }

#define AD_CMP(op)                                                                        \
    inline bool operator op(const adouble& a, const adouble& b) { return a.v op b.v; }   \
    inline bool operator op(const adouble& a, double b) { return a.v op b; }             \
    inline bool operator op(double a, const adouble& b) { return a op b.v; }
AD_CMP(<) AD_CMP(>) AD_CMP(<=) AD_CMP(>=) AD_CMP(==) AD_CMP(!=)
#undef AD_CMP

// the derivative on the branch the value takes (a kink's one-sided
// derivative is that of the branch evaluated)
inline adouble sqrt(const adouble& a) {
    double s = std::sqrt(a.v);
    return rec1(s, a, 0.5 / s);
}
inline adouble fabs(const adouble& a) { return rec1(std::fabs(a.v), a, a.v < 0.0 ? -1.0 : 1.0); }
inline adouble pow(const adouble& a, double e) {
    return rec1(std::pow(a.v, e), a, e == 2.0 ? 2.0 * a.v : e * std::pow(a.v, e - 1.0));
}
inline adouble atan(const adouble& a) { return rec1(std::atan(a.v), a, 1.0 / (1.0 + a.v * a.v)); }

}  // namespace ad
