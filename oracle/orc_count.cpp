// orc_count.cpp -- the oracle built with every floating-point value counted: an operation-count
// mode of oracle/rrtmg_oracle.c, for the algorithmic FLOP figure of the timed path.
//
// TEST INFRASTRUCTURE ONLY, like the oracle it wraps: built into oracle/liborc_count.so and run
// by oracle/count_flops.py and tests/test_flops.py, never by the product.
//
// The oracle's source is compiled unchanged as C++ with `double` standing for `cd`, a value type
// of the same size and layout whose +, -, *, / (binary, compound) and sqrt add one to fp_ops and
// whose exp, expm1, log and pow add one to fp_trans before computing in double exactly as the plain
// build does (same operands, same order, -ffp-contract=off), so its outputs are bit for bit the
// plain oracle's (tests/test_flops.py checks this). Negation, comparisons, min/max, fabs, table
// gathers and integer work are not counted. Arrays passed through the C interface are arrays of
// doubles either way.
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <cmath>

extern "C" {
uint64_t orc_fp_ops = 0, orc_fp_trans = 0;
}

struct cd {
  double v;
  cd() = default;
  constexpr cd(double x) : v(x) {}
  constexpr operator double() const { return v; }
  cd &operator+=(cd o) { orc_fp_ops++; v += o.v; return *this; }
  cd &operator-=(cd o) { orc_fp_ops++; v -= o.v; return *this; }
  cd &operator*=(cd o) { orc_fp_ops++; v *= o.v; return *this; }
  cd &operator/=(cd o) { orc_fp_ops++; v /= o.v; return *this; }
};
static_assert(sizeof(cd) == sizeof(double), "cd must have the layout of double");

#define CD_OP(op)                                                                   \
  inline cd operator op(cd a, cd b) { orc_fp_ops++; return cd(a.v op b.v); }      \
  inline cd operator op(cd a, double b) { orc_fp_ops++; return cd(a.v op b); }     \
  inline cd operator op(double a, cd b) { orc_fp_ops++; return cd(a op b.v); }     \
  inline cd operator op(cd a, int b) { orc_fp_ops++; return cd(a.v op b); }        \
  inline cd operator op(int a, cd b) { orc_fp_ops++; return cd(a op b.v); }
CD_OP(+)
CD_OP(-)
CD_OP(*)
CD_OP(/)
#undef CD_OP
inline cd operator-(cd a) { return cd(-a.v); }
inline cd sqrt(cd a) { orc_fp_ops++; return cd(std::sqrt(a.v)); }
inline cd exp(cd a) { orc_fp_trans++; return cd(std::exp(a.v)); }
inline cd log(cd a) { orc_fp_trans++; return cd(std::log(a.v)); }
inline cd expm1(cd a) { orc_fp_trans++; return cd(std::expm1(a.v)); }
inline cd pow(cd a, cd b) { orc_fp_trans++; return cd(std::pow(a.v, b.v)); }
inline cd pow(cd a, double b) { orc_fp_trans++; return cd(std::pow(a.v, b)); }
inline cd pow(double a, cd b) { orc_fp_trans++; return cd(std::pow(a, b.v)); }

#define double cd
extern "C" {
#include "rrtmg_oracle.c"

void orc_fp_counts(uint64_t *ops, uint64_t *trans) {
  *ops = orc_fp_ops;
  *trans = orc_fp_trans;
}
void orc_fp_reset(void) { orc_fp_ops = orc_fp_trans = 0; }
}
