// Prints second_order_tables.h: the second-order rearrangement
// (second_order.h) of the three tableaus (tableaus.h) as a brace initializer
// of exact hexadecimal floating-point literals.  Built and run with the host
// compiler by bqp_amd/build.py before the HIP units compile.
#include <cstdio>

#include "second_order.h"

static void row(FILE* f, const double* v, int n) {
  std::fputs("{", f);
  for (int i = 0; i < n; ++i) std::fprintf(f, "%s%a", i ? ", " : "", v[i]);
  std::fputs("}", f);
}

int main(int argc, char** argv) {
  FILE* f = argc > 1 ? std::fopen(argv[1], "w") : stdout;
  if (!f) return 1;
  static const bqp_tableau tabs[3] = {BQP_DOPRI5_TABLEAU, BQP_DOP853_TABLEAU, BQP_RK87_TABLEAU};
  std::fputs("/* Generated by gen_second_order.cpp from tableaus.h (do not edit). */\n"
             "#ifndef BQP_SECOND_ORDER_TABLES_H\n#define BQP_SECOND_ORDER_TABLES_H\n"
             "#define BQP_SECOND_ORDER_TABLES \\\n  { \\\n", f);
  for (int m = 0; m < 3; ++m) {
    bqp::SecondOrderTableau o;
    bqp::make_second_order(tabs[m], o);
    std::fputs("    { ", f);
    row(f, o.csum, BQP_MAX_STAGES);
    std::fputs(", \\\n      {", f);
    for (int i = 0; i < BQP_MAX_STAGES; ++i) {
      if (i) std::fputs(", \\\n       ", f);
      row(f, o.a2[i], BQP_MAX_STAGES);
    }
    std::fputs("}, \\\n      ", f);
    row(f, o.ba, BQP_MAX_STAGES);
    std::fputs(", \\\n      ", f);
    row(f, o.ea, BQP_MAX_STAGES);
    std::fputs(", \\\n      ", f);
    row(f, o.e3a, BQP_MAX_STAGES);
    std::fprintf(f, ", \\\n      %a, %a, %a, \\\n      {", o.bsum, o.esum, o.e3sum);
    for (int i = 0; i < BQP_MAX_STAGES; ++i) {
      if (i) std::fputs(", \\\n       ", f);
      std::fputs("{", f);
      for (int k = 0; k < bqp::STAGE_CHUNKS; ++k) {
        if (k) std::fputs(", ", f);
        row(f, o.sc[i][k], 2 * bqp::STAGE_CHUNK);
      }
      std::fputs("}", f);
    }
    std::fprintf(f, "} }%s \\\n", m < 2 ? "," : "");
  }
  std::fputs("  }\n#endif\n", f);
  if (f != stdout) std::fclose(f);
  return 0;
}
