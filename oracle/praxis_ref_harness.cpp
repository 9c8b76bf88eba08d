// ORACLE harness (test infrastructure only): exposes the reference's own
// PRAXIS, CUDA/praxis.cpp:879 `praxis(t0, h0, n, prin, x, f)`, compiled
// from the reference checkout by oracle/build_ref.py, through a C ABI so
// that tests can record its iterates on deterministic objectives and check
// the project's own PRAXIS-style minimizer against them.  Nothing here is
// part of the product.
#include <cstdint>

double praxis(double t0, double h0, int n, int prin, double x[], double f(double x[], int n));

namespace {
typedef double (*objective_t)(const double* x, int n);
objective_t g_obj = nullptr;
long g_nf = 0;
double* g_trace = nullptr;   // optional: nf rows of (x..., f)
long g_trace_cap = 0;
int g_n = 0;

double trampoline(double x[], int n) {
  double v = g_obj(x, n);
  if (g_trace && g_nf < g_trace_cap) {
    for (int i = 0; i < n; ++i) g_trace[g_nf * (n + 1) + i] = x[i];
    g_trace[g_nf * (n + 1) + n] = v;
  }
  ++g_nf;
  return v;
}
}  // namespace

extern "C" {
// Minimise obj from x (in/out).  Returns the final f; *nf = evaluations.
// trace (nullable) receives up to trace_cap rows of (x[0..n-1], f).
double praxis_ref_run(double t0, double h0, int n, double* x, objective_t obj, long* nf,
                      double* trace, long trace_cap) {
  g_obj = obj;
  g_nf = 0;
  g_trace = trace;
  g_trace_cap = trace_cap;
  g_n = n;
  double fx = praxis(t0, h0, n, 0, x, trampoline);
  if (nf) *nf = g_nf;
  g_obj = nullptr;
  g_trace = nullptr;
  return fx;
}
}
