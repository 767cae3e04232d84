// Reference-style host code on the drop-in classes (xflow_amd/host): the
// manufactured-solution Poisson check of T3dPoissonProblem
// (XProblem.cpp:554-817, here with u = (1 + x) sin x sin y sin z) written against
// the reference's call signatures, solved on the GPU.  Prints one JSON line;
// exits non-zero when a check fails.  Run by tests/test_gpu_dropin.py.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "xflow_dropin.h"

// u = (1 + x) sin x sin y sin z, -Lap u = 3 (1 + x) S - 2 cos x sin y sin z: zero on the
// boundary of [0, pi]^3 and not a discrete eigenvector, so the solvers iterate.
static double sol(double x, double y, double z) { return (1 + x) * sin(x) * sin(y) * sin(z); }
static double rhs_of(double x, double y, double z) {
  return 3 * (1 + x) * sin(x) * sin(y) * sin(z) - 2 * cos(x) * sin(y) * sin(z);
}

struct Result {
  int n;
  double trunc, err_cg, err_bicgstab, err_mr, err_fast, err_cycle;
  int it_cg, it_bicgstab, it_mr, it_fast, it_cycle;
};

static Result run(int cells) {
  Result R{};
  const double L = M_PI;
  double h = L / cells;
  TUniformSeparator *sx = new TUniformSeparator(0, L, h);
  TUniformSeparator *sy = new TUniformSeparator(0, L, h);
  TUniformSeparator *sz = new TUniformSeparator(0, L, h);
  const int N = sx->GetEndIndex(), Lc = sy->GetEndIndex(), M = sz->GetEndIndex();
  const double *x = sx->GetDimension(), *y = sy->GetDimension(), *z = sz->GetDimension();
  T3dNumberMask mask(N + 1, Lc + 1, M + 1);
  for (int i = 0; i <= N; i++)
    for (int j = 0; j <= Lc; j++)
      for (int k = 0; k <= M; k++) {
        bool border = i == 0 || j == 0 || k == 0 || i == N || j == Lc || k == M;
        mask[i][j][k] = border ? TDefinedBorderPoint : TActualPoint;
      }
  T3dNormalGrid grid(sx, sy, sz, mask, nullptr);
  // space mask (XProblem.cpp:4155-4182): Defined -> 0, Actual -> 1
  T3dNumberMask &spaceMask = mask.Copy();
  for (int i = 0; i <= N; i++)
    for (int j = 0; j <= Lc; j++)
      for (int k = 0; k <= M; k++) spaceMask[i][j][k] = mask[i][j][k] == TActualPoint ? 1 : 0;
  TMask *masks[1] = {&spaceMask};
  TRnRelease3dSpaceModel &model = *new TRnRelease3dSpaceModel(1, masks);
  const int name = 555;
  TRnSpace::AddModel(name, &model, true);

  TRnRelease3dSpace &rightPart = *new TRnRelease3dSpace(name);
  TRnRelease3dSpace &solution = *new TRnRelease3dSpace(name);
  TRnRelease3dSpace &approximation = *new TRnRelease3dSpace(name);
  TRnRelease3dSpace &error = *new TRnRelease3dSpace(name);
  for (int i = 0; i <= N; i++)
    for (int j = 0; j <= Lc; j++)
      for (int k = 0; k <= M; k++) {
        solution[i][j][k] = sol(x[i], y[j], z[k]);
        rightPart[i][j][k] = rhs_of(x[i], y[j], z[k]);
        // boundary values ride in the initial guess, as in the reference
        approximation[i][j][k] = mask[i][j][k] == TActualPoint ? 0.0 : sol(x[i], y[j], z[k]);
      }
  const TNodesGrid *grids[1] = {&grid};
  T3dLaplaceOperator &A = *new T3dLaplaceOperator(name, 1, grids);
  A.ConstructActivityTables();

  // truncation error of the operator on the exact solution (XProblem.cpp:694-698)
  error |= A(solution) - rightPart;
  R.trunc = error.GetAbsMaxElement();

  auto field_err = [&](const TRnSpace &res) {
    TRnRelease3dSpace &d = *new TRnRelease3dSpace(name);
    d |= res;
    d |= d - solution;
    double e = d.GetAbsMaxElement();
    delete &d;
    return e;
  };

  TSLAE slae(A, approximation, rightPart);
  slae.SolveByConjugateGradients();
  R.err_cg = field_err(slae.Result);
  R.it_cg = slae.LastStats().iters;

  TSLAE slae2(A, approximation, rightPart);
  slae2.SolveByBiConjugateGradientsStab();
  R.err_bicgstab = field_err(slae2.Result);
  R.it_bicgstab = slae2.LastStats().iters;

  TSLAE slae4(A, approximation, rightPart);
  slae4.SolveByFastConjugateGradients(1e-10, false);
  R.err_fast = field_err(slae4.Result);
  R.it_fast = slae4.LastStats().iters;

  TSLAE slae5(A, approximation, rightPart);
  // the cycle without a network: at 24^3 it has one level (a 64-cell x extent
  // is the coarsest), so this checks the drop-in call and its answer, not a gain
  slae5.SolveByLearnedPreconditioner(nullptr, 1e-10, false);
  R.err_cycle = field_err(slae5.Result);
  R.it_cycle = slae5.LastStats().inner_iters;

  if (cells <= 12) {
    TSLAE slae3(A, approximation, rightPart);
    slae3.SolveByMinimalResiduals();
    R.err_mr = field_err(slae3.Result);
    R.it_mr = slae3.LastStats().iters;
  }
  R.n = cells;
  delete &A;
  delete &rightPart;
  delete &solution;
  delete &approximation;
  delete &error;
  TRnSpace::RemoveModel(name);
  delete &spaceMask;
  return R;
}

int main() {
  try {
    Result a = run(12), b = run(24);
    // Thomas (XSLAE.cpp:416-445) on a 6-point system with a known answer
    double u[6] = {1, 0, 0, 0, 0, 2}, aa[6] = {0, -1, -1, -1, -1, 0}, bb[6] = {0, 2, 2, 2, 2, 0},
           cc[6] = {0, -1, -1, -1, -1, 0}, ff[6] = {0, 0, 0, 0, 0, 0};
    TSLAE::SolveByScalarRunning(5, u, aa, bb, cc, ff);  // linear interpolation 1 -> 2
    double thomas_err = 0;
    for (int q = 0; q <= 5; q++) thomas_err = fmax(thomas_err, fabs(u[q] - (1 + q / 5.0)));
    printf("{\"trunc12\": %.3e, \"trunc24\": %.3e, \"err_cg12\": %.3e, \"err_cg24\": %.3e, \"it_cg12\": %d, "
           "\"it_cg24\": %d, \"err_bicgstab24\": %.3e, \"it_bicgstab24\": %d, \"err_fast24\": %.3e, "
           "\"it_fast24\": %d, \"err_cycle24\": %.3e, \"it_cycle24\": %d, \"err_mr12\": %.3e, \"it_mr12\": %d, "
           "\"thomas_err\": %.3e}\n",
           a.trunc, b.trunc, a.err_cg, b.err_cg, a.it_cg, b.it_cg, b.err_bicgstab, b.it_bicgstab, b.err_fast,
           b.it_fast, b.err_cycle, b.it_cycle, a.err_mr, a.it_mr, thomas_err);
    // second-order operator and solution error (halving h divides by ~4)
    bool ok = b.trunc < a.trunc / 3 && b.err_cg < a.err_cg / 3 && b.err_cg < 5e-3 &&
              fabs(b.err_bicgstab - b.err_cg) < 1e-4 && fabs(b.err_fast - b.err_cg) < 1e-4 &&
              fabs(b.err_cycle - b.err_cg) < 1e-4 && b.it_cycle > 0 &&
              fabs(a.err_mr - a.err_cg) < 1e-4 && thomas_err < 1e-14 && a.it_cg >= 3 && b.it_cg > a.it_cg;
    return ok ? 0 : 1;
  } catch (const char *e) {
    fprintf(stderr, "thrown: %s\n", e);
    return 2;
  }
}
