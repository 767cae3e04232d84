// Reference-style host code for the packed operator on the drop-in classes:
// the Rn path of T3dPoissonProblem (XProblem.cpp:225-330) — a 1-D model with one
// entry per unknown, vectors copied through TRnSpace::operator[] (1-based Rn
// index), T1dPackOperator built from the 3-D operator with a NULL vector, TSLAE
// on it — plus BiCG with the transposed rows (operator~) on a non-symmetric
// mask.  Prints one JSON line; exits non-zero when a check fails.  Run by
// tests/test_gpu_dropin.py.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "xflow_dropin.h"

static double sol(double x, double y, double z) { return (1 + x) * sin(x) * sin(y) * sin(z); }
static double rhs_of(double x, double y, double z) {
  return 3 * (1 + x) * sin(x) * sin(y) * sin(z) - 2 * cos(x) * sin(y) * sin(z);
}

struct Out {
  double prod_diff = 0, row_diff = 0, err3 = 0, errRn = 0, sol_diff = 0;
  int it3 = 0, itRn = 0;
  double resid_bicg = 0, resid_bicg_t = 0, t_vs_a = 0;
  int it_bicg = 0, it_bicg_t = 0;
  bool threw_vector = false, threw_tilde_3d = false;
};

// predefined_face: the x = 0 face PreDefined (a non-symmetric row), else Defined
static void problem(int cells, bool predefined_face, int name, int nameRn, Out &o) {
  const double L = M_PI, h = L / cells;
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
        bool face0 = i == 0 && j > 0 && j < Lc && k > 0 && k < M;
        mask[i][j][k] = !border ? TActualPoint : (predefined_face && face0 ? TPreDefinedBorderPoint : TDefinedBorderPoint);
      }
  // outward normals (-1, 0, 0) on the PreDefined face: its row reads the
  // neighbour along the normal (XOperator.cpp:1034-1069)
  const size_t nc = (size_t)(N + 1) * (Lc + 1) * (M + 1);
  std::vector<T3dVector> nv(nc);
  std::vector<const T3dVector *> np_(nc, nullptr);
  std::vector<const T3dVector *const *> nj((size_t)(N + 1) * (Lc + 1));
  std::vector<const T3dVector *const *const *> ni(N + 1);
  for (int i = 0; i <= N; i++) {
    for (int j = 0; j <= Lc; j++) {
      for (int k = 0; k <= M; k++) {
        const size_t c = ((size_t)i * (Lc + 1) + j) * (M + 1) + k;
        if (mask[i][j][k] == TPreDefinedBorderPoint) {
          nv[c].X = -1;
          nv[c].Y = 0;
          nv[c].Z = 0;
          np_[c] = &nv[c];
        }
      }
      nj[(size_t)i * (Lc + 1) + j] = np_.data() + ((size_t)i * (Lc + 1) + j) * (M + 1);
    }
    ni[i] = nj.data() + (size_t)i * (Lc + 1);
  }
  T3dNormalGrid grid(sx, sy, sz, mask, predefined_face ? ni.data() : nullptr);
  // space mask (XProblem.cpp:4155-4182): Defined -> 0, Actual and PreDefined -> 1
  T3dNumberMask &spaceMask = mask.Copy();
  for (int i = 0; i <= N; i++)
    for (int j = 0; j <= Lc; j++)
      for (int k = 0; k <= M; k++) spaceMask[i][j][k] = mask[i][j][k] == TDefinedBorderPoint ? 0 : 1;
  TMask *masks[1] = {&spaceMask};
  TRnSpace::AddModel(name, new TRnRelease3dSpaceModel(1, masks), true);
  TRnRelease3dSpace &rightPart = *new TRnRelease3dSpace(name);
  TRnRelease3dSpace &solution = *new TRnRelease3dSpace(name);
  TRnRelease3dSpace &approximation = *new TRnRelease3dSpace(name);
  for (int i = 0; i <= N; i++)
    for (int j = 0; j <= Lc; j++)
      for (int k = 0; k <= M; k++) {
        solution[i][j][k] = sol(x[i], y[j], z[k]);
        rightPart[i][j][k] = mask[i][j][k] == TPreDefinedBorderPoint ? 0.0 : rhs_of(x[i], y[j], z[k]);
        approximation[i][j][k] = mask[i][j][k] == TActualPoint ? 0.0 : sol(x[i], y[j], z[k]);
      }
  const TNodesGrid *grids[1] = {&grid};
  T3dLaplaceOperator &A = *new T3dLaplaceOperator(name, 1, grids);

  // the Rn model and vectors, as XProblem.cpp:225-255
  const int Lr = approximation.GetEndIndex();
  T1dNumberMask &spaceMaskRn = *new T1dNumberMask(Lr);
  for (int l = 0; l < Lr; l++) spaceMaskRn[l] = TActualPoint;
  TMask *masksRn[1] = {&spaceMaskRn};
  TRnSpace::AddModel(nameRn, new TRnRelease1dSpaceModel(1, masksRn), true);
  TRnRelease1dSpace &solutionRn = *new TRnRelease1dSpace(nameRn);
  TRnRelease1dSpace &rightPartRn = *new TRnRelease1dSpace(nameRn);
  TRnRelease1dSpace &approximationRn = *new TRnRelease1dSpace(nameRn);
  TRnRelease1dSpace &errorRn = *new TRnRelease1dSpace(nameRn);
  for (int l = 1; l < Lr + 1; l++) {
    solutionRn[l - 1] = ((TRnSpace &)solution)[l];
    rightPartRn[l - 1] = ((TRnSpace &)rightPart)[l];
    approximationRn[l - 1] = ((TRnSpace &)approximation)[l];
  }
  // a non-NULL vector is the reference's unfinished Gauss transformation: it throws
  double *helpVector = new double[Lr]();
  try {
    T1dPackOperator bad(nameRn, A, &helpVector);
  } catch (const char *) {
    o.threw_vector = true;
  }
  delete[] helpVector;
  double *nullVector = nullptr;
  T1dPackOperator &ARn = *new T1dPackOperator(nameRn, A, &nullVector);

  if (!predefined_face) {
    // products agree with the 3-D operator at every unknown; rows through operator()(0, i)
    TRnRelease3dSpace &y3 = *new TRnRelease3dSpace(name);
    TRnRelease1dSpace &yRn = *new TRnRelease1dSpace(nameRn);
    y3 |= A(solution);
    yRn |= ARn(solutionRn);
    double scale = y3.GetAbsMaxElement();
    for (int l = 1; l <= Lr; l++) {
      o.prod_diff = fmax(o.prod_diff, fabs(((TRnSpace &)y3)[l] - yRn[l - 1]) / scale);
      o.row_diff = fmax(o.row_diff, fabs(ARn(0, l - 1) - yRn[l - 1]));
    }
    try {
      ~A;
    } catch (const char *) {
      o.threw_tilde_3d = true;
    }
    // CG3 (XSLAE.cpp:56-157) on the 3-D operator and on the packed one
    TSLAE slae3(A, approximation, rightPart);
    slae3.SolveByConjugateGradients();
    o.it3 = slae3.LastStats().iters;
    TRnRelease3dSpace &e3 = *new TRnRelease3dSpace(name);
    e3 |= slae3.Result;
    e3 |= e3 - solution;
    o.err3 = e3.GetAbsMaxElement();
    TSLAE &slae = *new TSLAE(ARn, approximationRn, rightPartRn);
    slae.SolveByConjugateGradients();
    o.itRn = slae.LastStats().iters;
    approximationRn |= slae.Result;
    errorRn |= solutionRn - approximationRn;
    o.errRn = errorRn.GetAbsMaxElement();
    for (int l = 1; l <= Lr; l++)
      o.sol_diff = fmax(o.sol_diff, fabs(slae3.GetResult()[l] - approximationRn[l - 1]));
    delete &slae;
    delete &y3;
    delete &yRn;
    delete &e3;
  } else {
    // BiCG with the transposed rows (XSLAE.cpp:159-348): A x = b, then A^T x = b
    TSLAE &s1 = *new TSLAE(ARn, approximationRn, rightPartRn);
    s1.SolveByBiConjugateGradients();
    o.it_bicg = s1.LastStats().iters;
    TRnRelease1dSpace &xa = *new TRnRelease1dSpace(nameRn);
    xa |= s1.Result;
    errorRn |= ARn(xa) - rightPartRn;
    o.resid_bicg = errorRn.GetAbsMaxElement();
    ~ARn;
    TSLAE &s2 = *new TSLAE(ARn, approximationRn, rightPartRn);
    s2.SolveByBiConjugateGradients();
    o.it_bicg_t = s2.LastStats().iters;
    TRnRelease1dSpace &xt = *new TRnRelease1dSpace(nameRn);
    xt |= s2.Result;
    errorRn |= ARn(xt) - rightPartRn;  // ARn is transposed now
    o.resid_bicg_t = errorRn.GetAbsMaxElement();
    errorRn |= xt - xa;
    o.t_vs_a = errorRn.GetAbsMaxElement() / xa.GetAbsMaxElement();
    delete &s1;
    delete &s2;
    delete &xa;
    delete &xt;
  }
  delete &ARn;
  delete &A;
  delete &rightPart;
  delete &solution;
  delete &approximation;
  delete &solutionRn;
  delete &rightPartRn;
  delete &approximationRn;
  delete &errorRn;
  TRnSpace::RemoveModel(nameRn);
  TRnSpace::RemoveModel(name);
  delete &spaceMaskRn;
  delete &spaceMask;
}

int main() {
  try {
    Out o;
    problem(12, false, 601, 602, o);
    problem(10, true, 603, 604, o);
    printf("{\"prod_diff\": %.3e, \"row_diff\": %.3e, \"err3\": %.3e, \"errRn\": %.3e, \"sol_diff\": %.3e, "
           "\"it3\": %d, \"itRn\": %d, \"resid_bicg\": %.3e, \"resid_bicg_t\": %.3e, \"t_vs_a\": %.3e, "
           "\"it_bicg\": %d, \"it_bicg_t\": %d, \"threw_vector\": %d, \"threw_tilde_3d\": %d}\n",
           o.prod_diff, o.row_diff, o.err3, o.errRn, o.sol_diff, o.it3, o.itRn, o.resid_bicg, o.resid_bicg_t,
           o.t_vs_a, o.it_bicg, o.it_bicg_t, (int)o.threw_vector, (int)o.threw_tilde_3d);
    bool ok = o.prod_diff < 1e-13 && o.row_diff == 0 && abs(o.it3 - o.itRn) <= 2 && fabs(o.err3 - o.errRn) < 1e-6 &&
              o.sol_diff < 1e-5 && o.resid_bicg < 1e-6 && o.resid_bicg_t < 1e-6 && o.t_vs_a > 1e-6 &&
              o.it_bicg > 3 && o.it_bicg_t > 3 && o.threw_vector && o.threw_tilde_3d;
    return ok ? 0 : 1;
  } catch (const char *e) {
    printf("{\"error\": \"%s\"}\n", e);
    return 2;
  }
}
