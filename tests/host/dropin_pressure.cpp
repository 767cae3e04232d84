// The pressure solve of a projection step written against the drop-in classes
// with the call forms TVPProblem::SolveBySplitting uses for it
// (XProblem.cpp:5516-5590): T3dVariableDivergentedOperator::withDensity,
// `p.EndIndex`, TRnSpace::Assign, `v |= 0`, `v |= ((TRnLinearOperator&)A)(z) -
// b`, `v |= (-1)*v`, TRnRelease1dSpace vectors filled through the 1-based
// `((TRnSpace&)v)[rn]`, TSLAE on a T1dPackOperator and on the 3-D operator,
// SolveByBiConjugateGradientsStab, `slae.Result` / `slae->Result`, UpdateRP /
// UpdateP on later steps; plus the Property accessors of masks, separators,
// grids and models, the six-argument TSLAE and UpdateA.  A channel with
// PreDefined inflow and outflow planes and a variable density (non-symmetric
// rows).  Prints one JSON line; exits non-zero when a check fails.  Run by
// tests/test_gpu_dropin.py.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "xflow_dropin.h"

static int g_failures = 0;
#define EXPECT(cond)                                              \
  do {                                                            \
    if (!(cond)) {                                                \
      fprintf(stderr, "check failed: %s (line %d)\n", #cond, __LINE__); \
      g_failures++;                                               \
    }                                                             \
  } while (0)

// max |A p - b| over the unknowns of the model
static double residual_max(T3dVariableDivergentedOperator &A, TRnRelease3dSpace &p, TRnRelease3dSpace &b,
                           int model) {
  TRnRelease3dSpace &r = *new TRnRelease3dSpace(model);
  r |= ((TRnLinearOperator &)A)(p) - b;
  double m = r.GetAbsMaxElement();
  delete &r;
  return m;
}

int main() {
  const int nx = 20, ny = 10, nz = 10;
  const double Lx = 2.0, h = Lx / nx;
  TUniformSeparator *sx = new TUniformSeparator(0, Lx, h);
  TUniformSeparator *sy = new TUniformSeparator(0, ny * h, h);
  TUniformSeparator *sz = new TUniformSeparator(0, nz * h, h);
  const int N = sx->EndIndex, L = sy->EndIndex, M = sz->EndIndex;
  EXPECT(N == nx && L == ny && M == nz && sx->CountNodes == nx + 1);
  EXPECT(std::fabs(sx->Separation[0] - h) < 1e-15 && sx->Dimension[0] == 0.0 && sx->EndValue == Lx);

  // pressure points: Actual inside, PreDefined inflow (i = 0) and outflow
  // (i = N) planes with outward normals, Fictive walls and edges
  T3dNumberMask mask(N + 1, L + 1, M + 1);
  const size_t nc = (size_t)(N + 1) * (L + 1) * (M + 1);
  std::vector<T3dVector> nv(nc);
  std::vector<const T3dVector *> nptr(nc, nullptr);
  std::vector<const T3dVector *const *> nrow((size_t)(N + 1) * (L + 1));
  std::vector<const T3dVector *const *const *> nplane(N + 1);
  for (int i = 0; i <= N; i++)
    for (int j = 0; j <= L; j++)
      for (int k = 0; k <= M; k++) {
        const bool wall = j == 0 || j == L || k == 0 || k == M;
        const bool end = i == 0 || i == N;
        int code = TActualPoint;
        if (wall) code = TFictivePoint;
        else if (end) code = TPreDefinedBorderPoint;
        mask[i][j][k] = code;
        const size_t c = ((size_t)i * (L + 1) + j) * (M + 1) + k;
        if (code == TPreDefinedBorderPoint) {
          nv[c].X = i == 0 ? -1 : 1;
          nv[c].Y = 0;
          nv[c].Z = 0;
          nptr[c] = &nv[c];
        }
      }
  for (size_t r = 0; r < nrow.size(); r++) nrow[r] = nptr.data() + r * (M + 1);
  for (int i = 0; i <= N; i++) nplane[i] = nrow.data() + (size_t)i * (L + 1);
  EXPECT(mask.N == N + 1 && mask.L == L + 1 && mask.M == M + 1);
  EXPECT(mask.Mask[0][3][4] == TPreDefinedBorderPoint && mask.Mask[5][0][4] == TFictivePoint);

  T3dNormalGrid grid(sx, sy, sz, mask, nplane.data());
  const T3dNumberMask &gm = grid.Mask;
  EXPECT(gm.N == N + 1 && grid.GetSeparator1().EndIndex == N && ((const TSeparator &)grid.Separator3).EndIndex == M);
  const T3dVector *const *const *const *gn = grid.Normals;
  EXPECT(gn[0][3][4] && gn[0][3][4]->X == -1 && gn[N][3][4]->X == 1 && gn[5][5][5] == nullptr);

  // the space mask of the pressure unknowns: PreDefined rows are unknowns
  // (XProblem.cpp:4155-4182)
  T3dNumberMask &spaceMask = mask.Copy();
  int unknowns = 0;
  for (int i = 0; i <= N; i++)
    for (int j = 0; j <= L; j++)
      for (int k = 0; k <= M; k++) {
        const int v = mask[i][j][k];
        const int s = (v == TActualPoint || v == TPreDefinedBorderPoint) ? 1 : 0;
        spaceMask[i][j][k] = s;
        unknowns += s;
      }
  const int modelP = 0, modelRn = 1;
  const TMask *masks[1] = {&spaceMask};
  TRnRelease3dSpaceModel &modelPObj = *new TRnRelease3dSpaceModel(1, masks);
  TRnSpace::AddModel(modelP, &modelPObj, false);
  EXPECT(modelPObj.EndIndex == unknowns);

  // density, right-hand side (interior values and boundary pressures), pressure
  TRnRelease3dSpace &density = *new TRnRelease3dSpace(modelP);
  TRnRelease3dSpace &rightPart = *new TRnRelease3dSpace(modelP);
  TRnRelease3dSpace &p = *new TRnRelease3dSpace(modelP);
  TRnRelease3dSpace &p2 = *new TRnRelease3dSpace(modelP);
  const double *x = sx->Dimension, *y = sy->Dimension, *z = sz->Dimension;
  for (int i = 0; i <= N; i++)
    for (int j = 0; j <= L; j++)
      for (int k = 0; k <= M; k++) {
        density[i][j][k] = 1.0 + 0.2 * sin(x[i]) * cos(y[j] + z[k]);
        double rp = 0;
        if (mask[i][j][k] == TActualPoint) rp = sin(3 * x[i]) * cos(2 * y[j]) * sin(z[k] + 0.3);
        else if (mask[i][j][k] == TPreDefinedBorderPoint) rp = i == 0 ? 6.0 : 1.0;
        rightPart[i][j][k] = rp;
        p[i][j][k] = 0;
        p2[i][j][k] = 0;
      }
  const TNodesGrid *grids[1] = {&grid};
  T3dVariableDivergentedOperator *APtr = new T3dVariableDivergentedOperator(modelP, 1, grids);

  // ---- the packed branch: unknowns through Rn vectors, the boundary
  // contributions of A moved to the right-hand side
  T3dVariableDivergentedOperator &A = *APtr;
  A.withDensity(density);
  int RN = p.EndIndex;
  EXPECT(RN == unknowns);
  T1dNumberMask &rnMask = *new T1dNumberMask(RN);
  for (int q = 0; q < RN; q++) rnMask[q] = TActualPoint;
  const TMask *rnMasks[1] = {&rnMask};
  TRnSpace::AddModel(modelRn, new TRnRelease1dSpaceModel(1, rnMasks), false);
  TRnRelease3dSpace &zero = *new TRnRelease3dSpace(modelP);
  TRnRelease3dSpace &shifted = *new TRnRelease3dSpace(modelP);
  zero.Assign(p);
  zero |= 0;
  shifted |= ((TRnLinearOperator &)A)(zero) - rightPart;
  shifted |= (-1) * shifted;
  delete &zero;
  TRnRelease1dSpace &bRn = *new TRnRelease1dSpace(modelRn);
  TRnRelease1dSpace &xRn = *new TRnRelease1dSpace(modelRn);
  for (int rn = 1; rn < RN + 1; rn++) {
    bRn[rn - 1] = ((TRnSpace &)shifted)[rn];
    xRn[rn - 1] = ((TRnSpace &)p)[rn];
  }
  delete &shifted;
  double *noVector = nullptr;
  T1dPackOperator *ARnPtr = new T1dPackOperator(modelRn, A, &noVector);
  T1dPackOperator &ARn = *ARnPtr;
  TSLAE &packed = *new TSLAE(ARn, xRn, bRn);
  packed.SolveByBiConjugateGradientsStab();
  const int itPack = packed.LastStats().iters;
  xRn |= packed.Result;
  delete &packed;
  for (int rn = 1; rn < RN + 1; rn++) ((TRnSpace &)p)[rn] = xRn[rn - 1];
  const double resPack = residual_max(A, p, rightPart, modelP);

  // ---- the grid branch, first step: a TSLAE on the 3-D operator
  TSLAE *slae = new TSLAE(*APtr, p2, rightPart);
  slae->SolveByBiConjugateGradientsStab();
  const int itGrid = slae->LastStats().iters;
  p2 |= slae->Result;
  const double resGrid = residual_max(A, p2, rightPart, modelP);
  double diff = 0, scale = 0;
  for (int rn = 1; rn < RN + 1; rn++) {
    diff = std::fmax(diff, std::fabs(((TRnSpace &)p)[rn] - ((TRnSpace &)p2)[rn]));
    scale = std::fmax(scale, std::fabs(((TRnSpace &)p2)[rn]));
  }

  // ---- the grid branch, a later step: new density and right-hand side,
  // the same TSLAE through UpdateA / UpdateRP / UpdateP
  for (int i = 0; i <= N; i++)
    for (int j = 0; j <= L; j++)
      for (int k = 0; k <= M; k++) {
        density[i][j][k] = 1.0 + 0.3 * cos(x[i] + y[j]) * sin(z[k]);
        if (mask[i][j][k] == TActualPoint) rightPart[i][j][k] *= 0.5;
      }
  APtr->withDensity(density);
  slae->UpdateA(*APtr);
  slae->UpdateRP(rightPart);
  slae->UpdateP(p2);
  slae->SolveByBiConjugateGradientsStab();
  p2 |= slae->Result;
  const double resStep2 = residual_max(*APtr, p2, rightPart, modelP);

  // the six-argument TSLAE (u, v, w held, unused) solves the same system
  TRnRelease3dSpace &u = *new TRnRelease3dSpace(modelP), &v = *new TRnRelease3dSpace(modelP),
                    &w = *new TRnRelease3dSpace(modelP);
  TRnRelease3dSpace &p3 = *new TRnRelease3dSpace(modelP);
  p3.Assign(p2);
  TSLAE slae6(*APtr, p3, u, v, w, rightPart);
  slae6.SolveByBiConjugateGradientsStab();
  TSLAE slae3(*APtr, p3, rightPart);
  slae3.SolveByBiConjugateGradientsStab();
  double d63 = 0;
  for (int rn = 1; rn < RN + 1; rn++) {
    TRnSpace &r6 = const_cast<TRnSpace &>((const TRnSpace &)slae6.Result);
    TRnSpace &r3 = const_cast<TRnSpace &>((const TRnSpace &)slae3.Result);
    d63 = std::fmax(d63, std::fabs(r6[rn] - r3[rn]));
  }

  // the fused BiCGStab (a GPU extension of TSLAE) on the step-2 system
  TRnRelease3dSpace &p4 = *new TRnRelease3dSpace(modelP);
  p4.Assign(p);  // from the first step's pressure, not the step-2 solution
  TSLAE slae7(*APtr, p4, rightPart);
  slae7.SolveByFastBiConjugateGradientsStab();
  const int itFused = slae7.LastStats().iters;
  p4 |= slae7.Result;
  const double resFused = residual_max(*APtr, p4, rightPart, modelP);

  // errors the reference throws as const char*
  bool assignThrew = false, updateThrew = false;
  try {
    xRn.Assign(p);
  } catch (const char *) {
    assignThrew = true;
  }
  try {
    slae->UpdateA(ARn);  // an operator on the Rn model, not the pressure model
  } catch (const char *) {
    updateThrew = true;
  }
  EXPECT(assignThrew && updateThrew);

  // BiCGStab stops on max |r| <= 1e-5 (XSLAE.cpp:669, 712)
  EXPECT(resPack <= 1e-4 && resGrid <= 1e-4 && resStep2 <= 1e-4 && resFused <= 1e-4);
  EXPECT(itFused > 0);
  EXPECT(diff <= 1e-5 * scale);  // both stop at max |r| <= 1e-5: their fields differ by about that
  EXPECT(d63 == 0);
  printf("{\"unknowns\": %d, \"it_pack\": %d, \"it_grid\": %d, \"res_pack\": %.3e, \"res_grid\": %.3e, "
         "\"res_step2\": %.3e, \"res_fused\": %.3e, \"it_fused\": %d, \"pack_vs_grid\": %.3e, \"six_vs_three\": %.3e, "
         "\"failures\": %d}\n",
         unknowns, itPack, itGrid, resPack, resGrid, resStep2, resFused, itFused, scale > 0 ? diff / scale : diff, d63,
         g_failures);
  return g_failures ? 1 : 0;
}
