/*
 * nla3d_oracle.c -- CPU restatement of nla3d's element, material and dense
 * kernels for the Newton-iteration hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product.  Built with -ffp-contract=off so that every a*b+c rounds twice,
 * as the reference's x86-64 build without FMA does.
 *
 * Each routine follows one reference function statement by statement, in
 * the same loop and accumulation order; the citation is on each function.
 * What pins each routine is listed in nla3d_oracle.h.
 */
#include "nla3d_oracle.h"

#include <math.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Dense kernels, math/Mat.h.  Matrices are row-major; MatSym<N> is packed    */
/* upper triangle row by row (Mat.h:576-622).                                 */
/* ------------------------------------------------------------------------ */

/* R_sym += coef * B^T D B, B is M x N, D sym M, R sym N.  math/Mat.h:686-741 */
void orc_matBTDBprod(int M, int N, const double* B, const double* Dsym, double coef, double* Rsym) {
  double A[64 * 64];
  memset(A, 0, sizeof(double) * (size_t)N * (size_t)M);
  const double* Dp = Dsym;
  /* A = B^T * D, upper part of D including the diagonal */
  for (int k = 0; k < M; k++) {
    for (int j = k; j < M; j++) {
      for (int i = 0; i < N; i++) A[i * M + j] += B[k * N + i] * (*Dp);
      Dp++;
    }
  }
  /* A += B^T * (strict upper of D)^T */
  Dp = Dsym;
  for (int j = 0; j < M; j++) {
    Dp++; /* leave the diagonal */
    for (int k = j + 1; k < M; k++) {
      for (int i = 0; i < N; i++) A[i * M + j] += B[k * N + i] * (*Dp);
      Dp++;
    }
  }
  /* R = A * B, upper triangle */
  double* Rp = Rsym;
  for (int i = 0; i < N; i++) {
    for (int j = i; j < N; j++) {
      for (int k = 0; k < M; k++) *Rp += A[i * M + k] * B[k * N + j] * coef;
      Rp++;
    }
  }
}

/* R += coef * B^T V, B is M x N.  math/Mat.h:744-756 */
void orc_matBTVprod(int M, int N, const double* B, const double* V, double coef, double* R) {
  for (int i = 0; i < N; i++)
    for (int j = 0; j < M; j++) R[i] += B[j * N + i] * V[j] * coef;
}

/* R += coef * B V, B is M x N.  math/Mat.h:758-770 */
void orc_matBVprod(int M, int N, const double* B, const double* V, double coef, double* R) {
  for (int i = 0; i < M; i++)
    for (int j = 0; j < N; j++) R[i] += B[i * N + j] * V[j] * coef;
}

/* R += coef * A B, A is M1 x N1, B is N1 x N2.  math/Mat.h:784-801 */
void orc_matABprod(int M1, int N1, int N2, const double* A, const double* B, double coef, double* R) {
  double* Rp = R;
  for (int i = 0; i < M1; i++) {
    for (int j = 0; j < N2; j++) {
      for (int k = 0; k < N1; k++) *Rp += A[i * N1 + k] * B[k * N2 + j] * coef;
      Rp++;
    }
  }
}

/* R += coef * A^T B, A is M1 x N1, B is M1 x N2.  math/Mat.h:804-822 */
void orc_matATBprod(int M1, int N1, int N2, const double* A, const double* B, double coef, double* R) {
  double* Rp = R;
  for (int i = 0; i < N1; i++) {
    for (int j = 0; j < N2; j++) {
      for (int k = 0; k < M1; k++) *Rp += A[k * N1 + i] * B[k * N2 + j] * coef;
      Rp++;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* 3x3 determinant and inverse, math/Mat.cpp:25-46                           */
/* ------------------------------------------------------------------------ */
static double det3(const double d[3][3]) {
  return d[0][0] * (d[1][1] * d[2][2] - d[1][2] * d[2][1]) -
         d[0][1] * (d[1][0] * d[2][2] - d[1][2] * d[2][0]) +
         d[0][2] * (d[1][0] * d[2][1] - d[1][1] * d[2][0]);
}

static void inv3(const double d[3][3], double det, double r[3][3]) {
  double t[9] = {d[1][1] * d[2][2] - d[1][2] * d[2][1], d[0][2] * d[2][1] - d[0][1] * d[2][2],
                 d[0][1] * d[1][2] - d[0][2] * d[1][1], d[2][0] * d[1][2] - d[1][0] * d[2][2],
                 d[0][0] * d[2][2] - d[0][2] * d[2][0], d[1][0] * d[0][2] - d[0][0] * d[1][2],
                 d[1][0] * d[2][1] - d[1][1] * d[2][0], d[0][1] * d[2][0] - d[0][0] * d[2][1],
                 d[0][0] * d[1][1] - d[0][1] * d[1][0]};
  double s = 1.0 / det;
  for (int i = 0; i < 9; i++) r[i / 3][i % 3] = t[i] * s;
}

/* ------------------------------------------------------------------------ */
/* Isoparametric hexahedron, elements/isoparametric.cpp:140-188, 211-221;    */
/* Gauss table elements/isoparametric.h:126-136                               */
/* ------------------------------------------------------------------------ */
static const double GP2 = 0.5773502691896258;

static void hex8_dN(double r, double s, double t, double dN[8][3]) {
  const double v[24] = {
      -0.125 * (1.0 - s) * (1.0 - t), -0.125 * (1.0 - r) * (1.0 - t), -0.125 * (1.0 - r) * (1.0 - s),
      +0.125 * (1.0 - s) * (1.0 - t), -0.125 * (1.0 + r) * (1.0 - t), -0.125 * (1.0 + r) * (1.0 - s),
      +0.125 * (1.0 + s) * (1.0 - t), +0.125 * (1.0 + r) * (1.0 - t), -0.125 * (1.0 + r) * (1.0 + s),
      -0.125 * (1.0 + s) * (1.0 - t), +0.125 * (1.0 - r) * (1.0 - t), -0.125 * (1.0 - r) * (1.0 + s),
      -0.125 * (1.0 - s) * (1.0 + t), -0.125 * (1.0 - r) * (1.0 + t), +0.125 * (1.0 - r) * (1.0 - s),
      +0.125 * (1.0 - s) * (1.0 + t), -0.125 * (1.0 + r) * (1.0 + t), +0.125 * (1.0 + r) * (1.0 - s),
      +0.125 * (1.0 + s) * (1.0 + t), +0.125 * (1.0 + r) * (1.0 + t), +0.125 * (1.0 + r) * (1.0 + s),
      -0.125 * (1.0 + s) * (1.0 + t), +0.125 * (1.0 - r) * (1.0 + t), +0.125 * (1.0 - r) * (1.0 + s)};
  for (int i = 0; i < 24; i++) dN[i / 3][i % 3] = v[i];
}

int orc_hex8_jacob(const double* X, double* NiXj, double* det) {
  memset(NiXj, 0, sizeof(double) * 8 * 8 * 3);
  for (int np = 0; np < 8; np++) {
    /* table order: r fastest, then s, then t (isoparametric.h:126-136) */
    double r = (np & 1) ? GP2 : -GP2;
    double s = (np & 2) ? GP2 : -GP2;
    double t = (np & 4) ? GP2 : -GP2;
    double dN[8][3];
    hex8_dN(r, s, t, dN);
    double J[3][3] = {{0}};
    for (int nod = 0; nod < 8; nod++)
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) J[i][j] += dN[nod][i] * X[nod * 3 + j];
    det[np] = det3(J);
    double Jinv[3][3];
    inv3(J, det[np], Jinv);
    for (int nod = 0; nod < 8; nod++)
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) NiXj[(np * 8 + nod) * 3 + i] += Jinv[i][j] * dN[nod][j];
  }
  return 8;
}

/* quad: elements/isoparametric.cpp:55-106, 123-128; table isoparametric.h:82-87 */
int orc_quad4_jacob(const double* X, double* NiXj, double* det) {
  memset(NiXj, 0, sizeof(double) * 4 * 4 * 2);
  for (int np = 0; np < 4; np++) {
    double r = (np & 1) ? GP2 : -GP2;
    double s = (np & 2) ? GP2 : -GP2;
    double dN[4][2] = {{-0.25 * (1.0 - s), -0.25 * (1.0 - r)},
                       {+0.25 * (1.0 - s), -0.25 * (1.0 + r)},
                       {+0.25 * (1.0 + s), +0.25 * (1.0 + r)},
                       {-0.25 * (1.0 + s), +0.25 * (1.0 - r)}};
    double J[2][2] = {{0}};
    for (int nod = 0; nod < 4; nod++)
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) J[i][j] += dN[nod][i] * X[nod * 3 + j];
    det[np] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    double s1 = 1.0 / det[np];
    double Jinv[2][2] = {{J[1][1] * s1, -J[0][1] * s1}, {-J[1][0] * s1, J[0][0] * s1}};
    for (int nod = 0; nod < 4; nod++)
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) NiXj[(np * 4 + nod) * 2 + i] += Jinv[i][j] * dN[nod][j];
  }
  return 4;
}

/* ------------------------------------------------------------------------ */
/* solidmech helpers, solidmech.cpp:10-42                                     */
/* ------------------------------------------------------------------------ */
static double J_C(const double* C) {
  return sqrt(C[0] * (C[3] * C[5] - C[4] * C[4]) - C[1] * (C[1] * C[5] - C[4] * C[2]) +
              C[2] * (C[1] * C[4] - C[3] * C[2]));
}

static void invC_C(const double* C, double J, double* iC) {
  double oo = 1.0 / (J * J);
  iC[0] = oo * (C[3] * C[5] - C[4] * C[4]);
  iC[1] = oo * (C[2] * C[4] - C[1] * C[5]);
  iC[2] = oo * (C[1] * C[4] - C[2] * C[3]);
  iC[3] = oo * (C[0] * C[5] - C[2] * C[2]);
  iC[4] = oo * (C[1] * C[2] - C[0] * C[4]);
  iC[5] = oo * (C[0] * C[3] - C[1] * C[1]);
}

static void IC_C(const double* C, double* IC) {
  IC[0] = C[0] + C[3] + C[5];
  IC[1] = C[0] * C[3] + C[3] * C[5] + C[0] * C[5] - C[1] * C[1] - C[4] * C[4] - C[2] * C[2];
  IC[2] = J_C(C);
}

/* ------------------------------------------------------------------------ */
/* Hyperelastic materials, materials/materials_hyperelastic.cpp              */
/* ------------------------------------------------------------------------ */
static const double Ivec[6] = {1.0, 0.0, 0.0, 1.0, 0.0, 1.0};       /* material.cpp:9 */
static const double II[6][6] = {{1.0, 0.0, 0.0, 0.0, 0.0, 0.0},     /* materials_hyperelastic.cpp:9-14 */
                                {0.0, 0.5, 0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.5, 0.0, 0.0, 0.0},
                                {0.0, 0.0, 0.0, 1.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0, 0.5, 0.0},
                                {0.0, 0.0, 0.0, 0.0, 0.0, 1.0}};

/* getK(): :151-153, :176-178, :201-203 */
double orc_mat_K(int type, const double* MC) {
  switch (type) {
    case ORC_MAT_NEO_HOOKEAN: return MC[1];
    case ORC_MAT_BIDERMAN: return MC[4];
    default: return MC[2];
  }
}

/* W_first/second_derivatives: NH :135-149, Biderman :161-174, MR :184-198 */
static void mat_alpha(int type, const double* MC, double I1, double* alpha, int second) {
  switch (type) {
    case ORC_MAT_NEO_HOOKEAN:
      alpha[0] = 0.5 * MC[0];
      alpha[1] = 0.0;
      if (second) { alpha[2] = 0.0; alpha[4] = 0.0; alpha[3] = 0.0; }
      break;
    case ORC_MAT_BIDERMAN:
      alpha[0] = MC[0] + 2 * MC[1] * (I1 - 3.0) + 3 * MC[2] * (I1 - 3.0) * (I1 - 3.0);
      alpha[1] = MC[3];
      if (second) { alpha[2] = 2 * MC[1] + 6 * MC[2] * (I1 - 3.0); alpha[4] = 0.0; alpha[3] = 0.0; }
      break;
    default:
      alpha[0] = MC[0];
      alpha[1] = MC[1];
      if (second) { alpha[2] = 0.0; alpha[4] = 0.0; alpha[3] = 0.0; }
      break;
  }
}

/* Mat_Hyper_Isotrop_General::getS_UP, :30-62 */
void orc_getS_UP(int type, const double* MC, const double* C, double press, double* S) {
  double alpha[2];
  double IC[3];
  IC_C(C, IC);
  double _13I1C = 1.0 / 3.0 * IC[0];
  double _23I2C = 2.0 / 3.0 * IC[1];
  double J = IC[2];
  double pp = pow(J, -2.0 / 3.0);
  double pppp = pp * pp;
  double C_inv[6];
  invC_C(C, J, C_inv);
  double a5[5];
  mat_alpha(type, MC, IC[0] * pp, a5, 0);
  alpha[0] = a5[0];
  alpha[1] = a5[1];
  double ko1 = 2 * alpha[0] * pp;
  double ko2 = 2 * alpha[1] * pppp;
  for (int i = 0; i < 6; i++) {
    double A = Ivec[i] - _13I1C * C_inv[i];
    double B = IC[0] * Ivec[i] - C[i] - _23I2C * C_inv[i];
    S[i] = ko1 * A + ko2 * B + press * J * C_inv[i];
  }
}

/* Mat_Hyper_Isotrop_General::getDdDp_UP, :65-129 */
void orc_getDdDp_UP(int type, const double* MC, const double* C, double press, double* Dd, double* Dp) {
  double alpha[5];
  double IC[3];
  IC_C(C, IC);
  double _13I1C = 1.0 / 3.0 * IC[0];
  double _23I2C = 2.0 / 3.0 * IC[1];
  double J = IC[2];
  double pp = pow(J, -2.0 / 3.0);
  double oo = pow(J, -2);
  double pppp = pp * pp;
  double Ci[6];
  invC_C(C, J, Ci);
  mat_alpha(type, MC, IC[0] * pp, alpha, 1);
  double A[6], B[6];
  for (int i = 0; i < 6; i++) {
    A[i] = Ivec[i] - _13I1C * Ci[i];
    B[i] = IC[0] * Ivec[i] - C[i] - _23I2C * Ci[i];
  }
  /* :97-108 */
  const double IIt[6][6] = {
      {-Ci[0] * Ci[0], -Ci[0] * Ci[1], -Ci[0] * Ci[2], -Ci[1] * Ci[1], -Ci[1] * Ci[2], -Ci[2] * Ci[2]},
      {-Ci[0] * Ci[1], -0.5 * (Ci[0] * Ci[3] + Ci[1] * Ci[1]), -0.5 * (Ci[0] * Ci[4] + Ci[2] * Ci[1]),
       -Ci[1] * Ci[3], -0.5 * (Ci[1] * Ci[4] + Ci[2] * Ci[3]), -Ci[2] * Ci[4]},
      {-Ci[0] * Ci[2], -0.5 * (Ci[0] * Ci[4] + Ci[1] * Ci[2]), -0.5 * (Ci[0] * Ci[5] + Ci[2] * Ci[2]),
       -Ci[1] * Ci[4], -0.5 * (Ci[1] * Ci[5] + Ci[2] * Ci[4]), -Ci[2] * Ci[5]},
      {-Ci[1] * Ci[1], -Ci[1] * Ci[3], -Ci[1] * Ci[4], -Ci[3] * Ci[3], -Ci[3] * Ci[4], -Ci[4] * Ci[4]},
      {-Ci[1] * Ci[2], -0.5 * (Ci[1] * Ci[4] + Ci[3] * Ci[2]), -0.5 * (Ci[1] * Ci[5] + Ci[4] * Ci[2]),
       -Ci[3] * Ci[4], -0.5 * (Ci[3] * Ci[5] + Ci[4] * Ci[4]), -Ci[4] * Ci[5]},
      {-Ci[2] * Ci[2], -Ci[2] * Ci[4], -Ci[2] * Ci[5], -Ci[4] * Ci[4], -Ci[4] * Ci[5], -Ci[5] * Ci[5]}};
  int ind = 0;
  for (int i = 0; i < 6; i++) {
    for (int j = i; j < 6; j++) {
      Dd[ind] = 4.0 * alpha[2] * pppp * A[i] * A[j] + 4.0 * alpha[3] * oo * (B[i] * A[j] + A[i] * B[j]) +
                4.0 * alpha[4] * pppp * pppp * B[i] * B[j] -
                4.0 / 3.0 * alpha[0] * pp *
                    (Ci[i] * A[j] + A[i] * Ci[j] + _13I1C * Ci[i] * Ci[j] + IC[0] * IIt[i][j]) -
                8.0 / 3.0 * alpha[1] * pppp *
                    (Ci[i] * B[j] + B[i] * Ci[j] + _23I2C * Ci[i] * Ci[j] - 3.0 / 2.0 * Ivec[i] * Ivec[j] +
                     3.0 / 2.0 * II[i][j] + IC[1] * IIt[i][j]) +
                press * J * (Ci[i] * Ci[j] + 2 * IIt[i][j]);
      Dd[ind] = Dd[ind] / 2.0;
      ind++;
    }
    Dp[i] = J * Ci[i];
  }
}

/* ------------------------------------------------------------------------ */
/* SOLID81, elements/SOLID81.cpp                                              */
/* ------------------------------------------------------------------------ */

/* make_B_NL, :136-152: 9 x 24 */
static void make_B_NL(const double* N /*[8][3]*/, double* B) {
  for (int i = 0; i < 8; i++) {
    B[0 * 24 + (i * 3 + 0)] += N[i * 3 + 0];
    B[1 * 24 + (i * 3 + 0)] += N[i * 3 + 1];
    B[2 * 24 + (i * 3 + 0)] += N[i * 3 + 2];
    B[3 * 24 + (i * 3 + 1)] += N[i * 3 + 0];
    B[4 * 24 + (i * 3 + 1)] += N[i * 3 + 1];
    B[5 * 24 + (i * 3 + 1)] += N[i * 3 + 2];
    B[6 * 24 + (i * 3 + 2)] += N[i * 3 + 0];
    B[7 * 24 + (i * 3 + 2)] += N[i * 3 + 1];
    B[8 * 24 + (i * 3 + 2)] += N[i * 3 + 2];
  }
}

/* make_B_L, :118-133: 6 x 24 */
static void make_B_L(const double* N, double* B) {
  for (int i = 0; i < 8; i++) {
    B[0 * 24 + (i * 3 + 0)] += 2 * N[i * 3 + 0];
    B[1 * 24 + (i * 3 + 0)] += 2 * N[i * 3 + 1];
    B[1 * 24 + (i * 3 + 1)] += 2 * N[i * 3 + 0];
    B[2 * 24 + (i * 3 + 0)] += 2 * N[i * 3 + 2];
    B[2 * 24 + (i * 3 + 2)] += 2 * N[i * 3 + 0];
    B[3 * 24 + (i * 3 + 1)] += 2 * N[i * 3 + 1];
    B[4 * 24 + (i * 3 + 1)] += 2 * N[i * 3 + 2];
    B[4 * 24 + (i * 3 + 2)] += 2 * N[i * 3 + 1];
    B[5 * 24 + (i * 3 + 2)] += 2 * N[i * 3 + 2];
  }
}

/* make_S, :156-176: MatSym<9> with S on the three 3x3 diagonal blocks */
static void make_S(const double* S, double* Sp) {
  Sp[0] += S[0];  Sp[1] += S[1];  Sp[2] += S[2];
  Sp[9] += S[3];  Sp[10] += S[4]; Sp[17] += S[5];
  Sp[24] += S[0]; Sp[25] += S[1]; Sp[26] += S[2];
  Sp[30] += S[3]; Sp[31] += S[4]; Sp[35] += S[5];
  Sp[39] += S[0]; Sp[40] += S[1]; Sp[41] += S[2];
  Sp[42] += S[3]; Sp[43] += S[4]; Sp[44] += S[5];
}

/* make_Omega, :179-191: 6 x 9 */
static void make_Omega(const double* O, double* Om) {
  for (int i = 0; i < 3; i++) {
    Om[0 * 9 + (i * 3 + 0)] = O[0 + i * 3];
    Om[1 * 9 + (i * 3 + 0)] = O[1 + i * 3];
    Om[1 * 9 + (i * 3 + 1)] = O[0 + i * 3];
    Om[2 * 9 + (i * 3 + 0)] = O[2 + i * 3];
    Om[2 * 9 + (i * 3 + 2)] = O[0 + i * 3];
    Om[3 * 9 + (i * 3 + 1)] = O[1 + i * 3];
    Om[4 * 9 + (i * 3 + 1)] = O[2 + i * 3];
    Om[4 * 9 + (i * 3 + 2)] = O[1 + i * 3];
    Om[5 * 9 + (i * 3 + 2)] = O[2 + i * 3];
  }
}

/* ElementSOLID81::update, :85-115 */
void orc_solid81_update(const double* NiXj, const double* U, double p_e, int type, const double* MC,
                        double* O, double* C, double* S) {
  for (int np = 0; np < 8; np++) {
    double B_NL[9 * 24];
    memset(B_NL, 0, sizeof(B_NL));
    make_B_NL(NiXj + np * 24, B_NL);
    double* o = O + np * 9;
    for (int i = 0; i < 9; i++) o[i] = 0.0;
    orc_matBVprod(9, 24, B_NL, U, 1.0, o);
    double* c = C + np * 6;
    c[0] = 1.0 + 2 * o[0] + o[0] * o[0] + o[3] * o[3] + o[6] * o[6];
    c[3] = 1.0 + 2 * o[4] + o[1] * o[1] + o[4] * o[4] + o[7] * o[7];
    c[5] = 1.0 + 2 * o[8] + o[2] * o[2] + o[5] * o[5] + o[8] * o[8];
    c[1] = o[1] + o[3] + o[0] * o[1] + o[3] * o[4] + o[6] * o[7];
    c[4] = o[5] + o[7] + o[1] * o[2] + o[4] * o[5] + o[7] * o[8];
    c[2] = o[2] + o[6] + o[0] * o[2] + o[3] * o[5] + o[6] * o[8];
    orc_getS_UP(type, MC, c, p_e, S + np * 6);
  }
}

/* ElementSOLID81::buildK, :29-82 (without the scatter, whose order is kept in out[]) */
void orc_solid81_buildK(const double* NiXj, const double* det, const double* O, const double* C,
                        const double* S, double p_e, int type, const double* MC, double* out) {
  double Kpp = 0.0, Fp = 0.0;
  double Kup[24] = {0}, Fu[24] = {0};
  double Kuu[300];
  memset(Kuu, 0, sizeof(Kuu));
  double k = orc_mat_K(type, MC);
  for (int np = 0; np < 8; np++) {
    double dWt = 1.0 * det[np]; /* intWeight: w * det, isoparametric.h:303-308 */
    double Dd[21], Dp[6];
    orc_getDdDp_UP(type, MC, C + np * 6, p_e, Dd, Dp);
    double J = J_C(C + np * 6);
    double matB[6 * 24], matS[45], matO[6 * 9], matB_NL[9 * 24];
    memset(matB, 0, sizeof(matB));
    memset(matS, 0, sizeof(matS));
    memset(matO, 0, sizeof(matO));
    memset(matB_NL, 0, sizeof(matB_NL));
    const double* N = NiXj + np * 24;
    make_B_L(N, matB);
    make_S(S + np * 6, matS);
    make_Omega(O + np * 9, matO);
    make_B_NL(N, matB_NL);
    orc_matABprod(6, 9, 24, matO, matB_NL, 2.0, matB);
    orc_matBTDBprod(6, 24, matB, Dd, 0.5 * dWt, Kuu);
    orc_matBTDBprod(9, 24, matB_NL, matS, dWt, Kuu);
    orc_matBTVprod(6, 24, matB, S + np * 6, -0.5 * dWt, Fu);
    orc_matBTVprod(6, 24, matB, Dp, 0.5 * dWt, Kup);
    Fp += -(J - 1 - p_e / k) * dWt;
    Kpp += -1.0 / k * dWt;
  }
  memcpy(out, Kuu, sizeof(Kuu));
  memcpy(out + 300, Kup, sizeof(Kup));
  out[324] = Kpp;
  memcpy(out + 325, Fu, sizeof(Fu));
  out[349] = Fp;
}

void orc_solid81_elements(int ne, const double* Xe, const double* Ue, const double* pe, int type,
                          const double* MC, double* out) {
  for (int e = 0; e < ne; e++) {
    double NiXj[8 * 8 * 3], det[8], O[72], C[48], S[48];
    orc_hex8_jacob(Xe + (size_t)e * 24, NiXj, det);
    orc_solid81_update(NiXj, Ue + (size_t)e * 24, pe[e], type, MC, O, C, S);
    orc_solid81_buildK(NiXj, det, O, C, S, pe[e], type, MC, out + (size_t)e * 350);
  }
}

/* ------------------------------------------------------------------------ */
/* QUADTH and SurfaceLINETH, elements/QUADTH.cpp                              */
/* ------------------------------------------------------------------------ */

/* ElementQUADTH::buildK :22-48 and buildC :51-67 */
void orc_quadth_element(const double* X, double k, double rho, double c, double volFlux, double* Ke,
                        double* Fe, double* Ce) {
  double NiXj[4 * 4 * 2], det[4];
  orc_quad4_jacob(X, NiXj, det);
  for (int i = 0; i < 10; i++) Ke[i] = Ce[i] = 0.0;
  for (int i = 0; i < 4; i++) Fe[i] = 0.0;
  double mat_k[3] = {k, 0.0, k};
  for (int np = 0; np < 4; np++) {
    double r = (np & 1) ? GP2 : -GP2;
    double s = (np & 2) ? GP2 : -GP2;
    double dWt = 1.0 * det[np];
    const double* N = NiXj + np * 8;
    double B[8] = {N[0], N[2], N[4], N[6], N[1], N[3], N[5], N[7]}; /* make_B, :69-72 */
    orc_matBTDBprod(2, 4, B, mat_k, dWt, Ke);
    double ff[4] = {0.25 * (1.0 - r) * (1.0 - s), 0.25 * (1.0 + r) * (1.0 - s),
                    0.25 * (1.0 + r) * (1.0 + s), 0.25 * (1.0 - r) * (1.0 + s)};
    if (volFlux != 0.0)
      for (int i = 0; i < 4; i++) Fe[i] += ff[i] * (volFlux * dWt);
    int idx = 0;
    for (int i = 0; i < 4; i++)
      for (int j = i; j < 4; j++) Ce[idx++] += ff[i] * ff[j] * rho * c * dWt;
  }
}

/* ------------------------------------------------------------------------ */
/* PLANE41 (plane-strain u/p quadrilateral), elements/PLANE41.cpp             */
/* ------------------------------------------------------------------------ */

/* r = a (ma x na) * b (na x nb), the loop order of Mat operator* (Mat.h:198-213) */
static void mat_mul(int ma, int na, int nb, const double* a, const double* b, double* r) {
  for (int i = 0; i < ma; i++)
    for (int j = 0; j < nb; j++) {
      double e = 0.0;
      for (int l = 0; l < na; l++) e += a[i * na + l] * b[l * nb + j];
      r[i * nb + j] = e;
    }
}

static void transpose(int m, int n, const double* a, double* t) {
  for (int i = 0; i < m; i++)
    for (int j = 0; j < n; j++) t[j * m + i] = a[i * n + j];
}

/* update() at U (:120-150) then buildK() (:31-104) of one element: out[0..35] the node-node
 * entries in assembleK's order (PLANE41.h:62-70: i <= j nodes, di, dj, the diagonal node
 * block's lower entries), out[36..43] node-pressure, out[44] pressure-pressure, out[45..53]
 * Fe = [-Qe (8), -Fp].  X: 4 nodes x 3 coordinates; U: [UX1 UY1 ... UX4 UY4]. */
void orc_plane41_element(const double* X, const double* U, double p_e, int type, const double* MC, double* out) {
  double NiXj[4 * 4 * 2], det[4];
  orc_quad4_jacob(X, NiXj, det);
  const int comp[3] = {0, 3, 1}; /* M_XX, M_YY, M_XY (PLANE41.cpp:11) */
  double k = orc_mat_K(type, MC);
  double Kuu[64] = {0}, Kup[8] = {0}, Qe[8] = {0}, Kpp = 0.0, Fp = 0.0;
  for (int np = 0; np < 4; np++) {
    const double* N = NiXj + np * 8; /* N[nod*2 + 0] = dN/dx, [1] = dN/dy */
    double Bom[32] = {0};            /* make_Bomega :113-118, 4 x 8 */
    for (int nod = 0; nod < 4; nod++) {
      Bom[0 * 8 + 2 * nod] = N[nod * 2 + 0];
      Bom[1 * 8 + 2 * nod] = N[nod * 2 + 1];
      Bom[2 * 8 + 2 * nod + 1] = N[nod * 2 + 0];
      Bom[3 * 8 + 2 * nod + 1] = N[nod * 2 + 1];
    }
    /* update (:136-149) */
    double O[4];
    for (int i = 0; i < 4; i++) {
      double e = 0.0;
      for (int l = 0; l < 8; l++) e += Bom[i * 8 + l] * U[l];
      O[i] = e;
    }
    double Cp[3];
    Cp[0] = 1.0 + 2 * O[0] + 1.0 * (O[0] * O[0] + O[2] * O[2]);
    Cp[1] = 1.0 + 2 * O[3] + 1.0 * (O[3] * O[3] + O[1] * O[1]);
    Cp[2] = O[1] + O[2] + O[0] * O[1] + O[2] * O[3];
    double CV[6] = {Cp[0], Cp[2], 0.0, Cp[1], 0.0, 1.0};
    double S6[6];
    orc_getS_UP(type, MC, CV, p_e, S6);
    double S[3] = {S6[comp[0]], S6[comp[1]], S6[comp[2]]};
    /* buildK (:50-87) */
    double dWt = 1.0 * det[np];
    double Dd6[21], Dp6[6];
    orc_getDdDp_UP(type, MC, CV, p_e, Dd6, Dp6);
    double E[9];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        int i = comp[a] < comp[b] ? comp[a] : comp[b], j = comp[a] < comp[b] ? comp[b] : comp[a];
        E[a * 3 + b] = Dd6[i * 6 - i * (i - 1) / 2 + (j - i)]; /* packed upper, row i */
      }
    double Ep[3] = {Dp6[comp[0]], Dp6[comp[1]], Dp6[comp[2]]};
    double J = J_C(CV);
    double B[24] = {0}; /* make_B :106-111, 3 x 8 */
    for (int nod = 0; nod < 4; nod++) {
      B[0 * 8 + 2 * nod] = N[nod * 2 + 0];
      B[1 * 8 + 2 * nod + 1] = N[nod * 2 + 1];
      B[2 * 8 + 2 * nod] = N[nod * 2 + 1];
      B[2 * 8 + 2 * nod + 1] = N[nod * 2 + 0];
    }
    double Sm[16] = {S[0], S[2], 0.0, 0.0, S[2], S[1], 0.0, 0.0, 0.0, 0.0, S[0], S[2], 0.0, 0.0, S[2], S[1]};
    double Om[12] = {O[0], 0.0, O[2], 0.0, 0.0, O[1], 0.0, O[3], O[1], O[0], O[3], O[2]};
    double Bl[24];
    mat_mul(3, 4, 8, Om, Bom, Bl);
    for (int i = 0; i < 24; i++) B[i] += Bl[i];
    double Bt[24], BtE[24], BtEB[64], Bomt[32], BomtS[32], BomtSBom[64];
    transpose(3, 8, B, Bt);
    mat_mul(8, 3, 3, Bt, E, BtE);
    mat_mul(8, 3, 8, BtE, B, BtEB);
    transpose(4, 8, Bom, Bomt);
    mat_mul(8, 4, 4, Bomt, Sm, BomtS);
    mat_mul(8, 4, 8, BomtS, Bom, BomtSBom);
    for (int i = 0; i < 64; i++) Kuu[i] += (BtEB[i] * 2.0 + BomtSBom[i]) * dWt;
    Fp += (J - 1 - p_e / k) * dWt;
    for (int i = 0; i < 8; i++) {
      double e = 0.0;
      for (int l = 0; l < 3; l++) e += Bt[i * 3 + l] * S[l];
      Qe[i] += e * dWt;
    }
    for (int i = 0; i < 8; i++) {
      double e = 0.0;
      for (int l = 0; l < 3; l++) e += Bt[i * 3 + l] * Ep[l];
      Kup[i] += e * dWt;
    }
    Kpp -= 1.0 / k * dWt;
  }
  /* assembleK order (PLANE41.h:55-89) */
  int s = 0;
  for (int i = 0; i < 4; i++)
    for (int j = i; j < 4; j++)
      for (int di = 0; di < 2; di++)
        for (int dj = 0; dj < 2; dj++) {
          if (i == j && dj > di) continue;
          out[s++] = Kuu[(i * 2 + di) * 8 + j * 2 + dj];
        }
  for (int i = 0; i < 8; i++) out[s++] = Kup[i];
  out[s++] = Kpp;
  for (int i = 0; i < 8; i++) out[s++] = -Qe[i];
  out[s++] = -Fp;
}

/* SurfaceLINETH::pre :79-91 and buildK :94-129; line makeJacob isoparametric.cpp:13-27 */
void orc_lineth_element(const double* X, double flux, double htc, double etemp0, double etemp1,
                        double* Ke, double* Fe) {
  double dx = X[3] - X[0], dy = X[4] - X[1], dz = X[5] - X[2];
  double det = sqrt(dx * dx + dy * dy + dz * dz) * 0.5;
  if (etemp1 == 0.0) etemp1 = etemp0;
  Ke[0] = Ke[1] = Ke[2] = 0.0;
  Fe[0] = Fe[1] = 0.0;
  const double rr[2] = {-GP2, GP2};
  if (flux != 0.0) {
    for (int np = 0; np < 2; np++) {
      double dWt = 1.0 * det;
      double Ns[2] = {0.5 * (1.0 - rr[np]), 0.5 * (1.0 + rr[np])};
      Fe[0] += Ns[0] * (flux * dWt);
      Fe[1] += Ns[1] * (flux * dWt);
    }
  }
  if (htc != 0.0) {
    for (int np = 0; np < 2; np++) {
      double dWt = 1.0 * det;
      double Ns[2] = {0.5 * (1.0 - rr[np]), 0.5 * (1.0 + rr[np])};
      double t00 = Ns[0] * Ns[0] * htc * dWt;
      double t01 = Ns[0] * Ns[1] * htc * dWt;
      double t11 = Ns[1] * Ns[1] * htc * dWt;
      Ke[0] += t00; Ke[1] += t01; Ke[2] += t11;
      Fe[0] += t00 * etemp0 * 1.0;
      Fe[0] += t01 * etemp1 * 1.0;
      Fe[1] += t01 * etemp0 * 1.0;
      Fe[1] += t11 * etemp1 * 1.0;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* TimeControl::nextStep, FESolver.cpp:34-56, called repeatedly with one    */
/* delta; NonlinearFESolver passes delta = (end - start) / N (FESolver.cpp:439) */
/* ------------------------------------------------------------------------ */
int orc_timecontrol(double start, double end, double delta, double* times, int max_steps) {
  double cur = 0.0; /* currentTime starts at 0.0 (FESolver.h:55) */
  (void)start;
  int n = 0;
  for (;;) {
    if (cur >= end) break;
    double d;
    if (cur + delta >= end) d = end - cur;
    else if (cur + 2 * delta >= end) d = (end - cur) / 2.0;
    else d = delta;
    cur += d;
    if (n < max_steps) times[n] = cur;
    n++;
  }
  return n;
}

/* ------------------------------------------------------------------------ */
/* Reference assembly loop for SOLID81 (CPU baseline and checker):            */
/* FEStorage::assembleGlobalEqMatrices element loop (FEStorage.cpp:33-38)     */
/* with ElementSOLID81::buildK (SOLID81.cpp:29-82, after update :85-115) and  */
/* assemble3 (SOLID81.h:54-94) -> BlockSparseSymMatrix::addValue              */
/* (BlockSparseMatrix.h:155-168) -> SparsityInfo::getIndex binary search      */
/* (SparseMatrix.cpp:117-154).  makeJacob is called per element here (the     */
/* reference caches it from pre()).                                           */
/* ------------------------------------------------------------------------ */
static unsigned getIndex(const unsigned* iofeir, const unsigned* cols, unsigned i, unsigned j) {
  unsigned st = iofeir[i - 1] - 1, en = iofeir[i] - 1;
  if (st == en) return 0xFFFFFFFFu;
  en--;
  for (;;) {
    if (en - st == 1) {
      if (cols[st] == j) return st;
      if (cols[en] == j) return en;
      return 0xFFFFFFFFu;
    }
    unsigned ind = (unsigned)((en + st) * 0.5);
    if (cols[ind] == j) return ind;
    if (en == st) return 0xFFFFFFFFu;
    if (cols[ind] > j) en = ind;
    else st = ind;
  }
}

static int addValue(unsigned i, unsigned j, double v, unsigned nC, const unsigned* io1, const unsigned* c1,
                    double* v1, const unsigned* io12, const unsigned* c12, double* v12, const unsigned* io2,
                    const unsigned* c2, double* v2) {
  if (i > j) { unsigned t = i; i = j; j = t; }
  unsigned idx;
  if (j <= nC) {
    idx = getIndex(io1, c1, i, j);
    if (idx == 0xFFFFFFFFu) return -1;
    v1[idx] += v;
  } else if (i <= nC) {
    idx = getIndex(io12, c12, i, j - nC);
    if (idx == 0xFFFFFFFFu) return -1;
    v12[idx] += v;
  } else {
    idx = getIndex(io2, c2, i - nC, j - nC);
    if (idx == 0xFFFFFFFFu) return -1;
    v2[idx] += v;
  }
  return 0;
}

int orc_solid81_assemble_ref(int ne, const double* Xe, const double* Ue, const double* pe, int type,
                             const double* MC, const unsigned* eq, unsigned nC, const unsigned* io1,
                             const unsigned* c1, double* v1, const unsigned* io12, const unsigned* c12,
                             double* v12, const unsigned* io2, const unsigned* c2, double* v2, double* F) {
  for (int e = 0; e < ne; e++) {
    double NiXj[8 * 8 * 3], det[8], O[72], C[48], S[48], out[350];
    orc_hex8_jacob(Xe + (size_t)e * 24, NiXj, det);
    orc_solid81_update(NiXj, Ue + (size_t)e * 24, pe[e], type, MC, O, C, S);
    orc_solid81_buildK(NiXj, det, O, C, S, pe[e], type, MC, out);
    const unsigned* q = eq + (size_t)e * 25;
    const double* kp = out;
    for (int i = 0; i < 8; i++)
      for (int di = 0; di < 3; di++)
        for (int j = i; j < 8; j++)
          for (int dj = 0; dj < 3; dj++) {
            if (i == j && dj < di) continue;
            if (addValue(q[3 * i + di], q[3 * j + dj], *kp, nC, io1, c1, v1, io12, c12, v12, io2, c2, v2)) return -1;
            kp++;
          }
    for (int a = 0; a < 24; a++)
      if (addValue(q[a], q[24], out[300 + a], nC, io1, c1, v1, io12, c12, v12, io2, c2, v2)) return -1;
    if (addValue(q[24], q[24], out[324], nC, io1, c1, v1, io12, c12, v12, io2, c2, v2)) return -1;
    for (int a = 0; a < 24; a++) F[q[a] - 1] += out[325 + a];
    F[q[24] - 1] += out[349];
  }
  return 0;
}
