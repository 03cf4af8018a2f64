/*
 * mndo_oracle.c -- TEST INFRASTRUCTURE ONLY (checker + CPU baseline).
 *
 * Plain-C FP64 restatement of the reference MNDO SCF path.  Every numeric
 * routine keeps the reference's operation order, so that with gcc -O2 on
 * x86-64 (no FMA contraction) its results agree bit for bit with the compiled
 * reference; tests/test_oracle_golden.py checks exactly that against fixtures
 * produced by oracle/_ref/ref_dump.  The only deliberate departures are
 *   - or_eigh, a textbook Householder + implicit-QR eigensolver standing in
 *     for LAPACK dsyev (eigenvectors agree up to sign / rotation within
 *     degenerate subspaces; energies agree to ~1e-12 eV), and
 *   - or_fock_nddo, the O(N_atoms^2) pair-blocked form of the same Fock sum.
 */
#include "mndo_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#define NELEM 20
#define ANG2BOHR 1.8897259886  /* mymath.h:84-87 */
#define BOHR2ANG 0.529177249   /* mymath.h:89-92 */
#define HARTREE2EV 27.211386   /* mymath.h:94-97 */

/* ------------------------------------------------------------------------ */
/* Parameter tables: effective (last-assigned) values of                     */
/* MNDOparameter::SetParameterValues (src/common/MNDO_parameters.cpp:48-121). */
/* ------------------------------------------------------------------------ */
typedef struct {
  double uss, upp, betaS, betaP, alpha, gss, gpp, gsp, gp2, hsp;
  double dd[2], pp[3];
} mndo_par;

static mndo_par PAR[NELEM];

/* STO-6G exponents/coefficients: the "Sparrow" blocks of STO-6G.cpp, which
 * overwrite the earlier assignments (H :42-54, C :92-118, N :155-181, O :219-245). */
typedef struct {
  double expS[6], expP[6], cS[6], cP[6];
} sto6g;

static sto6g STO[10];

/* 12-point Gauss-Hermite rule (overlap.cpp:26-37). */
static const double GH_X[12] = {
    -3.889724897869781919272, -3.020637025120889771711, -2.279507080501059900188,
    -1.59768263515260479671,  -0.9477883912401637437046, -0.314240376254359111277,
    3.889724897869781919272,  3.020637025120889771711,  2.279507080501059900188,
    1.59768263515260479671,   0.9477883912401637437046, 0.314240376254359111277};
static const double GH_W[12] = {
    2.65855168435630160602E-7, 8.5736870435878586546E-5, 0.00390539058462906185999,
    0.05160798561588392999187, 0.2604923102641611292334, 0.5701352362624795783471,
    2.65855168435630160602E-7, 8.5736870435878586546E-5, 0.00390539058462906185999,
    0.05160798561588392999187, 0.2604923102641611292334, 0.5701352362624795783471};

static int g_init = 0;

static void init_tables(void) {
  if (g_init) return;
  memset(PAR, 0, sizeof(PAR));
  memset(STO, 0, sizeof(STO));
  /* H */
  PAR[1].uss = -11.90627600; PAR[1].betaS = -6.98906400; PAR[1].alpha = 2.54413410;
  PAR[1].gss = 12.84800000; PAR[1].pp[0] = 1.05897351;
  /* C */
  PAR[6].uss = -52.27974500; PAR[6].upp = -39.20555800; PAR[6].betaS = -18.9850440;
  PAR[6].betaP = -7.93412200; PAR[6].alpha = 2.546380; PAR[6].gss = 12.23;
  PAR[6].gpp = 11.08; PAR[6].gsp = 11.47; PAR[6].gp2 = 9.84; PAR[6].hsp = 2.43;
  PAR[6].dd[0] = 0.80746618; PAR[6].dd[1] = 0.68515777;
  PAR[6].pp[0] = 1.11248501; PAR[6].pp[1] = 0.81309641; PAR[6].pp[2] = 0.74785493;
  /* N (dd[1] is the "original MNDO paper formula" value that overwrites 0.76788440) */
  PAR[7].uss = -71.93212200; PAR[7].upp = -57.17231900; PAR[7].betaS = -20.49575800;
  PAR[7].betaP = -20.49575800; PAR[7].alpha = 2.86134200; PAR[7].gss = 13.59;
  PAR[7].gpp = 12.98; PAR[7].gsp = 12.66; PAR[7].gp2 = 11.59; PAR[7].hsp = 3.14;
  PAR[7].dd[0] = 0.63990367; PAR[7].dd[1] = 0.54297627;
  PAR[7].pp[0] = 1.00115465; PAR[7].pp[1] = 0.63747362; PAR[7].pp[2] = 0.61528499;
  /* O (dd[1] overwritten the same way, 0.64152287 -> 0.45362517) */
  PAR[8].uss = -99.64430900; PAR[8].upp = -77.79747200; PAR[8].betaS = -32.68808200;
  PAR[8].betaP = -32.68808200; PAR[8].alpha = 3.16060400; PAR[8].gss = 15.42;
  PAR[8].gpp = 14.52; PAR[8].gsp = 14.48; PAR[8].gp2 = 12.98; PAR[8].hsp = 3.94;
  PAR[8].dd[0] = 0.53460239; PAR[8].dd[1] = 0.45362517;
  PAR[8].pp[0] = 0.88234058; PAR[8].pp[1] = 0.52124931; PAR[8].pp[2] = 0.52654944;

  static const double hS[6] = {4.098792193376860e+01, 7.515090619202102e+00,
                               2.102451537968084e+00, 7.222488471249888e-01,
                               2.804703624957924e-01, 1.155131838418812e-01};
  static const double hC[6] = {1.057956874127020e-01, 1.596794562489045e-01,
                               2.097264517638662e-01, 2.069126553471125e-01,
                               1.144014417359658e-01, 1.840525175585220e-02};
  static const double cSe[6] = {8.846144274163005e+01, 1.622292919221501e+01,
                                4.558993095148346e+00, 6.519461344814428e-01,
                                2.958932522507792e-01, 1.411098199523801e-01};
  static const double cSc[6] = {-8.534093658582062e-02, -1.190837759635380e-01,
                                -1.145234971659361e-01, 1.730334274890658e-01,
                                1.607236215755646e-01, 2.810826497961667e-02};
  static const double cPe[6] = {1.875086664726173e+01, 4.889844711496806e+00,
                                1.749633028685196e+00, 7.313800499385337e-01,
                                3.344367808801327e-01, 1.581099099775126e-01};
  static const double cPc[6] = {4.407314493338078e-01, 5.331737162464377e-01,
                                5.445171182695406e-01, 3.904445162839797e-01,
                                1.454563993779586e-01, 1.494840672526637e-02};
  static const double nSe[6] = {1.408553999528384e+02, 2.583144824406823e+01,
                                7.259194242116380e+00, 1.038080893483724e+00,
                                4.711449542017903e-01, 2.246863662931398e-01};
  static const double nSc[6] = {-1.209684224448265e-01, -1.687979660569579e-01,
                                -1.623338967119989e-01, 2.452701082384343e-01,
                                2.278212980714544e-01, 3.984268990094886e-02};
  static const double nPe[6] = {2.985663289232627e+01, 7.786002705798312e+00,
                                2.785905953918744e+00, 1.164561940873007e+00,
                                5.325170500259336e-01, 2.517552723103985e-01};
  static const double nPc[6] = {7.883124678886093e-01, 9.536589428843584e-01,
                                9.739482715823707e-01, 6.983669549124415e-01,
                                2.601699818271488e-01, 2.673740531657013e-02};
  static const double oSe[6] = {2.018091738330405e+02, 3.700975064350955e+01,
                                1.040053837613310e+01, 1.487300079059505e+00,
                                6.750282487920551e-01, 3.219171573709536e-01};
  static const double oSc[6] = {-1.584154624760238e-01, -2.210511414259479e-01,
                                -2.125860518260085e-01, 3.211960348235300e-01,
                                2.983457630220672e-01, 5.217641116074215e-02};
  static const double oPe[6] = {4.277679392805792e+01, 1.115531796470053e+01,
                                3.991479056714438e+00, 1.668514542173559e+00,
                                7.629585088943974e-01, 3.606998633356254e-01};
  static const double oPc[6] = {1.235684643746845e+00, 1.494866260646997e+00,
                                1.526669908217376e+00, 1.094694498739664e+00,
                                4.078180472887439e-01, 4.191104734371579e-02};
  memcpy(STO[1].expS, hS, sizeof hS); memcpy(STO[1].cS, hC, sizeof hC);
  memcpy(STO[6].expS, cSe, sizeof cSe); memcpy(STO[6].cS, cSc, sizeof cSc);
  memcpy(STO[6].expP, cPe, sizeof cPe); memcpy(STO[6].cP, cPc, sizeof cPc);
  memcpy(STO[7].expS, nSe, sizeof nSe); memcpy(STO[7].cS, nSc, sizeof nSc);
  memcpy(STO[7].expP, nPe, sizeof nPe); memcpy(STO[7].cP, nPc, sizeof nPc);
  memcpy(STO[8].expS, oSe, sizeof oSe); memcpy(STO[8].cS, oSc, sizeof oSc);
  memcpy(STO[8].expP, oPe, sizeof oPe); memcpy(STO[8].cP, oPc, sizeof oPc);
  g_init = 1;
}

static int supported(int z) { return z == 1 || z == 6 || z == 7 || z == 8; }

/* ------------------------------------------------------------------------ */
/* Atomic orbitals (ListAtomicOrbitals::SetOrbitals, atomicOrbitals.cpp:76-119) */
/* ------------------------------------------------------------------------ */
typedef struct {
  int atom, elem, l[3], lint;
  double c[3]; /* bohr */
} ao_t;

static int core_charge(int e) { /* AtomicOrbital::GetCoreCharge, atomicOrbitals.cpp:46-55 */
  if (e < 2) return e;
  if (e < 11) return e - 2;
  if (e < 19) return e - 10;
  return -1;
}
static int aos_of(int e) { return e < 2 ? 1 : (e < 11 ? 4 : -1); }

int or_n_ao(int n_atoms, const int* Z) {
  int n = 0;
  for (int i = 0; i < n_atoms; ++i) {
    if (!supported(Z[i])) return -1;
    n += aos_of(Z[i]);
  }
  return n;
}

static ao_t* build_aos(int n_atoms, const int* Z, const double* xyz, int* n_out) {
  int n = or_n_ao(n_atoms, Z);
  *n_out = n;
  if (n <= 0) return NULL;
  ao_t* ao = (ao_t*)calloc((size_t)n, sizeof(ao_t));
  int k = 0;
  for (int i = 0; i < n_atoms; ++i) {
    int cnt = aos_of(Z[i]);
    for (int t = 0; t < cnt; ++t, ++k) {
      ao[k].atom = i;
      ao[k].elem = Z[i];
      ao[k].l[0] = ao[k].l[1] = ao[k].l[2] = 0;
      if (t > 0) ao[k].l[t - 1] = 1;
      ao[k].lint = t;
      for (int c = 0; c < 3; ++c) ao[k].c[c] = xyz[3 * i + c] * ANG2BOHR;
    }
  }
  return ao;
}

static double dist3(const double* a, const double* b) { /* distancePointsV3, mymath.h:52-59 */
  double r = 0.0;
  for (int i = 0; i < 3; ++i) r += (a[i] - b[i]) * (a[i] - b[i]);
  return sqrt(r);
}

/* ------------------------------------------------------------------------ */
/* Multipole point-charge interactions, Hartree (multipole.cpp:20-498).       */
/* Argument order (R, element of the first orbital, element of the second).   */
/* ------------------------------------------------------------------------ */
static double addt(int ea, int l, int eb, int m) { /* AditiveTerms :494-498 */
  double a = PAR[ea].pp[l] + PAR[eb].pp[m];
  return a * a;
}
static double inv(double x) { return 1.0 / x; }

static double mp_qq(double R, int a, int b) {
  double t = addt(a, 0, b, 0);
  return 1.0 / sqrt(R * R + t);
}
static double mp_qUz(double R, int a, int b) {
  double t = addt(a, 0, b, 1), f = 0.0, x;
  x = R + PAR[b].dd[0]; f += inv(2.0 * sqrt(x * x + t));
  x = R - PAR[b].dd[0]; f -= inv(2.0 * sqrt(x * x + t));
  return f;
}
static double mp_Uzq(double R, int a, int b) { return mp_qUz(R, b, a) * -1.0; }
static double mp_qQpipi(double R, int a, int b) {
  double t = addt(a, 0, b, 2), f = 0.0, x;
  x = (2.0 * PAR[b].dd[1]) * (2.0 * PAR[b].dd[1]);
  f += inv(2.0 * sqrt(R * R + x + t));
  f -= inv(2.0 * sqrt(R * R + t));
  return f;
}
static double mp_Qpipiq(double R, int a, int b) { return mp_qQpipi(R, b, a); }
static double mp_qQzz(double R, int a, int b) {
  double t = addt(a, 0, b, 2), f = 0.0, x;
  x = R + 2.0 * PAR[b].dd[1]; f += inv(4.0 * sqrt(x * x + t));
  f -= inv(2.0 * sqrt(R * R + t));
  x = R - 2.0 * PAR[b].dd[1]; f += inv(4.0 * sqrt(x * x + t));
  return f;
}
static double mp_Qzzq(double R, int a, int b) { return mp_qQzz(R, b, a); }
static double mp_UpiUpi(double R, int a, int b) {
  double t = addt(a, 1, b, 1), f = 0.0, x;
  x = PAR[a].dd[0] - PAR[b].dd[0]; f += inv(2.0 * sqrt(R * R + x * x + t));
  x = PAR[a].dd[0] + PAR[b].dd[0]; f -= inv(2.0 * sqrt(R * R + x * x + t));
  return f;
}
static const double SG[2] = {1.0, -1.0};
static double mp_UzUz(double R, int a, int b) {
  static const double sgn[3] = {-1.0, 1.0, -1.0};
  double t = addt(a, 1, b, 1), f = 0.0, x;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      x = SG[i] * PAR[a].dd[0] + SG[j] * PAR[b].dd[0] + R;
      f += sgn[i + j] * inv(4.0 * sqrt(x * x + t));
    }
  return f;
}
static double mp_UpiQpiz(double R, int a, int b) {
  static const double sgn[3] = {-1.0, 1.0, -1.0};
  double t = addt(a, 1, b, 2), f = 0.0, x, y;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      x = R + SG[i] * PAR[b].dd[1];
      y = PAR[a].dd[0] + SG[j] * PAR[b].dd[1];
      f += sgn[i + j] * inv(4.0 * sqrt(x * x + y * y + t));
    }
  return f;
}
static double mp_QpizUpi(double R, int a, int b) { return mp_UpiQpiz(R, b, a) * -1.0; }
static double mp_UzQpipi(double R, int a, int b) {
  static const double sgn[2] = {-1.0, 1.0};
  double t = addt(a, 1, b, 2), f = 0.0, x, y;
  for (int i = 0; i < 2; ++i) {
    x = R + SG[i] * PAR[a].dd[0];
    y = 2.0 * PAR[b].dd[1];
    f += sgn[i] * inv(4.0 * sqrt(x * x + y * y + t));
  }
  for (int i = 0; i < 2; ++i) {
    x = R + SG[i] * PAR[a].dd[0];
    f += SG[i] * inv(4.0 * sqrt(x * x + t));
  }
  return f;
}
static double mp_QpipiUz(double R, int a, int b) { return mp_UzQpipi(R, b, a) * -1.0; }
static double mp_UzQz(double R, int a, int b) {
  static const double sgn[4] = {-1.0, -1.0, 1.0, 1.0};
  double t = addt(a, 1, b, 2), f = 0.0, x;
  int c = 0;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      x = R + SG[i] * PAR[a].dd[0] + SG[j] * 2.0 * PAR[b].dd[1];
      f += sgn[c++] * inv(8.0 * sqrt(x * x + t));
    }
  for (int i = 0; i < 2; ++i) {
    x = R + SG[i] * PAR[a].dd[0];
    f += SG[i] * inv(4.0 * sqrt(x * x + t));
  }
  return f;
}
static double mp_QzUz(double R, int a, int b) { return mp_UzQz(R, b, a) * -1.0; }
static double mp_QpipiQpipi(double R, int a, int b) {
  double t = addt(a, 2, b, 2), f = 0.0, x, y;
  for (int i = 0; i < 2; ++i) {
    y = PAR[a].dd[1] + SG[i] * PAR[b].dd[1];
    x = 4.0 * (y * y);
    f += inv(8.0 * sqrt(R * R + x + t));
  }
  y = 2.0 * PAR[a].dd[1]; x = y * y; f -= inv(4.0 * sqrt(R * R + x + t));
  y = 2.0 * PAR[b].dd[1]; x = y * y; f -= inv(4.0 * sqrt(R * R + x + t));
  f += inv(4.0 * sqrt(R * R + t));
  return f;
}
static double mp_QxxQyy(double R, int a, int b) {
  double t = addt(a, 2, b, 2), f = 0.0, x, ya = 2.0 * PAR[a].dd[1], yb = 2.0 * PAR[b].dd[1];
  x = ya * ya + yb * yb; f += inv(4.0 * sqrt(R * R + x + t));
  x = ya * ya; f -= inv(4.0 * sqrt(R * R + x + t));
  x = yb * yb; f -= inv(4.0 * sqrt(R * R + x + t));
  f += inv(4.0 * sqrt(R * R + t));
  return f;
}
static double mp_QpipiQzz(double R, int a, int b) {
  double t = addt(a, 2, b, 2), f = 0.0, ya = 2.0 * PAR[a].dd[1], aux, auxx, x;
  auxx = ya * ya + t;
  for (int i = 0; i < 2; ++i) {
    x = R + 2.0 * SG[i] * PAR[b].dd[1];
    aux = x * x;
    f += inv(8.0 * sqrt(aux + auxx));
    f -= inv(8.0 * sqrt(aux + t));
    f -= inv(8.0 * sqrt(R * R + auxx));
  }
  f += inv(4.0 * sqrt(R * R + t));
  return f;
}
static double mp_QzzQpipi(double R, int a, int b) { return mp_QpipiQzz(R, b, a); }
static double mp_QzzQzz(double R, int a, int b) {
  double t = addt(a, 2, b, 2), f = 0.0, x;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      x = R + SG[i] * 2.0 * PAR[a].dd[1] + SG[j] * 2.0 * PAR[b].dd[1];
      f += inv(16.0 * sqrt(x * x + t));
    }
  for (int i = 0; i < 2; ++i) {
    x = R + 2.0 * SG[i] * PAR[a].dd[1]; f -= inv(8.0 * sqrt(x * x + t));
    x = R + 2.0 * SG[i] * PAR[b].dd[1]; f -= inv(8.0 * sqrt(x * x + t));
  }
  f += inv(4.0 * sqrt(R * R + t));
  return f;
}
static double mp_QpizQpiz(double R, int a, int b) {
  static const double sgn[8] = {1.0, -1.0, -1.0, 1.0, -1.0, 1.0, 1.0, -1.0};
  double t = addt(a, 2, b, 2), f = 0.0, x, y;
  int c = 0;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k) {
        x = R + SG[i] * PAR[a].dd[1] + SG[j] * PAR[b].dd[1];
        y = PAR[a].dd[1] + SG[k] * PAR[b].dd[1];
        f += sgn[c++] * inv(8.0 * sqrt(x * x + y * y + t));
      }
  return f;
}

/* ------------------------------------------------------------------------ */
/* Local frame and rotation (twocenterintegral.cpp:468-513).                 */
/* ------------------------------------------------------------------------ */
static void rotation(const double* va, const double* vb, double m[3][3]) {
  double R = dist3(va, vb);
  for (int i = 0; i < 3; ++i) m[2][i] = (vb[i] - va[i]) / R;
  if (fabs(m[2][0] - 0.0) < 1e-10 && fabs(m[2][1] - 0.0) < 1e-10) {
    m[1][0] = 0.0; m[1][1] = 1.0; m[1][2] = 0.0;
    m[0][0] = 1.0; m[0][1] = 0.0; m[0][2] = 0.0;
  } else {
    double s = 1.0 / sqrt(m[2][0] * m[2][0] + m[2][1] * m[2][1]);
    m[1][0] = s * -m[2][1];
    m[1][1] = s * m[2][0];
    m[1][2] = 0.0;
    m[0][0] = m[2][1] * m[1][2] - m[2][2] * m[1][1];
    m[0][1] = m[2][2] * m[1][0] - m[2][0] * m[1][2];
    m[0][2] = m[2][0] * m[1][1] - m[2][1] * m[1][0];
  }
}
#define R1(a, g, v) ((v) * m[a][(g) - 1])
#define R2(a, b, g, h, v) ((v) * m[a][(g) - 1] * m[b][(h) - 1])
#define R3(a, b, c, g, h, k, v) ((v) * m[a][(g) - 1] * m[b][(h) - 1] * m[c][(k) - 1])
#define R4(a, b, c, d, g, h, k, l, v) \
  ((v) * m[a][(g) - 1] * m[b][(h) - 1] * m[c][(k) - 1] * m[d][(l) - 1])

/* IntegralType* (twocenterintegral.cpp:515-736); a, c = elements of the
 * first/second pair, t[] = angular types after CorrectOrderIntegral. */
static double it_ss_sp(double R, const ao_t* A, const ao_t* C, const int* t, int invf) {
  double m[3][3];
  rotation(A->c, C->c, m);
  double v = invf ? mp_Uzq(R, A->elem, C->elem) : mp_qUz(R, A->elem, C->elem);
  return R1(2, t[3], v);
}
static double it_ss_pp(double R, const ao_t* A, const ao_t* C, const int* t, int invf) {
  double m[3][3];
  rotation(A->c, C->c, m);
  int a = A->elem, c = C->elem;
  double qq = mp_qq(R, a, c), vp, vz;
  if (!invf) {
    vp = qq + mp_qQpipi(R, a, c);
    vz = qq + mp_qQzz(R, a, c);
  } else {
    vp = qq + mp_Qpipiq(R, a, c);
    vz = qq + mp_Qzzq(R, a, c);
  }
  vp = R2(0, 0, t[2], t[3], vp) + R2(1, 1, t[2], t[3], vp);
  vz = R2(2, 2, t[2], t[3], vz);
  return vp + vz;
}
static double it_sp_sp(double R, const ao_t* A, const ao_t* C, const int* t) {
  double m[3][3];
  rotation(A->c, C->c, m);
  double v1 = mp_UpiUpi(R, A->elem, C->elem);
  v1 = R2(0, 0, t[1], t[3], v1) + R2(1, 1, t[1], t[3], v1);
  double v2 = mp_UzUz(R, A->elem, C->elem);
  v2 = R2(2, 2, t[1], t[3], v2);
  return v1 + v2;
}
static double it_sp_pp(double R, const ao_t* A, const ao_t* C, const int* t, int invf) {
  double m[3][3];
  rotation(A->c, C->c, m);
  int a = A->elem, c = C->elem;
  double u, v1, v2, v3;
  if (!invf) {
    u = mp_Uzq(R, a, c);
    v1 = u + mp_UzQpipi(R, a, c);
    v2 = u + mp_UzQz(R, a, c);
    v3 = mp_UpiQpiz(R, a, c);
  } else {
    u = mp_qUz(R, a, c);
    v1 = u + mp_QpipiUz(R, a, c);
    v2 = u + mp_QzUz(R, a, c);
    v3 = mp_QpizUpi(R, a, c);
  }
  v1 = R3(2, 0, 0, t[1], t[2], t[3], v1) + R3(2, 1, 1, t[1], t[2], t[3], v1);
  v2 = R3(2, 2, 2, t[1], t[2], t[3], v2);
  v3 = R3(0, 2, 0, t[1], t[2], t[3], v3) + R3(1, 2, 1, t[1], t[2], t[3], v3) +
       R3(0, 2, 0, t[1], t[3], t[2], v3) + R3(1, 2, 1, t[1], t[3], t[2], v3);
  return v1 + v2 + v3;
}
static double it_pp_pp(double R, const ao_t* A, const ao_t* C, const int* t) {
  double m[3][3];
  rotation(A->c, C->c, m);
  int a = A->elem, c = C->elem;
  double qq = mp_qq(R, a, c), qQzz = mp_qQzz(R, a, c), Qzzq = mp_Qzzq(R, a, c);
  double zzzz = qq + qQzz + Qzzq + mp_QzzQzz(R, a, c);
  zzzz = R4(2, 2, 2, 2, t[0], t[1], t[2], t[3], zzzz);
  double qQpp = mp_qQpipi(R, a, c), Qppq = mp_Qpipiq(R, a, c);
  double pppp = qq + qQpp + Qppq + mp_QpipiQpipi(R, a, c);
  double xyxy = pppp;
  pppp = R4(0, 0, 0, 0, t[0], t[1], t[2], t[3], pppp) + R4(1, 1, 1, 1, t[0], t[1], t[2], t[3], pppp);
  double zzpp = qq + qQpp + Qzzq + mp_QzzQpipi(R, a, c);
  zzpp = R4(2, 2, 0, 0, t[0], t[1], t[2], t[3], zzpp) + R4(2, 2, 1, 1, t[0], t[1], t[2], t[3], zzpp);
  double ppzz = qq + Qppq + qQzz + mp_QpipiQzz(R, a, c);
  ppzz = R4(0, 0, 2, 2, t[0], t[1], t[2], t[3], ppzz) + R4(1, 1, 2, 2, t[0], t[1], t[2], t[3], ppzz);
  double pzpz = mp_QpizQpiz(R, a, c);
  pzpz = R4(0, 2, 0, 2, t[0], t[1], t[2], t[3], pzpz) + R4(0, 2, 0, 2, t[0], t[1], t[3], t[2], pzpz) +
         R4(0, 2, 0, 2, t[1], t[0], t[2], t[3], pzpz) + R4(0, 2, 0, 2, t[1], t[0], t[3], t[2], pzpz) +
         R4(1, 2, 1, 2, t[0], t[1], t[2], t[3], pzpz) + R4(1, 2, 1, 2, t[0], t[1], t[3], t[2], pzpz) +
         R4(1, 2, 1, 2, t[1], t[0], t[2], t[3], pzpz) + R4(1, 2, 1, 2, t[1], t[0], t[3], t[2], pzpz);
  double xxyy = qq + qQpp + Qppq + mp_QxxQyy(R, a, c);
  xyxy = (xyxy - xxyy) * 0.5;
  xxyy = R4(0, 0, 1, 1, t[0], t[1], t[2], t[3], xxyy) + R4(1, 1, 0, 0, t[0], t[1], t[2], t[3], xxyy);
  xyxy = R4(0, 1, 0, 1, t[0], t[1], t[2], t[3], xyxy) + R4(0, 1, 0, 1, t[0], t[1], t[3], t[2], xyxy) +
         R4(0, 1, 0, 1, t[1], t[0], t[2], t[3], xyxy) + R4(0, 1, 0, 1, t[1], t[0], t[3], t[2], xyxy);
  return zzzz + pppp + zzpp + ppzz + pzpz + xxyy + xyxy;
}

/* One-center integrals from the G parameters (:738-778), eV. */
static double self_pp_pp(int e, const int* t) {
  if (t[0] == t[1] && t[2] == t[3]) return t[0] == t[2] ? PAR[e].gpp : PAR[e].gp2;
  if (t[0] == t[2] && t[1] == t[3]) return 0.5 * (PAR[e].gpp - PAR[e].gp2);
  return 0.0;
}

static int pair_type(const ao_t* a, const ao_t* b) {
  return a->l[0] + a->l[1] + a->l[2] + b->l[0] + b->l[1] + b->l[2];
}

/* TwoCenterIntegral::ComputeTwoCenterIntegral (:144-227). */
static double two_center(const ao_t* A, const ao_t* B, const ao_t* C, const ao_t* D) {
  double val = 0.0;
  if (A->atom == B->atom && C->atom == D->atom) {
    int pa = pair_type(A, B), pb = pair_type(C, D);
    int t[4] = {A->lint, B->lint, C->lint, D->lint};
    if (A->atom == C->atom) {
      int e = A->elem;
      if (pa == pb) {
        if (pa == 0) return PAR[e].gss;
        if (pa == 1) return t[1] == t[3] ? PAR[e].hsp : 0.0;
        return self_pp_pp(e, t);
      } else {
        if (pa == 0 && pb == 2) return t[2] == t[3] ? PAR[e].gsp : 0.0;
        if (pa == 2 && pb == 0) return t[0] == t[1] ? PAR[e].gsp : 0.0;
      }
    } else {
      double R = dist3(A->c, C->c);
      int invf = 0;
      if (pa > pb) { /* CorrectOrderIntegral :785-796 */
        int s = pa; pa = pb; pb = s;
        s = t[0]; t[0] = t[2]; t[2] = s;
        s = t[1]; t[1] = t[3]; t[3] = s;
        invf = 1;
      }
      if (pa == 0) {
        if (pb == 0) val = mp_qq(R, A->elem, C->elem);
        else if (pb == 1) val = it_ss_sp(R, A, C, t, invf);
        else val = it_ss_pp(R, A, C, t, invf);
      } else if (pa == pb) {
        val = (pa == 1) ? it_sp_sp(R, A, C, t) : it_pp_pp(R, A, C, t);
      } else {
        val = it_sp_pp(R, A, C, t, invf);
      }
      return val * HARTREE2EV;
    }
  }
  return val;
}

/* ------------------------------------------------------------------------ */
/* Jagged pair-block storage (twocenterintegral.cpp:55-107).                 */
/* ------------------------------------------------------------------------ */
static int block_size(int za, int zb) {
  if (za == 1 && zb == 1) return 1;
  if (za > 1 && zb > 1) return 100;
  return 10;
}

long or_pair_block_count(int n_atoms, const int* Z) {
  long n = 0;
  for (int i = 0; i < n_atoms; ++i)
    for (int j = 0; j <= i; ++j) n += block_size(Z[i], Z[j]);
  return n;
}

static long* block_offsets(int n_atoms, const int* Z) {
  long* off = (long*)malloc(sizeof(long) * (size_t)n_atoms * (size_t)(n_atoms + 1) / 2 + 8);
  long o = 0;
  for (int i = 0; i < n_atoms; ++i)
    for (int j = 0; j <= i; ++j) {
      off[(long)i * (i + 1) / 2 + j] = o;
      o += block_size(Z[i], Z[j]);
    }
  return off;
}

/* MOPAC pair order SS,SX,XX,SY,XY,YY,SZ,XZ,YZ,ZZ (:399-402). */
static const int MOPA[10] = {0, 0, 1, 0, 1, 2, 0, 1, 2, 3};
static const int MOPB[10] = {0, 1, 1, 2, 2, 2, 3, 3, 3, 3};
static const int MOPIDX[4][4] = {{0, 1, 3, 6}, {1, 2, 4, 7}, {3, 4, 5, 8}, {6, 7, 8, 9}};

int or_pair_blocks(int n_atoms, const int* Z, const double* xyz, double* out) {
  init_tables();
  int n;
  ao_t* ao = build_aos(n_atoms, Z, xyz, &n);
  if (!ao) return -1;
  long* off = block_offsets(n_atoms, Z);
  /* ComputeAllTwoCenterIntegral (:229-279) with ComputePair_* (:388-466). */
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j <= i; ++j) {
      int ia = ao[i].atom, ib = ao[j].atom;
      double* blk = out + off[(long)ia * (ia + 1) / 2 + ib];
      int ei = ao[i].elem, ej = ao[j].elem;
      if (ei > 1 && ej > 1) { /* X-X and X-Y */
        for (int k = 0; k < 10; ++k)
          for (int l = 0; l < 10; ++l)
            blk[k * 10 + l] = two_center(&ao[i + MOPA[k]], &ao[i + MOPB[k]], &ao[j + MOPA[l]],
                                         &ao[j + MOPB[l]]);
        j += 3;
      } else if (ei == 1 && ej == 1) {
        blk[0] = two_center(&ao[i], &ao[i], &ao[j], &ao[j]);
      } else if (ei == 1) { /* H-X: (H H | X pair) with the H first */
        for (int k = 0; k < 10; ++k)
          blk[k] = two_center(&ao[i], &ao[i], &ao[j + MOPA[k]], &ao[j + MOPB[k]]);
        j += 3;
      } else { /* X-H: also (H H | X pair), H (lower atom) first */
        for (int k = 0; k < 10; ++k)
          blk[k] = two_center(&ao[j], &ao[j], &ao[i + MOPA[k]], &ao[i + MOPB[k]]);
      }
    }
    if (ao[i].elem > 1) i += 3;
  }
  free(off);
  free(ao);
  return 0;
}

/* TwoCenterIntegral::GetValueFromArray (:285-386). */
static double get_value(const ao_t* A, const ao_t* B, const ao_t* C, const ao_t* D,
                        const double* blocks, const long* off) {
  if (!(A->atom == B->atom && C->atom == D->atom)) return 0.0;
  int ia = A->atom, ic = C->atom;
  int la = A->lint, lb = B->lint, lc = C->lint, ld = D->lint, s;
  if (ia < ic) {
    s = ia; ia = ic; ic = s;
    s = la; la = lc; lc = s;
    s = lb; lb = ld; ld = s;
  }
  const double* blk = blocks + off[(long)ia * (ia + 1) / 2 + ic];
  if (A->elem > 1) {
    if (C->elem > 1) return blk[MOPIDX[la][lb] * 10 + MOPIDX[lc][ld]];
    return blk[MOPIDX[A->lint][B->lint]];
  }
  if (C->elem > 1) return blk[MOPIDX[C->lint][D->lint]];
  return blk[0];
}

double or_get_value(int n_atoms, const int* Z, const double* blocks, int a, int b, int c, int d) {
  init_tables();
  int n;
  double* xyz = (double*)calloc((size_t)3 * n_atoms, sizeof(double));
  ao_t* ao = build_aos(n_atoms, Z, xyz, &n);
  long* off = block_offsets(n_atoms, Z);
  double v = get_value(&ao[a], &ao[b], &ao[c], &ao[d], blocks, off);
  free(off);
  free(ao);
  free(xyz);
  return v;
}

/* ------------------------------------------------------------------------ */
/* STO-6G overlap (Overlap::ComputeOverlap, overlap.cpp:68-134, 177-192).    */
/* ------------------------------------------------------------------------ */
static double gh_1d(int la, int lb, double p, double pa, double pb) {
  double sp = sqrt(p);
  if (la == 0 && lb == 0) return sqrt(M_PI) / sp;
  double s = 0.0;
  for (int i = 0; i < 12; ++i) {
    double x = GH_X[i] / sp;
    double term = GH_W[i];
    if (la) term *= (x + pa);
    if (lb) term *= (x + pb);
    s += term;
  }
  return s / sp;
}

static double overlap(const ao_t* A, const ao_t* B) {
  double ab = 0.0;
  for (int i = 0; i < 3; ++i) ab += (A->c[i] - B->c[i]) * (A->c[i] - B->c[i]);
  int sa = A->l[0] + A->l[1] + A->l[2], sb = B->l[0] + B->l[1] + B->l[2];
  const sto6g* ga = &STO[A->elem];
  const sto6g* gb = &STO[B->elem];
  double tot = 0.0;
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      double al = sa == 0 ? ga->expS[i] : ga->expP[i];
      double be = sb == 0 ? gb->expS[j] : gb->expP[j];
      double p = al + be;
      double mu = al * be / p;
      double o = exp(-mu * ab);
      o *= sa == 0 ? ga->cS[i] : ga->cP[i];
      o *= sb == 0 ? gb->cS[j] : gb->cP[j];
      for (int k = 0; k < 3; ++k) {
        double P = al * A->c[k];
        P += be * B->c[k];
        P /= p;
        o *= gh_1d(A->l[k], B->l[k], p, P - A->c[k], P - B->c[k]);
      }
      tot += o;
    }
  return tot;
}

double or_overlap(int n_atoms, const int* Z, const double* xyz, int a, int b) {
  init_tables();
  int n;
  ao_t* ao = build_aos(n_atoms, Z, xyz, &n);
  double v = overlap(&ao[a], &ao[b]);
  free(ao);
  return v;
}

/* ------------------------------------------------------------------------ */
/* Hcore (hcore.cpp:25-99).                                                  */
/* ------------------------------------------------------------------------ */
static double hc_same_atom(const ao_t* ao, int n, int a, int b, const double* blocks,
                           const long* off) {
  double s = 0.0;
  for (int i = 0; i < n; ++i)
    if (ao[a].atom != ao[i].atom && ao[i].lint == 0)
      s += -(double)core_charge(ao[i].elem) * get_value(&ao[a], &ao[b], &ao[i], &ao[i], blocks, off);
  return s;
}

int or_hcore(int n_atoms, const int* Z, const double* xyz, const double* blocks, double* h) {
  init_tables();
  int n;
  ao_t* ao = build_aos(n_atoms, Z, xyz, &n);
  if (!ao) return -1;
  long* off = block_offsets(n_atoms, Z);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) {
      double v;
      const ao_t* A = &ao[i];
      const ao_t* B = &ao[j];
      if (i == j) {
        v = (A->lint == 0 ? PAR[A->elem].uss : PAR[A->elem].upp) +
            hc_same_atom(ao, n, i, i, blocks, off);
      } else if (A->atom == B->atom) {
        v = hc_same_atom(ao, n, i, j, blocks, off);
      } else {
        double s = overlap(A, B);
        double bsum = (A->lint == 0 ? PAR[A->elem].betaS : PAR[A->elem].betaP) +
                      (B->lint == 0 ? PAR[B->elem].betaS : PAR[B->elem].betaP);
        v = bsum * s * 0.5;
      }
      h[(long)i * (i + 1) / 2 + j] = v;
    }
  free(off);
  free(ao);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Density matrix (densitymatrix.cpp:22-47, 99-107).                         */
/* ------------------------------------------------------------------------ */
int or_n_electrons(int n_atoms, const int* Z) {
  int ne = 0;
  for (int i = 0; i < n_atoms; ++i) {
    if (!supported(Z[i])) return -1;
    ne += core_charge(Z[i]);
  }
  return ne;
}

int or_guess_density(int n_atoms, const int* Z, double* p) {
  int n;
  double* xyz = (double*)calloc((size_t)3 * n_atoms, sizeof(double));
  ao_t* ao = build_aos(n_atoms, Z, xyz, &n);
  free(xyz);
  if (!ao) return -1;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j)
      p[(long)i * (i + 1) / 2 + j] =
          i == j ? (double)core_charge(ao[i].elem) / aos_of(ao[i].elem) : 0.0;
  free(ao);
  return 0;
}

int or_density(int n, int n_occ, const double* C, double* p) {
  for (long i = 0; i < n; ++i)
    for (long j = 0; j <= i; ++j) {
      double v = 0.0;
      for (long k = 0; k < n_occ; ++k) v += C[n * k + i] * C[n * k + j];
      p[i * (i + 1) / 2 + j] = v * 2.0;
    }
  return 0;
}

static inline long pk(long r, long c) { return r >= c ? r * (r + 1) / 2 + c : c * (c + 1) / 2 + r; }

/* ------------------------------------------------------------------------ */
/* Fock matrix.                                                              */
/* ------------------------------------------------------------------------ */
int or_fock_ref(int n_atoms, const int* Z, const double* blocks, const double* h,
                const double* p, double* f) {
  init_tables();
  int n;
  double* xyz = (double*)calloc((size_t)3 * n_atoms, sizeof(double));
  ao_t* ao = build_aos(n_atoms, Z, xyz, &n);
  free(xyz);
  if (!ao) return -1;
  long* off = block_offsets(n_atoms, Z);
  for (int u = 0; u < n; ++u)
    for (int v = 0; v <= u; ++v) {
      double g = 0.0; /* ComputeGMatrix_UV (twoelectronmatrix.cpp:14-35) */
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
          double t = get_value(&ao[u], &ao[v], &ao[i], &ao[j], blocks, off);
          t -= 0.5 * get_value(&ao[u], &ao[i], &ao[v], &ao[j], blocks, off);
          t *= p[pk(i, j)];
          g += t;
        }
      double val = h[pk(u, v)];
      val += g;
      f[pk(u, v)] = val;
    }
  free(off);
  free(ao);
  return 0;
}

int or_fock_nddo(int n_atoms, const int* Z, const double* blocks, const double* h,
                 const double* p, double* f) {
  init_tables();
  int n;
  double* xyz = (double*)calloc((size_t)3 * n_atoms, sizeof(double));
  ao_t* ao = build_aos(n_atoms, Z, xyz, &n);
  free(xyz);
  if (!ao) return -1;
  long* off = block_offsets(n_atoms, Z);
  int* first = (int*)malloc(sizeof(int) * (size_t)(n_atoms + 1));
  for (int i = 0, k = 0; i <= n_atoms; ++i) {
    first[i] = k;
    if (i < n_atoms) k += aos_of(Z[i]);
  }
  for (long t = 0; t < (long)n * (n + 1) / 2; ++t) f[t] = h[t];
  for (int A = 0; A < n_atoms; ++A) {
    /* diagonal block: Coulomb over every atom B, one-center exchange on A */
    for (int u = first[A]; u < first[A + 1]; ++u)
      for (int v = first[A]; v <= u; ++v) {
        double g = 0.0;
        for (int B = 0; B < n_atoms; ++B)
          for (int i = first[B]; i < first[B + 1]; ++i)
            for (int j = first[B]; j < first[B + 1]; ++j)
              g += p[pk(i, j)] * get_value(&ao[u], &ao[v], &ao[i], &ao[j], blocks, off);
        for (int i = first[A]; i < first[A + 1]; ++i)
          for (int j = first[A]; j < first[A + 1]; ++j)
            g -= 0.5 * p[pk(i, j)] * get_value(&ao[u], &ao[i], &ao[v], &ao[j], blocks, off);
        f[pk(u, v)] += g;
      }
    /* off-diagonal blocks A > B: two-center exchange only */
    for (int B = 0; B < A; ++B)
      for (int u = first[A]; u < first[A + 1]; ++u)
        for (int v = first[B]; v < first[B + 1]; ++v) {
          double g = 0.0;
          for (int i = first[A]; i < first[A + 1]; ++i)
            for (int j = first[B]; j < first[B + 1]; ++j)
              g += p[pk(i, j)] * get_value(&ao[u], &ao[i], &ao[v], &ao[j], blocks, off);
          f[pk(u, v)] += -0.5 * g;
        }
  }
  free(first);
  free(off);
  free(ao);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Energies.                                                                 */
/* ------------------------------------------------------------------------ */
double or_energy(int n, const double* h, const double* f, const double* p) {
  double e = 0.0, t;
  for (long i = 0; i < n; ++i)
    for (long j = 0; j <= i; ++j) {
      long k = i * (i + 1) / 2 + j;
      t = h[k] + f[k];
      t *= p[k];
      t *= 2.0;
      e += t;
    }
  for (long i = 0; i < n; ++i) {
    long k = i * (i + 1) / 2 + i;
    t = h[k] + f[k];
    t *= p[k];
    e -= t;
  }
  return 0.5 * e;
}

static double adjustable(const ao_t* A, const ao_t* B) { /* corecorerepulsion.cpp:43-65 */
  double r = dist3(A->c, B->c);
  r = r * BOHR2ANG;
  if (B->elem == 1 && (A->elem == 7 || A->elem == 8))
    return 1.0 + r * exp(-PAR[A->elem].alpha * r) + exp(-PAR[B->elem].alpha * r);
  if (A->elem == 1 && (B->elem == 7 || B->elem == 8))
    return 1.0 + r * exp(-PAR[B->elem].alpha * r) + exp(-PAR[A->elem].alpha * r);
  return 1.0 + exp(-PAR[A->elem].alpha * r) + exp(-PAR[B->elem].alpha * r);
}

double or_core_core(int n_atoms, const int* Z, const double* xyz, const double* blocks) {
  init_tables();
  int n;
  ao_t* ao = build_aos(n_atoms, Z, xyz, &n);
  if (!ao) return NAN;
  long* off = block_offsets(n_atoms, Z);
  double e = 0.0;
  for (int i = 0; i < n; ++i) {
    for (int j = i; j < n; ++j) {
      if (ao[i].atom != ao[j].atom) {
        double t = (double)(core_charge(ao[i].elem) * core_charge(ao[j].elem));
        t *= get_value(&ao[i], &ao[i], &ao[j], &ao[j], blocks, off);
        t *= adjustable(&ao[i], &ao[j]);
        e += t;
        if (ao[j].elem > 1) j += 3;
      }
    }
    if (ao[i].elem > 1) i += 3;
  }
  free(off);
  free(ao);
  return e;
}

/* ------------------------------------------------------------------------ */
/* Dense symmetric eigensolver: Householder tridiagonalisation + implicit    */
/* symmetric QR with Wilkinson shifts (Golub & Van Loan 8.3), written here.  */
/* ------------------------------------------------------------------------ */
int or_eigh(int n, double* a, double* w) {
#define A_(i, j) a[(long)(j) * n + (i)]
  if (n <= 0) return 0;
  for (int j = 0; j < n; ++j)
    for (int i = j + 1; i < n; ++i) A_(j, i) = A_(i, j);
  double* d = (double*)calloc((size_t)n, sizeof(double));
  double* e = (double*)calloc((size_t)n, sizeof(double));
  double* tau = (double*)calloc((size_t)n, sizeof(double));
  double* pv = (double*)calloc((size_t)n, sizeof(double));
  /* reduction: column k, reflector on rows k+1..n-1 */
  for (int k = 0; k + 2 < n; ++k) {
    double x0 = A_(k + 1, k), s2 = 0.0;
    for (int i = k + 2; i < n; ++i) s2 += A_(i, k) * A_(i, k);
    if (s2 == 0.0) {
      tau[k] = 0.0;
      e[k] = x0;
      continue;
    }
    double nrm = sqrt(x0 * x0 + s2);
    double beta = x0 >= 0.0 ? -nrm : nrm;
    tau[k] = (beta - x0) / beta;
    double sc = 1.0 / (x0 - beta);
    for (int i = k + 2; i < n; ++i) A_(i, k) *= sc; /* v = [1, A(k+2:,k)] */
    A_(k + 1, k) = 1.0;
    e[k] = beta;
    /* p = tau * A22 v */
    for (int i = k + 1; i < n; ++i) {
      double s = 0.0;
      for (int j = k + 1; j < n; ++j) s += A_(i, j) * A_(j, k);
      pv[i] = tau[k] * s;
    }
    double kk = 0.0;
    for (int i = k + 1; i < n; ++i) kk += pv[i] * A_(i, k);
    kk *= 0.5 * tau[k];
    for (int i = k + 1; i < n; ++i) pv[i] -= kk * A_(i, k);
    for (int j = k + 1; j < n; ++j)
      for (int i = k + 1; i < n; ++i) A_(i, j) -= A_(i, k) * pv[j] + pv[i] * A_(j, k);
  }
  for (int i = 0; i < n; ++i) d[i] = A_(i, i);
  if (n >= 2) e[n - 2] = A_(n - 1, n - 2);
  /* accumulate Q = H_0 ... H_{n-3} into z (backward) */
  double* z = (double*)calloc((size_t)n * n, sizeof(double));
#define Z_(i, j) z[(long)(j) * n + (i)]
  for (int i = 0; i < n; ++i) Z_(i, i) = 1.0;
  for (int k = n - 3; k >= 0; --k) {
    if (tau[k] == 0.0) continue;
    /* v_i = A_(i,k) for i >= k+1 (A_(k+1,k) = 1) */
    for (int j = k + 1; j < n; ++j) {
      double s = 0.0;
      for (int i = k + 1; i < n; ++i) s += A_(i, k) * Z_(i, j);
      s *= tau[k];
      for (int i = k + 1; i < n; ++i) Z_(i, j) -= s * A_(i, k);
    }
  }
  /* implicit symmetric QR on (d, e) with e[i] = T(i+1, i) */
  int status = 0;
  const double eps = 2.220446049250313e-16;
  int hi = n - 1, iters = 0;
  while (hi > 0) {
    for (int i = 0; i < hi; ++i)
      if (fabs(e[i]) <= eps * (fabs(d[i]) + fabs(d[i + 1]))) e[i] = 0.0;
    while (hi > 0 && e[hi - 1] == 0.0) --hi;
    if (hi == 0) break;
    int lo = hi - 1;
    while (lo > 0 && e[lo - 1] != 0.0) --lo;
    if (++iters > 60 * n) {
      status = hi + 1;
      break;
    }
    double dl = (d[hi - 1] - d[hi]) * 0.5, b = e[hi - 1];
    double den = dl + (dl >= 0.0 ? hypot(dl, b) : -hypot(dl, b));
    double mu = d[hi] - b * b / den;
    double x = d[lo] - mu, y = e[lo];
    for (int k = lo; k < hi; ++k) {
      double r = hypot(x, y);
      double c = 1.0, s = 0.0;
      if (r != 0.0) {
        c = x / r;
        s = y / r;
      }
      if (k > lo) e[k - 1] = r;
      double dk = d[k], dk1 = d[k + 1], ek = e[k];
      d[k] = c * c * dk + 2.0 * c * s * ek + s * s * dk1;
      d[k + 1] = s * s * dk - 2.0 * c * s * ek + c * c * dk1;
      e[k] = c * s * (dk1 - dk) + (c * c - s * s) * ek;
      if (k + 1 < hi) {
        y = s * e[k + 1];
        e[k + 1] = c * e[k + 1];
        x = e[k];
      }
      for (int i = 0; i < n; ++i) {
        double zk = Z_(i, k), zk1 = Z_(i, k + 1);
        Z_(i, k) = c * zk + s * zk1;
        Z_(i, k + 1) = -s * zk + c * zk1;
      }
    }
  }
  /* ascending order (selection sort on values, swapping vectors) */
  for (int i = 0; i < n; ++i) {
    int m = i;
    for (int j = i + 1; j < n; ++j)
      if (d[j] < d[m]) m = j;
    if (m != i) {
      double t = d[i]; d[i] = d[m]; d[m] = t;
      for (int r = 0; r < n; ++r) {
        t = Z_(r, i); Z_(r, i) = Z_(r, m); Z_(r, m) = t;
      }
    }
  }
  memcpy(a, z, sizeof(double) * (size_t)n * n);
  memcpy(w, d, sizeof(double) * (size_t)n);
  free(z); free(d); free(e); free(tau); free(pv);
  return status;
#undef A_
#undef Z_
}

/* ------------------------------------------------------------------------ */
/* SCF driver (SCFCalculation::ComputeSCF / InitialGuess /                   */
/* GetEigenValVecOfFmatrix, scfcalculation.cpp:66-203).                      */
/* ------------------------------------------------------------------------ */
int or_scf(int n_atoms, const int* Z, const double* xyz, double threshold, int max_steps,
           int fock_mode, or_scf_result* res, double* energies, double* p_out, double* f_out,
           double* eig_out) {
  init_tables();
  memset(res, 0, sizeof(*res));
  int n = or_n_ao(n_atoms, Z);
  if (n <= 0) {
    res->status = -1;
    return -1;
  }
  long T = (long)n * (n + 1) / 2;
  double* blocks = (double*)malloc(sizeof(double) * (size_t)or_pair_block_count(n_atoms, Z));
  double* h = (double*)malloc(sizeof(double) * (size_t)T);
  double* p = (double*)malloc(sizeof(double) * (size_t)T);
  double* f = (double*)malloc(sizeof(double) * (size_t)T);
  double* c = (double*)malloc(sizeof(double) * (size_t)n * n);
  double* w = (double*)malloc(sizeof(double) * (size_t)n);
  or_pair_blocks(n_atoms, Z, xyz, blocks);
  or_hcore(n_atoms, Z, xyz, blocks, h);
  res->e_cc = or_core_core(n_atoms, Z, xyz, blocks);
  int n_occ = or_n_electrons(n_atoms, Z) / 2;
  int (*fock)(int, const int*, const double*, const double*, const double*, double*) =
      fock_mode == 0 ? or_fock_ref : or_fock_nddo;

  or_guess_density(n_atoms, Z, p);
  fock(n_atoms, Z, blocks, h, p, f);
  double e = or_energy(n, h, f, p), e_old;
  if (energies) energies[0] = e;
  int steps = 0, conv = 0;
  for (int i = 0; i < n; ++i) w[i] = 0.0;
  while (!conv && max_steps > steps) {
    for (long i = 0; i < n; ++i)
      for (long j = 0; j <= i; ++j) c[j * n + i] = f[i * (i + 1) / 2 + j];
    int st = or_eigh(n, c, w);
    if (st != 0) {
      res->status = st;
      break;
    }
    or_density(n, n_occ, c, p);
    fock(n_atoms, Z, blocks, h, p, f);
    e_old = e;
    e = or_energy(n, h, f, p);
    conv = fabs(e_old - e) < threshold;
    steps++;
    if (energies) energies[steps] = e;
  }
  res->e_el = e;
  res->e_tot = e + res->e_cc;
  res->steps = steps;
  res->converged = conv;
  if (p_out) memcpy(p_out, p, sizeof(double) * (size_t)T);
  if (f_out) memcpy(f_out, f, sizeof(double) * (size_t)T);
  if (eig_out) memcpy(eig_out, w, sizeof(double) * (size_t)n);
  free(blocks); free(h); free(p); free(f); free(c); free(w);
  return res->status;
}
