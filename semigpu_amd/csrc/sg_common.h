// sg_common.h -- device-side constants and helpers shared by the HIP kernels.
//
// MNDO parameters are the effective (last-assigned) values of
// MNDOparameter::SetParameterValues (src/common/MNDO_parameters.cpp:48-121),
// STO-6G exponents/coefficients the effective blocks of STO_6G
// (src/common/STO-6G.cpp:42-54, 92-118, 155-181, 219-245).  Tables are indexed by
// atomic number; only H, C, N, O are supported (the host rejects others).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sg {

constexpr double kHartreeToEv = 27.211386;  // mymath.h:94-97
constexpr double kBohrToAng = 0.529177249;  // mymath.h:89-92
constexpr int kNz = 9;

struct Params {
  double uss[kNz], upp[kNz], betaS[kNz], betaP[kNz], alpha[kNz];
  double gss[kNz], gpp[kNz], gsp[kNz], gp2[kNz], hsp[kNz];
  double dd0[kNz], dd1[kNz], pp0[kNz], pp1[kNz], pp2[kNz];
};

static __constant__ Params c_par = {
    /* uss   */ {0, -11.90627600, 0, 0, 0, 0, -52.27974500, -71.93212200, -99.64430900},
    /* upp   */ {0, 0, 0, 0, 0, 0, -39.20555800, -57.17231900, -77.79747200},
    /* betaS */ {0, -6.98906400, 0, 0, 0, 0, -18.9850440, -20.49575800, -32.68808200},
    /* betaP */ {0, 0, 0, 0, 0, 0, -7.93412200, -20.49575800, -32.68808200},
    /* alpha */ {0, 2.54413410, 0, 0, 0, 0, 2.546380, 2.86134200, 3.16060400},
    /* gss   */ {0, 12.848, 0, 0, 0, 0, 12.23, 13.59, 15.42},
    /* gpp   */ {0, 0, 0, 0, 0, 0, 11.08, 12.98, 14.52},
    /* gsp   */ {0, 0, 0, 0, 0, 0, 11.47, 12.66, 14.48},
    /* gp2   */ {0, 0, 0, 0, 0, 0, 9.84, 11.59, 12.98},
    /* hsp   */ {0, 0, 0, 0, 0, 0, 2.43, 3.14, 3.94},
    /* dd0   */ {0, 0, 0, 0, 0, 0, 0.80746618, 0.63990367, 0.53460239},
    /* dd1   */ {0, 0, 0, 0, 0, 0, 0.68515777, 0.54297627, 0.45362517},
    /* pp0   */ {0, 1.05897351, 0, 0, 0, 0, 1.11248501, 1.00115465, 0.88234058},
    /* pp1   */ {0, 0, 0, 0, 0, 0, 0.81309641, 0.63747362, 0.52124931},
    /* pp2   */ {0, 0, 0, 0, 0, 0, 0.74785493, 0.61528499, 0.52654944},
};

struct Sto6g {
  double expS[kNz][6], expP[kNz][6], cS[kNz][6], cP[kNz][6];
};

static __constant__ Sto6g c_sto = {
    /* expS */
    {{0},
     {4.098792193376860e+01, 7.515090619202102e+00, 2.102451537968084e+00,
      7.222488471249888e-01, 2.804703624957924e-01, 1.155131838418812e-01},
     {0}, {0}, {0}, {0},
     {8.846144274163005e+01, 1.622292919221501e+01, 4.558993095148346e+00,
      6.519461344814428e-01, 2.958932522507792e-01, 1.411098199523801e-01},
     {1.408553999528384e+02, 2.583144824406823e+01, 7.259194242116380e+00,
      1.038080893483724e+00, 4.711449542017903e-01, 2.246863662931398e-01},
     {2.018091738330405e+02, 3.700975064350955e+01, 1.040053837613310e+01,
      1.487300079059505e+00, 6.750282487920551e-01, 3.219171573709536e-01}},
    /* expP */
    {{0}, {0}, {0}, {0}, {0}, {0},
     {1.875086664726173e+01, 4.889844711496806e+00, 1.749633028685196e+00,
      7.313800499385337e-01, 3.344367808801327e-01, 1.581099099775126e-01},
     {2.985663289232627e+01, 7.786002705798312e+00, 2.785905953918744e+00,
      1.164561940873007e+00, 5.325170500259336e-01, 2.517552723103985e-01},
     {4.277679392805792e+01, 1.115531796470053e+01, 3.991479056714438e+00,
      1.668514542173559e+00, 7.629585088943974e-01, 3.606998633356254e-01}},
    /* cS */
    {{0},
     {1.057956874127020e-01, 1.596794562489045e-01, 2.097264517638662e-01,
      2.069126553471125e-01, 1.144014417359658e-01, 1.840525175585220e-02},
     {0}, {0}, {0}, {0},
     {-8.534093658582062e-02, -1.190837759635380e-01, -1.145234971659361e-01,
      1.730334274890658e-01, 1.607236215755646e-01, 2.810826497961667e-02},
     {-1.209684224448265e-01, -1.687979660569579e-01, -1.623338967119989e-01,
      2.452701082384343e-01, 2.278212980714544e-01, 3.984268990094886e-02},
     {-1.584154624760238e-01, -2.210511414259479e-01, -2.125860518260085e-01,
      3.211960348235300e-01, 2.983457630220672e-01, 5.217641116074215e-02}},
    /* cP */
    {{0}, {0}, {0}, {0}, {0}, {0},
     {4.407314493338078e-01, 5.331737162464377e-01, 5.445171182695406e-01,
      3.904445162839797e-01, 1.454563993779586e-01, 1.494840672526637e-02},
     {7.883124678886093e-01, 9.536589428843584e-01, 9.739482715823707e-01,
      6.983669549124415e-01, 2.601699818271488e-01, 2.673740531657013e-02},
     {1.235684643746845e+00, 1.494866260646997e+00, 1.526669908217376e+00,
      1.094694498739664e+00, 4.078180472887439e-01, 4.191104734371579e-02}},
};

// MOPAC pair order inside a heavy atom: SS, SX, XX, SY, XY, YY, SZ, XZ, YZ, ZZ
// (twocenterintegral.cpp:399-402); MOPIDX[a][b] is the slot of AO pair (a, b).
__host__ __device__ constexpr int mop_a(int k) {
  return k == 0 ? 0 : k == 1 ? 0 : k == 2 ? 1 : k == 3 ? 0 : k == 4 ? 1 : k == 5 ? 2 : k == 6 ? 0
                                                                        : k == 7 ? 1 : k == 8 ? 2 : 3;
}
__host__ __device__ constexpr int mop_b(int k) {
  return k == 0 ? 0 : k <= 2 ? 1 : k <= 5 ? 2 : 3;
}
__host__ __device__ constexpr int mop_idx(int a, int b) {
  return a > b ? mop_idx(b, a) : (b == 0 ? 0 : b == 1 ? a + 1 : b == 2 ? a + 3 : a + 6);
}

__host__ __device__ inline int core_charge(int z) { return z < 2 ? z : (z < 11 ? z - 2 : z - 10); }
__host__ __device__ inline int n_ao_of(int z) { return z < 2 ? 1 : 4; }
__host__ __device__ inline int block_size(int za, int zb) {
  return (za == 1 && zb == 1) ? 1 : ((za > 1 && zb > 1) ? 100 : 10);
}

// Per-molecule metadata on the device (structure of arrays in the batch).
struct MolDesc {
  int atom0;       // first atom (global)
  int n_atoms;
  int n_ao;        // N
  int n_occ;       // floor(n_electrons / 2)
  int64_t mat0;    // offset of this molecule's N*N matrices
  int64_t blk0;    // offset of this molecule's pair blocks
  int64_t vec0;    // offset of this molecule's length-N vectors
};

// Per-atom metadata.
struct AtomDesc {
  double x, y, z;  // bohr
  int zn;          // atomic number
  int mol;         // molecule index
  int local;       // index inside the molecule
  int ao0;         // first AO inside the molecule
  int hv_before;   // heavy atoms before this one in its molecule
  int pad;
  int64_t row0;    // offset (relative to the molecule's blk0) of block (this, 0)
};

// Offset (relative to the molecule's blk0) of the block of atoms i >= j
// in the reference jagged order (twocenterintegral.cpp:55-107).
__host__ __device__ inline int64_t block_offset(const AtomDesc& ai, const AtomDesc& aj) {
  int64_t h_before = aj.local - aj.hv_before;
  return ai.zn > 1 ? ai.row0 + 100 * (int64_t)aj.hv_before + 10 * h_before
                   : ai.row0 + 10 * (int64_t)aj.hv_before + h_before;
}

}  // namespace sg
