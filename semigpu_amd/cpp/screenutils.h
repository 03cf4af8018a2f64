// screenutils.h -- the console helpers main() uses (reference: src/common/screenutils.h).
#ifndef SEMIGPU_SCREENUTILS_H
#define SEMIGPU_SCREENUTILS_H

#include <cstddef>

class ScreenUtils {
 public:
  static void PrintScrStarLine(void);         // 80 '*'
  static void PrintScrCharLine(char h);       // 80 h
  static void PrintVectorN(size_t n, const double* v);
  static void PrintMatrixNxN(size_t n, const double* m);              // column-major
  static void PrintMatrixNxNSymmetric(size_t n, const double* packed);  // lower triangle
};

#endif  // SEMIGPU_SCREENUTILS_H
